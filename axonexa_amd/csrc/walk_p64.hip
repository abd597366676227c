// walk_p64.hip — walk kernels for 64 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 64
#include "walk_pick.h"
