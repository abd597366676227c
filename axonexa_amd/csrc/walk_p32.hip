// walk_p32.hip — walk kernels for 32 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 32
#include "walk_pick.h"
