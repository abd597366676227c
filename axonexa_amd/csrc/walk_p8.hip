// walk_p8.hip — walk kernels for 8 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 8
#include "walk_pick.h"
