// walk_p2.hip — walk kernels for 2 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 2
#include "walk_pick.h"
