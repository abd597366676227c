// walk_p1.hip — walk kernels for 1 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 1
#include "walk_pick.h"
