// walk_p4.hip — walk kernels for 4 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 4
#include "walk_pick.h"
