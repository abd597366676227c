// walk_p16.hip — walk kernels for 16 profile column(s): one translation unit per
// profile width, so the widths compile in parallel.
#define AXX_WALK_P 16
#include "walk_pick.h"
