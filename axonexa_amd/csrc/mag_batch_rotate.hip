// mag_batch_rotate.hip — batched magnetization walk kernels, integrator rotate.
#include "mag_batch_kernel.h"

namespace axx {

mag_batch_fn pick_mag_batch_rotate(int mb) {
    switch (mb) {
    case 1: return mag_batch_walk_kernel<1, kMagRotate>;
    case 2: return mag_batch_walk_kernel<2, kMagRotate>;
    case 3: return mag_batch_walk_kernel<3, kMagRotate>;
    case 4: return mag_batch_walk_kernel<4, kMagRotate>;
    case 5: return mag_batch_walk_kernel<5, kMagRotate>;
    case 6: return mag_batch_walk_kernel<6, kMagRotate>;
    case 7: return mag_batch_walk_kernel<7, kMagRotate>;
    default: return mag_batch_walk_kernel<8, kMagRotate>;
    }
}

}  // namespace axx
