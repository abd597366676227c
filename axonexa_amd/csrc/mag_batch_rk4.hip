// mag_batch_rk4.hip — batched magnetization walk kernels, integrator rk4.
#include "mag_batch_kernel.h"

namespace axx {

mag_batch_fn pick_mag_batch_rk4(int mb) {
    switch (mb) {
    case 1: return mag_batch_walk_kernel<1, kMagRK4>;
    case 2: return mag_batch_walk_kernel<2, kMagRK4>;
    case 3: return mag_batch_walk_kernel<3, kMagRK4>;
    case 4: return mag_batch_walk_kernel<4, kMagRK4>;
    case 5: return mag_batch_walk_kernel<5, kMagRK4>;
    case 6: return mag_batch_walk_kernel<6, kMagRK4>;
    case 7: return mag_batch_walk_kernel<7, kMagRK4>;
    default: return mag_batch_walk_kernel<8, kMagRK4>;
    }
}

}  // namespace axx
