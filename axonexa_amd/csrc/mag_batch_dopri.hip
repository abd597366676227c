// mag_batch_dopri.hip — batched magnetization walk kernels, integrator dopri.
#include "mag_batch_kernel.h"

namespace axx {

mag_batch_fn pick_mag_batch_dopri(int mb) {
    switch (mb) {
    case 1: return mag_batch_walk_kernel<1, kMagDopri>;
    case 2: return mag_batch_walk_kernel<2, kMagDopri>;
    case 3: return mag_batch_walk_kernel<3, kMagDopri>;
    case 4: return mag_batch_walk_kernel<4, kMagDopri>;
    case 5: return mag_batch_walk_kernel<5, kMagDopri>;
    case 6: return mag_batch_walk_kernel<6, kMagDopri>;
    case 7: return mag_batch_walk_kernel<7, kMagDopri>;
    default: return mag_batch_walk_kernel<8, kMagDopri>;
    }
}

}  // namespace axx
