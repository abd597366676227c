// mag_batch_dopri.hip — batched magnetization walk kernels, integrator dopri.
#include "mag_batch_kernel.h"

namespace axx {

template <int MB>
static mag_batch_fn pick_cls(int cls) {
    switch (cls) {
    case 0: return mag_batch_walk_kernel<MB, kMagDopri, 0>;
    case 1: return mag_batch_walk_kernel<MB, kMagDopri, 1>;
    case 2: return mag_batch_walk_kernel<MB, kMagDopri, 2>;
    case 3: return mag_batch_walk_kernel<MB, kMagDopri, 3>;
    default: return mag_batch_walk_kernel<MB, kMagDopri, -1>;
    }
}

// cls: -1 mixed, 0..3 the class every run of the batch has (isotropic steps)
mag_batch_fn pick_mag_batch_dopri(int mb, int cls) {
    switch (mb) {
    case 1: return pick_cls<1>(cls);
    case 2: return pick_cls<2>(cls);
    case 3: return pick_cls<3>(cls);
    case 4: return pick_cls<4>(cls);
    case 5: return pick_cls<5>(cls);
    case 6: return pick_cls<6>(cls);
    case 7: return pick_cls<7>(cls);
    default: return pick_cls<8>(cls);
    }
}

}  // namespace axx
