// mag_dopri.hip — magnetization walk kernels, integrator dopri; one translation
// unit per integrator so the instantiations compile in parallel.
#include "mag_kernel.h"

namespace axx {

template <int MB, bool GEN>
static mag_fn pick_rej(bool rej) {
    return rej ? (mag_fn)mag_walk_kernel<MB, kMagDopri, GEN, true> : (mag_fn)mag_walk_kernel<MB, kMagDopri, GEN, false>;
}
template <int MB>
static mag_fn pick_gen(bool gen, bool rej) {
    return gen ? pick_rej<MB, true>(rej) : pick_rej<MB, false>(rej);
}

mag_fn pick_mag_dopri(int mb, bool gen, bool rej) {
    switch (mb) {
    case 1: return pick_gen<1>(gen, rej);
    case 2: return pick_gen<2>(gen, rej);
    case 3: return pick_gen<3>(gen, rej);
    case 4: return pick_gen<4>(gen, rej);
    case 5: return pick_gen<5>(gen, rej);
    case 6: return pick_gen<6>(gen, rej);
    case 7: return pick_gen<7>(gen, rej);
    default: return pick_gen<8>(gen, rej);
    }
}

}  // namespace axx
