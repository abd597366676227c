// mag_rk4.hip — magnetization walk kernels, integrator rk4; one translation
// unit per integrator so the instantiations compile in parallel.
#include "mag_kernel.h"

namespace axx {

template <int MB, bool GEN, bool REJ>
static mag_fn pick_alt(bool alt) {
    return alt ? (mag_fn)mag_walk_kernel<MB, kMagRK4, GEN, REJ, true>
               : (mag_fn)mag_walk_kernel<MB, kMagRK4, GEN, REJ, false>;
}
template <int MB, bool GEN>
static mag_fn pick_rej(bool rej, bool alt) {
    return rej ? pick_alt<MB, GEN, true>(alt) : pick_alt<MB, GEN, false>(alt);
}
template <int MB>
static mag_fn pick_gen(bool gen, bool rej, bool alt) {
    return gen ? pick_rej<MB, true>(rej, alt) : pick_rej<MB, false>(rej, alt);
}

mag_fn pick_mag_rk4(int mb, bool gen, bool rej, bool alt) {
    switch (mb) {
    case 1: return pick_gen<1>(gen, rej, alt);
    case 2: return pick_gen<2>(gen, rej, alt);
    case 3: return pick_gen<3>(gen, rej, alt);
    case 4: return pick_gen<4>(gen, rej, alt);
    case 5: return pick_gen<5>(gen, rej, alt);
    case 6: return pick_gen<6>(gen, rej, alt);
    case 7: return pick_gen<7>(gen, rej, alt);
    default: return pick_gen<8>(gen, rej, alt);
    }
}

}  // namespace axx
