// walk_p1.hip — walk kernels for 1 profile column(s); one translation
// unit per profile width so the instantiations compile in parallel.
#include "walk_kernel.h"

namespace axx {

template <int AX, bool SPH>
static walk_fn pick_rej(bool rej) {
    return rej ? (walk_fn)walk_kernel<1, AX, SPH, true> : (walk_fn)walk_kernel<1, AX, SPH, false>;
}
template <int AX>
static walk_fn pick_sph(bool sph, bool rej) {
    return sph ? pick_rej<AX, true>(rej) : pick_rej<AX, false>(rej);
}

walk_fn pick_walk_p1(int ax, bool sph, bool rej) {
    switch (ax) {
    case 1: return pick_sph<1>(sph, rej);
    case 3: return pick_sph<3>(sph, rej);
    case 5: return pick_sph<5>(sph, rej);
    default: return pick_sph<7>(sph, rej);
    }
}

}  // namespace axx
