// walk_p4.hip — walk kernels for 4 profile column(s); one translation
// unit per profile width so the instantiations compile in parallel.
#include "walk_kernel.h"

namespace axx {

template <int AX, bool GEN, bool REJ>
static walk_fn pick_alt(bool alt) {
    return alt ? (walk_fn)walk_kernel<4, AX, GEN, REJ, true> : (walk_fn)walk_kernel<4, AX, GEN, REJ, false>;
}
template <int AX, bool GEN>
static walk_fn pick_rej(bool rej, bool alt) {
    return rej ? pick_alt<AX, GEN, true>(alt) : pick_alt<AX, GEN, false>(alt);
}
template <int AX>
static walk_fn pick_gen(bool gen, bool rej, bool alt) {
    return gen ? pick_rej<AX, true>(rej, alt) : pick_rej<AX, false>(rej, alt);
}

walk_fn pick_walk_p4(int ax, bool gen, bool rej, bool alt) {
    switch (ax) {
    case 1: return pick_gen<1>(gen, rej, alt);
    case 3: return pick_gen<3>(gen, rej, alt);
    case 5: return pick_gen<5>(gen, rej, alt);
    default: return pick_gen<7>(gen, rej, alt);
    }
}

}  // namespace axx
