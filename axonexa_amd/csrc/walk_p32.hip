// walk_p32.hip — walk kernels for 32 profile column(s); one translation
// unit per profile width so the instantiations compile in parallel.
#include "walk_kernel.h"

namespace axx {

template <int AX, bool GEN>
static walk_fn pick_rej(bool rej) {
    return rej ? (walk_fn)walk_kernel<32, AX, GEN, true> : (walk_fn)walk_kernel<32, AX, GEN, false>;
}
template <int AX>
static walk_fn pick_gen(bool gen, bool rej) {
    return gen ? pick_rej<AX, true>(rej) : pick_rej<AX, false>(rej);
}

walk_fn pick_walk_p32(int ax, bool gen, bool rej) {
    switch (ax) {
    case 1: return pick_gen<1>(gen, rej);
    case 3: return pick_gen<3>(gen, rej);
    case 5: return pick_gen<5>(gen, rej);
    default: return pick_gen<7>(gen, rej);
    }
}

}  // namespace axx
