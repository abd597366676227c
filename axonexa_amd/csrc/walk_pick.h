// walk_pick.h — the walk-kernel instantiations of one profile width
// AXX_WALK_P and their dispatch function pick_walk_p<AXX_WALK_P>(); each
// walk_p*.hip defines the width and includes this file, so the widths
// compile as separate translation units in parallel.
#pragma once
#include "walk_kernel.h"

#ifndef AXX_WALK_P
#error "define AXX_WALK_P before including walk_pick.h"
#endif

#define AXX_PICK_NAME2(p) pick_walk_p##p
#define AXX_PICK_NAME(p) AXX_PICK_NAME2(p)

namespace axx {

template <int AX, bool GEN, bool REJ>
static walk_fn pick_alt(bool alt) {
    return alt ? (walk_fn)walk_kernel<AXX_WALK_P, AX, GEN, REJ, true>
               : (walk_fn)walk_kernel<AXX_WALK_P, AX, GEN, REJ, false>;
}
template <int AX, bool GEN>
static walk_fn pick_rej(bool rej, bool alt) {
    return rej ? pick_alt<AX, GEN, true>(alt) : pick_alt<AX, GEN, false>(alt);
}
template <int AX>
static walk_fn pick_gen(bool gen, bool rej, bool alt) {
    return gen ? pick_rej<AX, true>(rej, alt) : pick_rej<AX, false>(rej, alt);
}

// ax: 1 = x, 3 = x and y, 5 = x and z, otherwise all three axes
walk_fn AXX_PICK_NAME(AXX_WALK_P)(int ax, bool gen, bool rej, bool alt) {
    switch (ax) {
    case 1: return pick_gen<1>(gen, rej, alt);
    case 3: return pick_gen<3>(gen, rej, alt);
    case 5: return pick_gen<5>(gen, rej, alt);
    default: return pick_gen<7>(gen, rej, alt);
    }
}

}  // namespace axx
