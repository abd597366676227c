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

// Lookup-group mask: spheres and planes (GEN) always take groups of two;
// isotropic cylinder walks take the group as a compile-time constant; the
// axis-sign and Gaussian generators (ALT) read it from the arguments.
template <int AX, bool GEN, bool REJ>
static walk_fn pick_alt(bool alt, int gmask) {
    if (alt) return (walk_fn)walk_kernel<AXX_WALK_P, AX, GEN, REJ, true, 0>;
    if (GEN || gmask == 1) return (walk_fn)walk_kernel<AXX_WALK_P, AX, GEN, REJ, false, 1>;
    if (gmask == 3) return (walk_fn)walk_kernel<AXX_WALK_P, AX, false, REJ, false, 3>;
    return (walk_fn)walk_kernel<AXX_WALK_P, AX, false, REJ, false, 7>;
}
template <int AX, bool GEN>
static walk_fn pick_rej(bool rej, bool alt, int gmask) {
    return rej ? pick_alt<AX, GEN, true>(alt, gmask) : pick_alt<AX, GEN, false>(alt, gmask);
}
template <int AX>
static walk_fn pick_gen(bool gen, bool rej, bool alt, int gmask) {
    return gen ? pick_rej<AX, true>(rej, alt, gmask) : pick_rej<AX, false>(rej, alt, gmask);
}

// ax: 1 = x, 3 = x and y, 5 = x and z, otherwise all three axes;
// gmask: steps per lookup - 1 (1, 3 or 7; 1 whenever gen)
walk_fn AXX_PICK_NAME(AXX_WALK_P)(int ax, bool gen, bool rej, bool alt, int gmask) {
    switch (ax) {
    case 1: return pick_gen<1>(gen, rej, alt, gmask);
    case 3: return pick_gen<3>(gen, rej, alt, gmask);
    case 5: return pick_gen<5>(gen, rej, alt, gmask);
    default: return pick_gen<7>(gen, rej, alt, gmask);
    }
}

}  // namespace axx
