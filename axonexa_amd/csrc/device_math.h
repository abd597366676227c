// device_math.h — random numbers and geometry predicates of the walk.
//
// Every operation is written out explicitly (fmaf where a fused multiply-add
// is intended; the library is compiled with -ffp-contract=off) so that the
// arithmetic is fully specified; sqrtf and '/' are the correctly rounded IEEE
// operations.
//
// Reference counterparts: curand_uniform / acos / sin / cos of the step
// direction (src/kernels/kernelLattice.cuh:52-57), Cylinder_XY::intersect
// (src/primitives/cylinderXY.cuh:82-122), Sphere::intersect
// (src/primitives/Sphere.cuh:102-139), Lattice::lmod (src/substrate/
// lattice.cuh:66-88).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axx {

// Philox4x32-10 (Salmon et al., SC'11).  Replaces cuRAND XORWOW
// (src/kernels/kernelSetup.cuh:1-5): stateless, keyed by (seed), counter =
// (walker, index, stream<<24 | sub, instance).
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                          uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        // three-way xor in one v_bitop3_b32 (truth table 0x96); the
        // compiler otherwise emits two v_xor_b32
        uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return u32x4{c0, c1, c2, c3};
}

constexpr uint32_t kStreamStep = 0u;
constexpr uint32_t kStreamColl = 1u << 24;
constexpr uint32_t kStreamInit = 2u << 24;
constexpr uint32_t kStreamGauss = 3u << 24;   // extra words of Gaussian steps

// Uniform in (0,1): odd multiples of 2^-24 (exactly representable).
__device__ __forceinline__ float unif(uint32_t w) {
    // ((w >> 8) | 1) 2^-24 = (w >> 9) 2^-23 + 2^-24, exact either way
    return fmaf((float)(w >> 9), 0x1p-23f, 0x1p-24f);
}

// sin and cos of 2 pi v 2^-24 for a 24-bit integer v: integer quarter-turn
// reduction, Taylor polynomials of sin(pi/2 r), cos(pi/2 r) for r in
// [-1/2, 1/2].  Used to build the azimuth table below.
__device__ __forceinline__ void sincos_u24(uint32_t v24, float& s, float& c) {
    // v = v24 + 2^21; quarter turn q = v >> 22 (0..4); r = 4 v24 2^-24 - q in
    // [-1/2, 1/2) from the low 22 bits of v, exact
    const uint32_t v = v24 + (1u << 21);
    const uint32_t q = v >> 22;
    float r = fmaf((float)(v & 0x3FFFFFu), 0x1p-22f, -0.5f);
    float y = r * r;
    float ps = fmaf(y, 0.00016044118478735982f, -0.004681754135318687f);
    ps = fmaf(y, ps, 0.07969262624616704f);
    ps = fmaf(y, ps, -0.6459640975062462f);
    ps = fmaf(y, ps, 1.5707963267948966f);
    float sp = r * ps;
    float pc = fmaf(y, -0.000025202042373060605f, 0.00091926027483942659f);
    pc = fmaf(y, pc, -0.020863480763352960f);
    pc = fmaf(y, pc, 0.25366950790104802f);
    pc = fmaf(y, pc, -1.2337005501361697f);
    float cp = fmaf(y, pc, 1.0f);
    uint32_t qq = q & 3u;
    float a0 = (qq & 1u) ? cp : sp;   // sin magnitude source
    float b0 = (qq & 1u) ? sp : cp;   // cos magnitude source
    s = (qq >= 2u) ? -a0 : a0;
    c = (qq == 1u || qq == 2u) ? -b0 : b0;
}

// Azimuth table: entry j = (cos, sin) of 2 pi (j + 1/2) / 256, i.e.
// sincos_u24((2j + 1) << 15), j = 0..255 (2 KB, built in LDS by each
// workgroup).
constexpr int kAzimuthEntries = 256;
__device__ __forceinline__ float2 azimuth_entry(int j) {
    float s, c;
    sincos_u24((2u * (uint32_t)j + 1u) << 15, s, c);
    return make_float2(c, s);
}

// sin and cos of 2 pi u(w), u(w) = ((w >> 8) | 1) 2^-24: the table entry of
// the top 8 bits of the 24-bit value v, turned by d = 2 pi (v 2^-24 - (j +
// 1/2)/256) = (f - 2^15) K1 with f the low 16 bits of v and K1 = the float
// of 2 pi times 2^-24 (|d| <= pi/256): cos d = 1 - d^2/2, sin d = d (1 -
// d^2/6) (truncation below 1e-9), then the angle-sum formulas.  Replaces
// phi = 2 PI curand_uniform; sin(phi), cos(phi) (kernelLattice.cuh:52-57).
__device__ __forceinline__ void sincos_tab(const float2* tab, uint32_t w, float& s, float& c) {
    const uint32_t v = (w >> 8) | 1u;
    const float2 e = tab[v >> 16];
    // (f - 2^15) K1 in one rounding: 2^15 K1 is exact
    const float d = fmaf((float)(v & 0xFFFFu), 0x1.921fb6p-22f, -0x1.921fb6p-7f);
    const float d2 = d * d;
    const float cd = fmaf(d2, -0.5f, 1.0f);
    const float sd = d * fmaf(d2, -0.16666667f, 1.0f);
    c = fmaf(-e.y, sd, e.x * cd);
    s = fmaf(e.x, sd, e.y * cd);
}

// -2 ln u(w) for the Box-Muller transform of Gaussian steps: u = k 2^-24
// with k odd, ln u = (e - 24) ln 2 + ln m for k = 2^e m, m in [sqrt(1/2),
// sqrt(2)), ln m = 2 atanh(s) with s = (m - 1) / (m + 1), |s| < 0.172, by
// its odd series to s^9 (truncation < 4e-10).  The result lies in
// [1.19e-7, 33.3].
__device__ __forceinline__ float neg2log_unif(uint32_t w) {
    const float x = (float)((w >> 8) | 1u);
    const uint32_t bits = __float_as_uint(x);
    int e = (int)(bits >> 23) - 127;
    float m = __uint_as_float((bits & 0x007FFFFFu) | 0x3F800000u);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = fmaf(s2, 0.11111111f, 0.14285714f);
    p = fmaf(s2, p, 0.2f);
    p = fmaf(s2, p, 0.33333334f);
    p = fmaf(s2, p, 1.0f);
    const float lnm = (2.0f * s) * p;
    const float lnu = fmaf((float)(e - 24), 0.69314718f, lnm);
    return -2.0f * lnu;
}

// Periodic wrap of one coordinate (Lattice::lmod, strict comparisons).
__device__ __forceinline__ float wrap1(float v, float a) {
    if (v > a) v = v - a;
    if (v < 0.0f) v = v + a;
    return v;
}

// A placement coordinate v of the substrate's cell taken into the unit cell
// of width w of a repeating substrate: v - floor(v / w) w in one rounding,
// then wrapped as a walk position (every operation correctly rounded).
__device__ __forceinline__ float unit_wrap(float v, float w) {
    const float k = floorf(v / w);
    return wrap1(fmaf(-k, w, v), w);
}

// Crossing of the segment s + t*u, t in [0, ell], with a circle (cylinder,
// xy only) or sphere surface centred at o-offsets (ox, oy, oz) = s - centre.
// The walker's side comes from its tracked compartment; the start point is
// clamped onto that side, which replaces the reference's epsilon exclusion
// of the root at the reflection point.
// The walker is inside this image iff it is in the image's compartment and
// within half a cell (H2 = squared half of the smallest cell side) of its
// centre: other images of the same primitive are at least that far away.
template <bool SPH>
__device__ __forceinline__ bool hit_test(float ox, float oy, float oz, float ux, float uy,
                                         float uz, float A, float R2, bool same_comp, float H2,
                                         float ell, float& t_out, bool& ins_out) {
    float B, q;
    if (SPH) {
        B = fmaf(oz, uz, fmaf(oy, uy, ox * ux));
        q = fmaf(oz, oz, fmaf(oy, oy, ox * ox));
    } else {
        B = fmaf(oy, uy, ox * ux);
        q = fmaf(oy, oy, ox * ox);
    }
    const float C = q - R2;
    const bool inside = same_comp && q < H2;
    ins_out = inside;
    if (!(A > 0.0f)) return false;
    if (inside) {
        float Cp = C < 0.0f ? C : 0.0f;
        float fe = fmaf(fmaf(A, ell, B + B), ell, Cp);
        if (!(fe > 0.0f)) return false;
        float disc = fmaf(B, B, -(A * Cp));
        float sq = sqrtf(disc);
        float t;
        if (B < 0.0f) {
            t = (sq - B) / A;
        } else {
            float den = B + sq;
            t = den > 0.0f ? (-Cp) / den : 0.0f;
        }
        t_out = t < ell ? t : ell;
        return true;
    }
    if (!(B < 0.0f)) return false;
    float Cp = C > 0.0f ? C : 0.0f;
    if (Cp == 0.0f) {
        t_out = 0.0f;
        return true;
    }
    float fe = fmaf(fmaf(A, ell, B + B), ell, Cp);
    float disc = fmaf(B, B, -(A * Cp));
    if (!(fe < 0.0f) && !((-B) < A * ell && disc > 0.0f)) return false;
    float sq = sqrtf(disc > 0.0f ? disc : 0.0f);
    float t = Cp / (sq - B);
    t_out = t < ell ? t : ell;
    return true;
}


// Correctly rounded sqrt for arguments known to be normal and finite
// (here 1 - c^2 in [2^-22, 1]): hardware estimate plus the one-ulp
// residual correction; equal to IEEE sqrtf on that domain.
__device__ __forceinline__ float sqrt_unit(float x) {
    float y = __builtin_amdgcn_sqrtf(x);
    const float ydn = __int_as_float(__float_as_int(y) - 1);
    const float yup = __int_as_float(__float_as_int(y) + 1);
    const float rdn = fmaf(-ydn, y, x);
    const float rup = fmaf(-yup, y, x);
    y = rdn <= 0.0f ? ydn : y;
    y = rup > 0.0f ? yup : y;
    return y;
}

// The crossing test of hit_test split in two: a branch-free decision (does
// the segment cross this surface within [0, ell]?) and the root evaluation
// for a surface that is crossed.  Same operations and operands as hit_test,
// so the same results.
struct Quad {
    float A, B, Cp, fe, disc;
    bool inside;
    unsigned long long m_in;   // MASK quads: the wave's lane mask of inside
};

// MASK: also the lane mask of `inside` (for quad_hits_masked), made from the
// two compares with ballots.
template <bool SPH, bool MASK = false>
__device__ __forceinline__ Quad quad_of(float ox, float oy, float oz, float ux, float uy, float uz,
                                        float A, float R2, bool same_comp, float H2, float ell) {
    Quad Q;
    float q;
    if (SPH) {
        Q.B = fmaf(oz, uz, fmaf(oy, uy, ox * ux));
        q = fmaf(oz, oz, fmaf(oy, oy, ox * ox));
    } else {
        Q.B = fmaf(oy, uy, ox * ux);
        q = fmaf(oy, oy, ox * ox);
    }
    const float C = q - R2;
    Q.A = A;
    if (MASK) {
        Q.m_in = __builtin_amdgcn_ballot_w64(same_comp) & __builtin_amdgcn_ballot_w64(q < H2);
        Q.inside = __builtin_amdgcn_inverse_ballot_w64(Q.m_in);
    } else {
        Q.m_in = 0ull;
        Q.inside = same_comp && q < H2;
    }
    Q.Cp = Q.inside ? (C < 0.0f ? C : 0.0f) : (C > 0.0f ? C : 0.0f);
    Q.fe = fmaf(fmaf(A, ell, Q.B + Q.B), ell, Q.Cp);
    Q.disc = fmaf(Q.B, Q.B, -(A * Q.Cp));
    return Q;
}

__device__ __forceinline__ bool quad_hits(const Quad& Q, float ell) {
    const bool hit_in = Q.fe > 0.0f;
    // bitwise, not short-circuit: every operand is a compare, and evaluating
    // them all avoids the exec-mask branches the && / || chain compiles to
    const bool hit_out = (Q.B < 0.0f) & ((Q.Cp == 0.0f) | (Q.fe < 0.0f) | (((-Q.B) < Q.A * ell) & (Q.disc > 0.0f)));
    return (Q.A > 0.0f) & (Q.inside ? hit_in : hit_out);
}

// quad_hits for a MASK quad when A > 0 holds in every lane (cylinder-only
// walks: A = ux^2 + uy^2 of a drawn direction is at least about 2^-22 --
// sin^2 theta of the isotropic draw, 2/3 for axis signs, r1^2 / |n|^2 for
// Gaussian steps -- and a reflection about a normal with nz = 0 keeps it to
// rounding; tests/test_rng.py checks the bound over the generators' whole
// word grids).  Each compare is taken as the lane mask of its v_cmp (ballot),
// the masks are combined by scalar instructions and the result read back as
// the lane's condition (inverse ballot): the same decision as quad_hits
// without turning booleans into 0/1 vector values.  valid: lanes whose
// quad is a real candidate (others answer false).
__device__ __forceinline__ bool quad_hits_masked(const Quad& Q, float ell, unsigned long long valid) {
    typedef unsigned long long u64;
    const u64 hit_in = __builtin_amdgcn_ballot_w64(Q.fe > 0.0f);
    const u64 hit_out = __builtin_amdgcn_ballot_w64(Q.B < 0.0f) &
                        (__builtin_amdgcn_ballot_w64(Q.Cp == 0.0f) | __builtin_amdgcn_ballot_w64(Q.fe < 0.0f) |
                         (__builtin_amdgcn_ballot_w64((-Q.B) < Q.A * ell) & __builtin_amdgcn_ballot_w64(Q.disc > 0.0f)));
    return __builtin_amdgcn_inverse_ballot_w64(valid & ((Q.m_in & hit_in) | (~Q.m_in & hit_out)));
}

// inside, B < 0: (sq - B) / A;  inside, B >= 0: -Cp / (B + sq);
// outside: Cp / (sq - B); clamped to ell
__device__ __forceinline__ float quad_root(const Quad& Q, float ell) {
    const float sq = sqrtf(Q.disc > 0.0f ? Q.disc : 0.0f);
    const bool far_root = Q.inside && Q.B < 0.0f;
    // both sums in every lane (the empty asm keeps the compiler from sinking
    // them into exec-mask branches): the choices below stay v_cndmask selects
    float smb = sq - Q.B, bps = Q.B + sq;
    asm volatile("" : "+v"(smb), "+v"(bps));
    const float num = far_root ? smb : (Q.inside ? -Q.Cp : Q.Cp);
    const float den = far_root ? Q.A : (Q.inside ? bps : smb);
    const float t = den > 0.0f ? num / den : 0.0f;
    return t < ell ? t : ell;
}

}  // namespace axx
