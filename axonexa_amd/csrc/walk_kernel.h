// walk_kernel.h — the random-walk kernel for gfx950.
//
// One lane owns one walker for every step of its chunk: position,
// displacement, compartment, phase moments and counters stay in VGPRs; the
// substrate (primitives with periodic images + cell bins of candidate
// surfaces) is staged once per workgroup into LDS; the per-step profile
// samples are wave-uniform and come through the scalar cache.
//
// Reference: updateWalkersLatticePhase (src/kernels/kernelLattice.cuh:1-104)
// with boundaryCheckImpermeable / boundaryCheckPermeable
// (src/kernels/boundaryCheck.cuh:34-64, :66-113) and the Lattice scans
// (src/substrate/lattice.cuh:90-247).  Differences by design: Philox
// counters instead of per-thread XORWOW state, moments per (profile, axis)
// instead of M global read-modify-writes per step, LDS cell bins instead of
// the O(n_basis) scan, the segment form (start, direction, remaining length)
// of the reflection loop, side-tracked crossing tests, and true T2 time split.
#pragma once
#include <type_traits>

#include "device_math.h"

namespace axx {

struct WalkArgs {
    float a, b, c;                          // cell the positions wrap in
    float ca, cb;                           // the substrate's cell along x, y (placement); != a, b
                                            // when the walk runs in the unit cell of a repeating substrate
    float H2c, H2s;                         // squared half of the smallest cell side (xy; xyz)
    float plane_eps;                        // planes count as hit within this distance
    int nbx, nby, nbz;
    float ibz;                              // z bins per unit length (spheres)
    float ibx4, iby4;                       // sub-bins (4 x 4 per bin) per unit length
    int n_img, n_bins, n_cand;
    const float4* __restrict__ g_geom;      // [n_img] cx, cy, cz, R (planes: nx, ny, nz, d)
    const int* __restrict__ g_meta;         // [n_img] compartment | kSphereBit | kPlaneBit
    const float4* __restrict__ g_cyl;       // [n_img] cx, cy, R, meta bits (cylinder-only substrates)
    const float* __restrict__ g_inv;        // [n_img] 1 / R (float division; 0 for planes)
    const float2* __restrict__ g_thr;       // [n_img] transmission hazards -2 ln(1 - p) of the image's
                                            // primitive, p = 4 kappa / v_inside, 4 kappa / v_lattice
                                            // (+inf for p >= 1; 0 for planes)
    const uint4* __restrict__ g_bins;       // [n_bins] {offset << 8 | count, sub-bin nibbles 0-7, 8-15, 0}
    const uint16_t* __restrict__ g_cand;    // [n_cand] image indices, ascending per bin
    const float* __restrict__ prof;         // [n_steps][Pt] unit profile samples
    const float* __restrict__ compL;        // [n_comp] step length
    const float* __restrict__ rate;         // [n_comp] dt / T2
    uint32_t* __restrict__ state;           // SoA, AXX_STATE_FIELDS(P) x N
    long long N;
    uint32_t k0, k1, inst, woff;
    int mode, init_mode, init_arg, init_img;
    int step_mode;                          // 0 isotropic, 1 axis signs, 2 Gaussian
    int i0, i1, cap;
    int gmask;                              // steps per cell-bin lookup - 1 (1, 3 or 7): positions wrap
                                            // into the cell after step i when (i - 1) & gmask == gmask
    int* __restrict__ err;
};

constexpr int kSphereBit = 1 << 30;
constexpr int kPlaneBit = 1 << 29;
constexpr int kCompMask = (1 << 29) - 1;
// Sub-bin nibble: 0..6 = position in the bin's list of the single surface
// within reach, 7 = none, 8 (bit 3) = several (test the whole list).
constexpr uint32_t kNibNone = 7u;
constexpr uint32_t kNibMany = 8u;

// state field indices (include/axonexa.h)
enum {
    F_X = 0, F_Y, F_Z, F_DX, F_DY, F_DZ, F_X0, F_Y0, F_Z0, F_E, F_LF,
    F_COMP, F_COMP0, F_LI, F_NCOLL, F_NTRAN, F_NCAP, F_STEP, F_BUD, F_MOM
};

// g = (cx, cy, cz, R); inside test against R^2 = R * R (cylinderXY.cuh:69-74,
// Sphere.cuh:89-94)
__device__ __forceinline__ bool inside_at(float4 g, bool sph, float x, float y, float z) {
    const float R2 = g.w * g.w;
    float ox = x - g.x, oy = y - g.y;
    if (sph) {
        float oz = z - g.z;
        return fmaf(oz, oz, fmaf(oy, oy, ox * ox)) < R2;
    }
    return fmaf(oy, oy, ox * ox) < R2;
}

// Region of a point: compartment of the containing primitive (unique by the
// non-overlap rule), else the lattice (lattice.cuh:134-148).  Planes bound
// no region (plane.cuh:33).
__device__ __forceinline__ int region_of(const float4* geom, const int* meta, int n_img, float x,
                                         float y, float z) {
    for (int i = 0; i < n_img; ++i) {
        int mt = meta[i];
        if (mt & kPlaneBit) continue;
        if (inside_at(geom[i], (mt & kSphereBit) != 0, x, y, z)) return mt & kCompMask;
    }
    return 0;
}

// Plane g = (nx, ny, nz, d), unit normal n and offset n . point
// (plane.cuh:18-28 restated).  The walker's side is the side of the step's
// start point (x0, y0, z0); the segment start (sx, sy, sz) is clamped onto
// it.  The plane counts as hit when the segment comes closer than eps to it;
// t is the time of reaching the plane, at most ell.  Together with the push
// after a plane reflection this keeps every step end about eps or more away
// on the walker's side.
__device__ __forceinline__ bool plane_hit(float4 g, float x0, float y0, float z0, float sx, float sy,
                                          float sz, float ux, float uy, float uz, float ell, float eps,
                                          float& t) {
    const float un = fmaf(g.z, uz, fmaf(g.y, uy, g.x * ux));
    const float sd = fmaf(g.z, sz, fmaf(g.y, sy, g.x * sx)) - g.w;
    const float s0 = fmaf(g.z, z0, fmaf(g.y, y0, g.x * x0)) - g.w;
    const float sg = s0 >= 0.0f ? 1.0f : -1.0f;
    float dd = sg * sd;
    dd = dd > 0.0f ? dd : 0.0f;
    const float v = -(sg * un);
    if (!(v > 0.0f)) return false;
    if (!(fmaf(-ell, v, dd) < eps)) return false;
    const float tt = dd / v;
    t = tt < ell ? tt : ell;
    return true;
}

// Substrate tables in LDS: the azimuth table (2 KB), bins (16 B each), then one float4 per image
// (cylinder-only: cx, cy, R, meta; otherwise cx, cy, cz, R plus a meta
// array), the transmission hazards of each image's primitive (float2:
// from inside, from outside), the inverse radii, then the u16 candidate
// lists.
struct LdsTables {
    float2* sc;
    uint4* bins;
    float4* geom;
    float2* thr;
    int* meta;
    float* inv;
    uint16_t* cand;
};

template <bool GEN>
__device__ __forceinline__ LdsTables stage_tables(const WalkArgs& A, unsigned char* lds) {
    LdsTables T;
    T.sc = reinterpret_cast<float2*>(lds);
    T.bins = reinterpret_cast<uint4*>(T.sc + kAzimuthEntries);
    T.geom = reinterpret_cast<float4*>(T.bins + A.n_bins);
    T.thr = reinterpret_cast<float2*>(T.geom + A.n_img);
    T.meta = reinterpret_cast<int*>(T.thr + A.n_img);
    T.inv = GEN ? reinterpret_cast<float*>(T.meta + A.n_img) : reinterpret_cast<float*>(T.meta);
    T.cand = reinterpret_cast<uint16_t*>(T.inv + A.n_img);
    for (int i = threadIdx.x; i < kAzimuthEntries; i += blockDim.x) T.sc[i] = azimuth_entry(i);
    for (int i = threadIdx.x; i < A.n_bins; i += blockDim.x) T.bins[i] = A.g_bins[i];
    for (int i = threadIdx.x; i < A.n_img; i += blockDim.x) {
        if (GEN) {
            T.geom[i] = A.g_geom[i];
            T.meta[i] = A.g_meta[i];
        } else {
            T.geom[i] = A.g_cyl[i];
        }
        T.inv[i] = A.g_inv[i];
        T.thr[i] = A.g_thr[i];
    }
    for (int i = threadIdx.x; i < A.n_cand; i += blockDim.x) T.cand[i] = A.g_cand[i];
    __syncthreads();
    return T;
}

// Image k as (cx, cy, cz, R) or (nx, ny, nz, d) and its meta word.
template <bool GEN>
__device__ __forceinline__ float4 image_at(const LdsTables& T, int k, int& mt) {
    const float4 g = T.geom[k];
    if (GEN) {
        mt = T.meta[k];
        return g;
    }
    mt = __float_as_int(g.w);
    return make_float4(g.x, g.y, 0.0f, g.z);
}

// Profile samples are read through the constant address space: the address
// is wave-uniform, so these become scalar loads (s_load) into SGPRs.
typedef const __attribute__((address_space(4))) float* const_fptr;

// Work of one walk launch as tasks t = k * n_groups + g: group g is walkers
// [64 g, 64 g + 64), time chunk k is steps [b + k chunk, b + (k + 1) chunk)
// with b = max(i0, 1), clipped to [i0, i1).  Waves take tasks from a counter in that order (every
// group's chunk k before any group's chunk k + 1), so all groups advance
// together and the launch ends one chunk after the counter runs out, whatever
// the walker count; a wave that takes chunk k > 0 of a group first waits for
// that group's chunk k - 1 (taken by some wave before it) to be published.
struct WalkQueue {
    unsigned int* __restrict__ q;   // [0] task counter, [1 + g] chunks of group g finished;
                                    // zeroed before each launch
    uint32_t n_groups, n_tasks;
    int chunk;                      // steps per task (even: step pairs are never split)
    uint32_t wait_ticks;            // limit of one wait, in 100 MHz real-time ticks
};
constexpr int kErrWaitLimit = 2;    // walk error bit: a task gave up waiting for its group

typedef void (*walk_fn)(WalkArgs, WalkQueue, const float*);
walk_fn pick_walk_p1(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p2(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p4(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p8(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p16(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p32(int ax, bool gen, bool rej, bool alt, int gmask);
walk_fn pick_walk_p64(int ax, bool gen, bool rej, bool alt, int gmask);

// Wide profile sets keep up to 192 moments per lane: those instantiations are
// limited to 512 threads per block so that each lane may use 256 VGPRs.
constexpr int walk_max_block(int P) { return P >= 32 ? 512 : 1024; }

// Phase moments of P gradient profiles over the axes in AX
// (kernelLattice.cuh:81-85): m_p += g_p(i) * (r_i - r_0) per step, with the
// wave-uniform profile samples of step i.
template <int P, int AX>
struct PhaseAcc {
    static constexpr bool kOneSite = false;
    const_fptr prof;
    float m[P][3];

    __device__ __forceinline__ void load(const float* fs, long long N, uint32_t wo) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            m[p][0] = (AX & 1) ? (fs + (F_MOM + 3 * p + 0) * N)[wo] : 0.0f;
            m[p][1] = (AX & 2) ? (fs + (F_MOM + 3 * p + 1) * N)[wo] : 0.0f;
            m[p][2] = (AX & 4) ? (fs + (F_MOM + 3 * p + 2) * N)[wo] : 0.0f;
        }
    }
    // the position at t = 0 carries no moment
    __device__ __forceinline__ void at_start(float, float, float, int) {}
    __device__ __forceinline__ void step(int i, float dx, float dy, float dz, int) {
        const_fptr g = prof + (size_t)i * P;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const float gp = g[p];
            if (AX & 1) m[p][0] = fmaf(gp, dx, m[p][0]);
            if (AX & 2) m[p][1] = fmaf(gp, dy, m[p][1]);
            if (AX & 4) m[p][2] = fmaf(gp, dz, m[p][2]);
        }
    }
    __device__ __forceinline__ void store(float* fs, long long N, uint32_t wo) const {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (AX & 1) (fs + (F_MOM + 3 * p + 0) * N)[wo] = m[p][0];
            if (AX & 2) (fs + (F_MOM + 3 * p + 1) * N)[wo] = m[p][1];
            if (AX & 4) (fs + (F_MOM + 3 * p + 2) * N)[wo] = m[p][2];
        }
    }
};

// One lane walks walker wo of A over steps [max(A.i0, 1), A.i1) with the
// substrate tables T in LDS; acc sees the walker at t = 0 (when A.i0 <= 0)
// and after every step.  GEN: substrates with spheres or planes (3D bins,
// per-image meta words); otherwise cylinders only.  ALT: the axis-sign and
// Gaussian step generators are compiled in (selected by A.step_mode);
// otherwise isotropic steps only.
// GM: the lookup-group mask when fixed at compile time (1, 3 or 7), 0 = A.gmask
template <class Acc, bool GEN, bool REJ, bool ALT, int GM = 0>
__device__ __forceinline__ void walk_lane(const WalkArgs& A, const LdsTables& T, Acc& acc, const uint32_t wo) {
    const int gm = GM > 0 ? GM : A.gmask;
    const long long N = A.N;
    uint32_t* st = A.state;
    float* fs = reinterpret_cast<float*>(st);
    int* is = reinterpret_cast<int*>(st);
    const uint32_t wid = A.woff + wo;

    // field f of this walker: base + f*N (wave-uniform) + wo (32-bit lane offset)
    float x, y, z, dx, dy, dz, E, lf;
    int comp, li;
    uint32_t ncoll, ntran, ncap;
    x = (fs + F_X * N)[wo]; y = (fs + F_Y * N)[wo]; z = (fs + F_Z * N)[wo];
    dx = (fs + F_DX * N)[wo]; dy = (fs + F_DY * N)[wo]; dz = (fs + F_DZ * N)[wo];
    E = (fs + F_E * N)[wo]; lf = (fs + F_LF * N)[wo];
    comp = (is + F_COMP * N)[wo]; li = (is + F_LI * N)[wo];
    ncoll = (st + F_NCOLL * N)[wo]; ntran = (st + F_NTRAN * N)[wo];
    ncap = (st + F_NCAP * N)[wo];
    float bud = (fs + F_BUD * N)[wo];
    acc.load(fs, N, wo);

    const int cap = A.cap;
    const bool permeable = A.mode == 1;
    float Lc = A.compL[comp];

    // Surfaces within reach of the current start point: the bin's list,
    // narrowed by the sub-bin nibble to one entry (or the whole list when
    // several).  The lists cover a lookup group of gmask + 1 steps (two,
    // four or eight): one lookup serves the group (steps g (gmask + 1) + 1 ..; the
    // position wraps into the cell after its last step), and a task's first
    // steps inside a group take a lookup of their own; such a start may lie
    // up to gmask steps outside the cell (clamped to the edge sub-bin, still
    // within reach of the steps left in the group).
    struct Cands {
        int nrel;
        const uint16_t* cl;
    };
    auto lookup = [&]() {
        const int sbx = min(max((int)(x * A.ibx4), 0), 4 * A.nbx - 1);
        const int sby = min(max((int)(y * A.iby4), 0), 4 * A.nby - 1);
        const int bz = GEN ? min(max((int)(z * A.ibz), 0), A.nbz - 1) : 0;
        const uint32_t row = __umul24((uint32_t)bz, (uint32_t)A.nby) + ((uint32_t)sby >> 2);
        const uint4 rec = T.bins[__umul24(row, (uint32_t)A.nbx) + ((uint32_t)sbx >> 2)];
        const int sbi = (sbx & 3) | ((sby & 3) << 2);
        const uint32_t nib = ((sbi < 8 ? rec.y : rec.z) >> ((sbi & 7) << 2)) & 0xFu;
        const bool many = (nib & kNibMany) != 0u;
        Cands c;
        c.nrel = many ? (int)(rec.x & 0xFFu) : (nib == kNibNone ? 0 : 1);
        c.cl = T.cand + (rec.x >> 8) + (many ? 0u : nib);
        return c;
    };

    // One step i: direction from the two Philox words, boundary handling,
    // displacement, phase moments; the position is wrapped into the cell
    // after the last step of each lookup group (kernelLattice.cuh:46-99).
    auto step = [&](const int i, const uint32_t wa, const uint32_t wb, const uint32_t wc, const uint32_t wd,
                    const Cands C, const bool wrap) {
        const int nrel = C.nrel;
        const uint16_t* cl = C.cl;
        float ux, uy, uz, gs = 1.0f;   // direction; step length = gs * (compartment step)
        if (!ALT || A.step_mode == 0) {
            // isotropic direction (sphere point picking, kernelLattice.cuh:52-57)
            // cos(theta) = 2u - 1 and sin^2 = 1 - c^2, each in one rounding
            // 2 u(wb) - 1 = ((wb >> 8) | 1) 2^-23 - 1 = (wb >> 9) 2^-22 + (2^-23 - 1):
            // the same exact value under the one rounding of the fma
            const float ct = fmaf((float)(wb >> 9), 0x1p-22f, 0x1p-23f - 1.0f);
            const float sth = sqrt_unit(fmaf(-ct, ct, 1.0f));
            float sphi, cphi;
            sincos_tab(T.sc, wa, sphi, cphi);
            ux = sth * cphi;
            uy = sth * sphi;
            uz = ct;
        } else if (A.step_mode == 1) {
            // a random sign per axis, +-sqrt(2 D dt) each (kernelPhase.cuh:31-35
            // with the per-axis length that keeps D)
            const float k3 = 0.57735026f;
            ux = (wa & 1u) ? -k3 : k3;
            uy = (wa & 2u) ? -k3 : k3;
            uz = (wa & 4u) ? -k3 : k3;
        } else {
            // Gaussian step of variance 2 D dt per axis (kernelWC.cuh:56-67):
            // Box-Muller on (wa, wb) and (wc, wd); length |n| sqrt(2 D dt)
            const float r1 = sqrt_unit(neg2log_unif(wa));
            const float r2 = sqrt_unit(neg2log_unif(wc));
            float s1, c1, s2, c2;
            sincos_tab(T.sc, wb, s1, c1);
            sincos_tab(T.sc, wd, s2, c2);
            const float n1 = r1 * c1, n2 = r1 * s1, n3 = r2 * c2;
            const float nn = sqrtf(fmaf(n3, n3, fmaf(n2, n2, n1 * n1)));
            const float inv = 1.0f / nn;
            ux = n1 * inv;
            uy = n2 * inv;
            uz = n3 * inv;
            gs = nn * 0.57735026f;
        }
        float ell = gs * Lc;
        float sx = x, sy = y, sz = z;

        int lastpl = -1;     // image of the plane of the step's last reflection
        if (REJ) {
            // simple rejection (kernelLattice.cuh:73-78); a plane rejects the
            // step that would come within plane_eps of it
            const float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
            bool rej = false;
            for (int r = 0; r < nrel; ++r) {
                int mt;
                const float4 g = image_at<GEN>(T, cl[r], mt);
                const bool sk = GEN && (mt & kSphereBit);
                float tp;
                const bool cross = (GEN && (mt & kPlaneBit))
                                       ? plane_hit(g, x, y, z, sx, sy, sz, ux, uy, uz, ell, A.plane_eps, tp)
                                       : inside_at(g, sk, sx, sy, sz) != inside_at(g, sk, ex, ey, ez);
                rej = rej || cross;
            }
            if (rej) { ncoll++; ell = 0.0f; }
        } else {
            float Axy = fmaf(uy, uy, ux * ux);
            float A3 = fmaf(uz, uz, Axy);
            int nb = 0;
            int excl = -1;   // image just left outward: a convex surface cannot be hit again
            bool live = nrel > 0;
            // no lane of the wave has two or more candidates: no list loop
            const bool single = __ballot(nrel > 1) == 0ull;
            while (live) {
                // nearest crossing over the candidates (lattice.cuh:205-247)
                float best = __builtin_inff();
                int bk = -1;
                bool bins_ = false;
                // FIRST: the only test of this pass (best is still +inf and
                // every crossing time is finite), so the t < best compare
                // and its branch are left out
                auto test = [&](const int k, auto first_tag) {
                    constexpr bool FIRST = decltype(first_tag)::value;
                    int mt;
                    const float4 g = image_at<GEN>(T, k, mt);
                    if (GEN && (mt & kPlaneBit)) {
                        float t;
                        if (plane_hit(g, x, y, z, sx, sy, sz, ux, uy, uz, ell, A.plane_eps, t) && (FIRST || t < best)) {
                            best = t;
                            bk = k;
                            bins_ = false;
                        }
                        return;
                    }
                    const bool sk = GEN && (mt & kSphereBit);
                    const bool same = comp == (mt & kCompMask);
                    const Quad Q = sk ? quad_of<true>(sx - g.x, sy - g.y, sz - g.z, ux, uy, uz, A3, g.w * g.w,
                                                      same, A.H2s, ell)
                                      : quad_of<false, !GEN>(sx - g.x, sy - g.y, 0.0f, ux, uy, uz, Axy, g.w * g.w,
                                                             same, A.H2c, ell);
                    if (GEN ? quad_hits(Q, ell) : quad_hits_masked(Q, ell, ~0ull)) {
                        const float t = quad_root(Q, ell);
                        if (FIRST || t < best) {
                            best = t;
                            bk = k;
                            bins_ = Q.inside;
                        }
                    }
                };
                if (single && !GEN) {
                    // cylinders, at most one candidate per lane: nrel == 1 in
                    // every lane of this loop (live starts as nrel > 0), so
                    // the candidate is tested in every lane without branches
                    // (the surface just left outward masked out) and the pass
                    // result is set once
                    const int k = (int)cl[0];
                    int mt;
                    const float4 g = image_at<GEN>(T, k, mt);
                    const bool same = comp == (mt & kCompMask);
                    const Quad Q = quad_of<false, true>(sx - g.x, sy - g.y, 0.0f, ux, uy, uz, Axy, g.w * g.w, same,
                                                        A.H2c, ell);
                    const bool hit = quad_hits_masked(Q, ell, __builtin_amdgcn_ballot_w64(k != excl));
                    float t = ell;
                    if (hit) t = quad_root(Q, ell);
                    best = hit ? t : __builtin_inff();
                    bk = hit ? k : -1;
                    bins_ = hit && Q.inside;
                } else if (single) {
                    if (nrel > 0 && (int)cl[0] != excl) test(cl[0], std::true_type{});
                } else {
                    for (int r = 0; r < nrel; ++r) {
                        const int k = cl[r];
                        if (k != excl) test(k, std::false_type{});
                    }
                }
                live = bk >= 0;
                if (!live) break;
                int mt;
                const float4 g = image_at<GEN>(T, bk, mt);
                const bool sk = GEN && (mt & kSphereBit);
                const bool pl = GEN && (mt & kPlaneBit);
                const int pcomp = mt & kCompMask;
                const float t = best;
                ncoll++;
                const float px = fmaf(t, ux, sx), py = fmaf(t, uy, sy), pz = fmaf(t, uz, sz);
                const bool ins = bins_;
                bool transmit = false;
                if (permeable && !pl) {     // planes never transmit (plane.cuh:39)
                    // transmission with p = 4 kappa / v (boundaryCheck.cuh:94):
                    // the budget, a 2 Exp(1) draw, pays this side's hazard
                    // h = -2 ln(1 - p); the event transmits when the budget
                    // runs out, which happens with probability p whatever
                    // came before (the exponential has no memory)
                    const float2 hz = T.thr[bk];   // this surface's (from inside, from outside), from LDS
                    bud = bud - (ins ? hz.x : hz.y);
                    transmit = bud <= 0.0f;
                }
                lastpl = pl ? bk : -1;
                // left outward (outside reflection, transmission from inside,
                // plane reflection): the next scan skips this surface
                excl = (pl || ins == transmit) ? bk : -1;
                if (transmit) {
                    // continue on the far side with its speed (boundaryCheck.cuh:104-111)
                    const int nc = ins ? 0 : pcomp;
                    const float f = (ell - t) / (gs * Lc);
                    const float ph = 1.0f - f;
                    E = fmaf((float)(i - li) + (ph - lf), A.rate[comp], E);
                    li = i;
                    lf = ph;
                    comp = nc;
                    Lc = A.compL[nc];
                    ell = f * (gs * Lc);
                    ntran++;
                    // the next budget: transmission ordinal ntran
                    bud = neg2log_unif(philox10(wid, ntran, kStreamColl, A.inst, A.k0, A.k1).x);
                } else {
                    // specular reflection about the surface normal (p - c) / R, or
                    // the plane's normal (boundaryCheck.cuh:59-62, :96-103,
                    // cylinderXY.cuh:125-131, plane.cuh:32)
                    float nx, ny, nz, kd;
                    if (pl) {
                        nx = g.x; ny = g.y; nz = g.z;
                        kd = fmaf(nz, uz, fmaf(ny, uy, nx * ux));
                    } else {
                        const float inv = T.inv[bk];   // 1.0f / R
                        nx = (px - g.x) * inv;
                        ny = (py - g.y) * inv;
                        nz = sk ? (pz - g.z) * inv : 0.0f;
                        kd = sk ? fmaf(nz, uz, fmaf(ny, uy, nx * ux)) : fmaf(ny, uy, nx * ux);
                    }
                    const float k2 = kd + kd;
                    ux = fmaf(-k2, nx, ux);
                    uy = fmaf(-k2, ny, uy);
                    uz = fmaf(-k2, nz, uz);
                    Axy = fmaf(uy, uy, ux * ux);
                    A3 = fmaf(uz, uz, Axy);
                    ell = ell - t;
                }
                sx = px; sy = py; sz = pz;
                if (++nb >= cap) {
                    ncap++;
                    ell = 0.0f;
                    live = false;
                } else if (nrel == 1) {
                    if (excl >= 0) {
                        live = false;   // the only surface within reach was just left
                    } else {
                        // the only surface within reach, from inside: decide here
                        // whether the rest of the step hits it again, with the
                        // operands the next pass's test would use, so a wave
                        // whose lanes all miss leaves the loop without that pass
                        const bool same2 = comp == pcomp;
                        const Quad Q2 = sk ? quad_of<true>(sx - g.x, sy - g.y, sz - g.z, ux, uy, uz, A3, g.w * g.w,
                                                           same2, A.H2s, ell)
                                           : quad_of<false, !GEN>(sx - g.x, sy - g.y, 0.0f, ux, uy, uz, Axy,
                                                                  g.w * g.w, same2, A.H2c, ell);
                        live = GEN ? quad_hits(Q2, ell) : quad_hits_masked(Q2, ell, ~0ull);
                    }
                }
            }
        }
        float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
        if (GEN && lastpl >= 0) {
            // after a plane reflection keep at least plane_eps from the plane on
            // the walker's side, so that rounding cannot put the next start on
            // the far side
            int mt;
            const float4 g = image_at<GEN>(T, lastpl, mt);
            const float s0 = fmaf(g.z, z, fmaf(g.y, y, g.x * x)) - g.w;
            const float se = fmaf(g.z, ez, fmaf(g.y, ey, g.x * ex)) - g.w;
            const float sg = s0 >= 0.0f ? 1.0f : -1.0f;
            const float gap = A.plane_eps - sg * se;
            if (gap > 0.0f) {
                ex = fmaf(sg * gap, g.x, ex);
                ey = fmaf(sg * gap, g.y, ey);
                ez = fmaf(sg * gap, g.z, ez);
            }
        }
        dx = dx + (ex - x);
        dy = dy + (ey - y);
        dz = dz + (ez - z);
        x = wrap ? wrap1(ex, A.a) : ex;
        y = wrap ? wrap1(ey, A.b) : ey;
        z = (wrap && GEN) ? wrap1(ez, A.c) : ez;   // cylinders only: z enters no geometry
        acc.step(i, dx, dy, dz, comp);
    };

    // Steps are paired: one Philox call, counter (walker, (i-1)/2, 0, instance),
    // gives words (x, y) to step i with i-1 even and (z, w) to the next one.
    // Gaussian steps take two more words each from a second call, counter
    // (walker, (i-1)/2, 3 << 24, instance).
    auto gauss = [&](const int ii) {
        u32x4 g{0u, 0u, 0u, 0u};
        if (ALT && A.step_mode == 2)
            g = philox10(wid, (uint32_t)(ii - 1) >> 1, kStreamGauss, A.inst, A.k0, A.k1);
        return g;
    };
    if (A.i0 < 1 && A.i1 > 0) acc.at_start(dx, dy, dz, comp);
    int i = A.i0 < 1 ? 1 : A.i0;
    if (Acc::kOneSite) {
        // one step per iteration (draws and lookup on the first step of each
        // pair), so that a heavy per-step update is inlined once
        u32x4 o{0u, 0u, 0u, 0u}, g{0u, 0u, 0u, 0u};
        Cands c{0, T.cand};
        const int ifirst = i;
        for (; i < A.i1; ++i) {
            const bool second = ((i - 1) & 1) != 0;
            if (!second || i == ifirst) {
                o = philox10(wid, (uint32_t)(i - 1) >> 1, kStreamStep, A.inst, A.k0, A.k1);
                g = gauss(i);
                if (GM == 1) c = lookup();
            }
            if (GM != 1 && (((i - 1) & gm) == 0 || i == ifirst)) c = lookup();
            step(i, second ? o.z : o.x, second ? o.w : o.y, second ? g.z : g.x, second ? g.w : g.y, c,
                 ((i - 1) & gm) == gm);
        }
    } else {
        if (((i - 1) & 1) && i < A.i1) {
            const u32x4 o = philox10(wid, (uint32_t)(i - 1) >> 1, kStreamStep, A.inst, A.k0, A.k1);
            const u32x4 g = gauss(i);
            step(i, o.z, o.w, g.z, g.w, lookup(), GM == 1 || ((i - 1) & gm) == gm);
            ++i;
        }
        // pairs: a lookup at each group's first pair and at the loop's first
        Cands c{0, T.cand};
        const int ipair = i;
        for (; i + 1 < A.i1; i += 2) {
            const u32x4 o = philox10(wid, (uint32_t)(i - 1) >> 1, kStreamStep, A.inst, A.k0, A.k1);
            const u32x4 g = gauss(i);
            if (GM == 1 || ((i - 1) & gm) == 0 || i == ipair) c = lookup();
            step(i, o.x, o.y, g.x, g.y, c, false);
            step(i + 1, o.z, o.w, g.z, g.w, c, GM == 1 || (i & gm) == gm);
        }
        if (i < A.i1) {
            const u32x4 o = philox10(wid, (uint32_t)(i - 1) >> 1, kStreamStep, A.inst, A.k0, A.k1);
            const u32x4 g = gauss(i);
            step(i, o.x, o.y, g.x, g.y, lookup(), false);
        }
    }

    // Re-derive the state addresses after the loop (the empty asm hides
    // their equality with the pre-loop ones, so no 64-bit address stays live
    // across the step loop).  The base is wave-uniform (one instance per
    // wave); readfirstlane puts it in SGPRs whatever the compiler holds it in.
    const uint64_t fsu = reinterpret_cast<uint64_t>(fs);
    float* fso = reinterpret_cast<float*>(
        ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(fsu >> 32)) << 32) |
        (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)fsu));
    uint32_t wo2 = wo;
    asm volatile("" : "+s"(fso));
    asm volatile("" : "+v"(wo2));
    int* iso = reinterpret_cast<int*>(fso);
    uint32_t* sto = reinterpret_cast<uint32_t*>(fso);
    (fso + F_X * N)[wo2] = x; (fso + F_Y * N)[wo2] = y; (fso + F_Z * N)[wo2] = z;
    (fso + F_DX * N)[wo2] = dx; (fso + F_DY * N)[wo2] = dy; (fso + F_DZ * N)[wo2] = dz;
    (fso + F_E * N)[wo2] = E; (fso + F_LF * N)[wo2] = lf;
    (iso + F_COMP * N)[wo2] = comp; (iso + F_LI * N)[wo2] = li;
    (sto + F_NCOLL * N)[wo2] = ncoll; (sto + F_NTRAN * N)[wo2] = ntran;
    (sto + F_NCAP * N)[wo2] = ncap;
    (iso + F_STEP * N)[wo2] = A.i1;
    (fso + F_BUD * N)[wo2] = bud;
    acc.store(fso, N, wo2);
}

// Register budget per instantiation (waves per SIMD the compiler must fit;
// 1 = its own choice).  The isotropic 8-column kernels with up to two
// moment axes come out at 81-83 VGPRs, just past the 80 that six waves per
// SIMD allow; held to 80 they spill at most a few dwords outside the step
// loop, and the hexagonal workload's walk runs 4.5 % faster (2.18e11 ->
// 2.28e11 walker-steps/s).
constexpr int walk_waves(int P, int AX, bool ALT) { return (P == 8 && AX <= 5 && !ALT) ? 6 : 1; }

// Wait until chunk k - 1 of group g is published (lane 0 polls; bounded by
// Q.wait_ticks of real time), then acquire: the group's state written by the
// wave that walked chunk k - 1, on any XCD, is visible to the loads that follow.
__device__ __forceinline__ bool wait_group(const WalkQueue& Q, uint32_t g, uint32_t k, uint32_t lane) {
    uint32_t ok = 1u;
    if (lane == 0u) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while (__hip_atomic_load(Q.q + 1 + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
            __builtin_amdgcn_s_sleep(4);
            if (__builtin_amdgcn_s_memrealtime() - t0 > (uint64_t)Q.wait_ticks) {
                ok = 0u;
                break;
            }
        }
    }
    ok = (uint32_t)__builtin_amdgcn_readfirstlane((int)ok);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok != 0u;
}

// Persistent walk: the grid is at most what stays resident; the substrate
// tables are staged once per workgroup, then each wave takes tasks (group,
// chunk) until the counter passes Q.n_tasks; make_acc() gives a task's
// accumulator.  With one chunk per walk there are no waits and nothing to
// publish.  The queue fields and each wave's task sit in LDS and are read
// back around every task (the empty asm with a memory clobber forces the
// reads), so none of them holds an SGPR through the step loop, whose SGPR
// budget is already full.
constexpr int kWalkStaticLds = 64 + 4 * 16;   // static LDS of walk_tasks (bytes)

template <class Acc, bool GEN, bool REJ, bool ALT, int GM, class MakeAcc>
__device__ __forceinline__ void walk_tasks(const WalkArgs& A, const WalkQueue& Q, unsigned char* lds,
                                           MakeAcc make_acc) {
    __shared__ WalkQueue s_q;
    __shared__ int s_range[2];
    __shared__ uint32_t s_task[16];
    static_assert(sizeof(WalkQueue) + sizeof(s_range) <= 64, "static LDS of walk_tasks");
    if (threadIdx.x == 0) {
        s_q = Q;
        s_range[0] = A.i0;
        s_range[1] = A.i1;
    }
    const LdsTables T = stage_tables<GEN>(A, lds);   // ends with a barrier
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wv = threadIdx.x >> 6;
    for (;;) {
        asm volatile("" ::: "memory");
        const uint32_t n_groups = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_q.n_groups);
        const uint32_t n_tasks = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_q.n_tasks);
        uint32_t t = 0u;
        if (lane == 0u) t = atomicAdd(s_q.q, 1u);
        t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
        if (t >= n_tasks) break;
        const uint32_t k = t / n_groups;
        const uint32_t g = t - k * n_groups;
        if (k > 0u && !wait_group(s_q, g, k, lane)) {
            if (lane == 0u) atomicOr(A.err, kErrWaitLimit);
            break;
        }
        if (lane == 0u) s_task[wv] = t;
        // chunk k: steps [b + k chunk, b + (k + 1) chunk) with b = max(i0, 1)
        // (step pairs start at odd steps), chunk 0 from i0 (step 0 is the
        // accumulator's t = 0 term)
        WalkArgs B = A;
        const int chunk = __builtin_amdgcn_readfirstlane(s_q.chunk);
        const int r0 = __builtin_amdgcn_readfirstlane(s_range[0]);
        const int b = max(r0, 1);
        B.i0 = k == 0u ? r0 : b + (int)k * chunk;
        B.i1 = min(__builtin_amdgcn_readfirstlane(s_range[1]), b + (int)(k + 1u) * chunk);
        const uint32_t w = g * 64u + lane;
        if ((long long)w < A.N) {
            Acc acc = make_acc();
            walk_lane<Acc, GEN, REJ, ALT, GM>(B, T, acc, w);
        }
        asm volatile("" ::: "memory");
        const uint32_t ng = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_q.n_groups);
        if ((uint32_t)__builtin_amdgcn_readfirstlane((int)s_q.n_tasks) > ng) {
            // release: this wave's state stores reach every XCD before the
            // group's next chunk can be taken up
            const uint32_t tt = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_task[wv]);
            const uint32_t kk = tt / ng;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (lane == 0u)
                __hip_atomic_store(s_q.q + 1 + (tt - kk * ng), kk + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// GM: lookup-group mask fixed at compile time (1, 3 or 7), or 0 to read it
// from the arguments (the axis-sign and Gaussian walks, where either occurs)
template <int P, int AX, bool GEN, bool REJ, bool ALT, int GM>
__global__ __launch_bounds__(walk_max_block(P)) __attribute__((amdgpu_waves_per_eu(walk_waves(P, AX, ALT)))) void
walk_kernel(WalkArgs A, WalkQueue Q, const float* __restrict__ prof_g) {
    extern __shared__ __align__(16) unsigned char lds[];
    walk_tasks<PhaseAcc<P, AX>, GEN, REJ, ALT, GM>(A, Q, lds, [&]() {
        PhaseAcc<P, AX> acc;
        acc.prof = (const_fptr)prof_g;
        return acc;
    });
}

}  // namespace axx
