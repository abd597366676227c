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
// (src/kernels/boundaryCheck.cuh:138-217) and the Lattice scans
// (src/substrate/lattice.cuh:90-247).  Differences by design: Philox
// counters instead of per-thread XORWOW state, moments per (profile, axis)
// instead of M global read-modify-writes per step, LDS cell bins instead of
// the O(n_basis) scan, the segment form (start, direction, remaining length)
// of the reflection loop, side-tracked crossing tests, and true T2 time split.
#pragma once
#include "device_math.h"

namespace axx {

struct WalkArgs {
    float a, b, c;
    float H2c, H2s;                         // squared half of the smallest cell side (xy; xyz)
    int nbx, nby, nbz;
    float ibx, iby, ibz;
    int n_img, n_bins, n_cand;
    const float4* __restrict__ g_geom;      // [n_img] cx, cy, cz, R^2
    const int* __restrict__ g_meta;         // [n_img] compartment | kSphereBit
    const uint32_t* __restrict__ g_bins;    // [n_bins] offset << 8 | count
    const uint16_t* __restrict__ g_cand;    // [n_cand] image indices, ascending per bin
    const float* __restrict__ compL;        // [n_comp] step length
    const float* __restrict__ rate;         // [n_comp] dt / T2
    const float* __restrict__ thr_out;      // [n_prims] 4 kappa / v_inside
    const float* __restrict__ thr_in;       // [n_prims] 4 kappa / v_lattice
    const float* __restrict__ prof;         // [n_steps][P] (P = template width)
    uint32_t* __restrict__ state;           // SoA, AXX_STATE_FIELDS(P) x N
    long long N;
    uint32_t k0, k1, inst, woff;
    int mode, init_mode, init_arg, init_img;
    int i0, i1, cap;
    int* __restrict__ err;
};

constexpr int kSphereBit = 1 << 30;
constexpr int kCompMask = (1 << 30) - 1;

// state field indices (include/axonexa.h)
enum {
    F_X = 0, F_Y, F_Z, F_DX, F_DY, F_DZ, F_X0, F_Y0, F_Z0, F_E, F_LF,
    F_COMP, F_COMP0, F_LI, F_NCOLL, F_NTRAN, F_NCAP, F_STEP, F_MOM
};

__device__ __forceinline__ bool inside_at(float4 g, bool sph, float x, float y, float z) {
    float ox = x - g.x, oy = y - g.y;
    if (sph) {
        float oz = z - g.z;
        return fmaf(oz, oz, fmaf(oy, oy, ox * ox)) < g.w;
    }
    return fmaf(oy, oy, ox * ox) < g.w;
}

// Region of a point: compartment of the containing primitive (unique by the
// non-overlap rule), else the lattice (lattice.cuh:134-148).
__device__ __forceinline__ int region_of(const float4* geom, const int* meta, int n_img, float x,
                                         float y, float z) {
    for (int i = 0; i < n_img; ++i) {
        int mt = meta[i];
        if (inside_at(geom[i], (mt & kSphereBit) != 0, x, y, z)) return mt & kCompMask;
    }
    return 0;
}

template <int P, int AX, bool SPH>
__global__ __launch_bounds__(1024) void walk_kernel(WalkArgs A) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* geom = reinterpret_cast<float4*>(lds);
    int* meta = reinterpret_cast<int*>(geom + A.n_img);
    uint32_t* bins = reinterpret_cast<uint32_t*>(meta + A.n_img);
    uint16_t* cand = reinterpret_cast<uint16_t*>(bins + A.n_bins);
    for (int i = threadIdx.x; i < A.n_img; i += blockDim.x) {
        geom[i] = A.g_geom[i];
        meta[i] = A.g_meta[i];
    }
    for (int i = threadIdx.x; i < A.n_bins; i += blockDim.x) bins[i] = A.g_bins[i];
    for (int i = threadIdx.x; i < A.n_cand; i += blockDim.x) cand[i] = A.g_cand[i];
    __syncthreads();

    const long long N = A.N;
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= N) return;
    uint32_t* st = A.state;
    float* fs = reinterpret_cast<float*>(st);
    int* is = reinterpret_cast<int*>(st);
    const uint32_t wid = A.woff + (uint32_t)w;

    float x, y, z, dx, dy, dz, E, lf;
    int comp, li;
    uint32_t ncoll, ntran, ncap;
    float m[P][3];

    if (A.i0 == 0) {
        // Initial placement (kernelLattice.cuh:29-37, lattice.cuh:31-58,
        // cylinderXY.cuh:176-183), all draws on the init stream.
        bool done = false;
        x = y = z = 0.0f;
        comp = 0;
        if (A.init_mode == 0) {
            u32x4 o = philox10(wid, 0u, kStreamInit, A.inst, A.k0, A.k1);
            x = unif(o.x) * A.a;
            y = unif(o.y) * A.b;
            z = unif(o.z) * A.c;
            comp = region_of(geom, meta, A.n_img, x, y, z);
            done = true;
        } else {
            for (uint32_t it = 0; it < (1u << 20) && !done; ++it) {
                u32x4 o = philox10(wid, it, kStreamInit, A.inst, A.k0, A.k1);
                if (A.init_mode == 1) {
                    float px = unif(o.x) * A.a, py = unif(o.y) * A.b, pz = unif(o.z) * A.c;
                    int cm = region_of(geom, meta, A.n_img, px, py, pz);
                    if (cm != A.init_arg) continue;
                    x = px; y = py; z = pz; comp = cm;
                    done = true;
                } else {
                    float4 gk = geom[A.init_img];
                    bool sk = (meta[A.init_img] & kSphereBit) != 0;
                    float Rf = sqrtf(gk.w);
                    float px = gk.x + (unif(o.x) * 2.0f - 1.0f) * Rf;
                    float py = gk.y + (unif(o.y) * 2.0f - 1.0f) * Rf;
                    float pz = gk.z + (unif(o.z) * 2.0f - 1.0f) * Rf;
                    if (!inside_at(gk, sk, px, py, pz)) continue;
                    px = wrap1(px, A.a); py = wrap1(py, A.b); pz = wrap1(pz, A.c);
                    x = px; y = py; z = pz;
                    comp = region_of(geom, meta, A.n_img, px, py, pz);
                    done = true;
                }
            }
        }
        if (!done) atomicOr(A.err, 1);
        dx = dy = dz = 0.0f;
        E = 0.0f; lf = 0.0f; li = 1;
        ncoll = ntran = ncap = 0u;
#pragma unroll
        for (int p = 0; p < P; ++p) m[p][0] = m[p][1] = m[p][2] = 0.0f;
        fs[(size_t)F_X0 * N + w] = x;
        fs[(size_t)F_Y0 * N + w] = y;
        fs[(size_t)F_Z0 * N + w] = z;
        is[(size_t)F_COMP0 * N + w] = comp;
    } else {
        x = fs[(size_t)F_X * N + w]; y = fs[(size_t)F_Y * N + w]; z = fs[(size_t)F_Z * N + w];
        dx = fs[(size_t)F_DX * N + w]; dy = fs[(size_t)F_DY * N + w]; dz = fs[(size_t)F_DZ * N + w];
        E = fs[(size_t)F_E * N + w]; lf = fs[(size_t)F_LF * N + w];
        comp = is[(size_t)F_COMP * N + w]; li = is[(size_t)F_LI * N + w];
        ncoll = st[(size_t)F_NCOLL * N + w]; ntran = st[(size_t)F_NTRAN * N + w];
        ncap = st[(size_t)F_NCAP * N + w];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            m[p][0] = (AX & 1) ? fs[(size_t)(F_MOM + 3 * p + 0) * N + w] : 0.0f;
            m[p][1] = (AX & 2) ? fs[(size_t)(F_MOM + 3 * p + 1) * N + w] : 0.0f;
            m[p][2] = (AX & 4) ? fs[(size_t)(F_MOM + 3 * p + 2) * N + w] : 0.0f;
        }
    }

    const float a = A.a, b = A.b, c = A.c;
    const int cap = A.cap;
    const bool permeable = A.mode == 1;
    const bool rejection = A.mode == 2;
    float Lc = A.compL[comp];
    const int ibeg = A.i0 < 1 ? 1 : A.i0;
    u32x4 o = {0u, 0u, 0u, 0u};

    for (int i = ibeg; i < A.i1; ++i) {
        const uint32_t h = (uint32_t)(i - 1) & 1u;
        if (h == 0u || i == ibeg) o = philox10(wid, (uint32_t)(i - 1) >> 1, kStreamStep, A.inst, A.k0, A.k1);
        const uint32_t wa = h ? o.z : o.x;
        const uint32_t wb = h ? o.w : o.y;
        // isotropic direction (sphere point picking, kernelLattice.cuh:52-57)
        const float ct = unif(wb) * 2.0f - 1.0f;
        const float sth = sqrtf((1.0f - ct) * (1.0f + ct));
        float sphi, cphi;
        sincos_turn(wa, sphi, cphi);
        float ux = sth * cphi, uy = sth * sphi, uz = ct;
        float ell = Lc;
        float sx = x, sy = y, sz = z;
        // candidate surfaces of the start bin
        int bx = (int)(x * A.ibx), by = (int)(y * A.iby), bz = SPH ? (int)(z * A.ibz) : 0;
        bx = bx > A.nbx - 1 ? A.nbx - 1 : bx;
        by = by > A.nby - 1 ? A.nby - 1 : by;
        bz = bz > A.nbz - 1 ? A.nbz - 1 : bz;
        bx = bx < 0 ? 0 : bx;
        by = by < 0 ? 0 : by;
        bz = bz < 0 ? 0 : bz;
        const uint32_t rec = bins[(bz * A.nby + by) * A.nbx + bx];
        const int cnt = (int)(rec & 0xFFu);
        const uint16_t* cl = cand + (rec >> 8);

        if (rejection) {
            const float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
            bool rej = false;
            for (int j = 0; j < cnt && !rej; ++j) {
                const int k = cl[j];
                const float4 g = geom[k];
                const bool sk = SPH && (meta[k] & kSphereBit);
                rej = inside_at(g, sk, sx, sy, sz) != inside_at(g, sk, ex, ey, ez);
            }
            if (rej) { ncoll++; ell = 0.0f; }
        } else if (cnt > 0) {
            float Axy = fmaf(uy, uy, ux * ux);
            float A3 = fmaf(uz, uz, Axy);
            int nb = 0, kc = 0;
            for (;;) {
                float best = __builtin_inff();
                int bk = -1;
                bool bins_ = false;
                for (int j = 0; j < cnt; ++j) {
                    const int k = cl[j];
                    const float4 g = geom[k];
                    const int mt = meta[k];
                    const bool sk = SPH && (mt & kSphereBit);
                    const bool same = comp == (mt & kCompMask);
                    float t;
                    bool hit, ins;
                    if (sk)
                        hit = hit_test<true>(sx - g.x, sy - g.y, sz - g.z, ux, uy, uz, A3, g.w, same,
                                             A.H2s, ell, t, ins);
                    else
                        hit = hit_test<false>(sx - g.x, sy - g.y, 0.0f, ux, uy, uz, Axy, g.w, same,
                                              A.H2c, ell, t, ins);
                    if (hit && t < best) {
                        best = t;
                        bk = k;
                        bins_ = ins;
                    }
                }
                if (bk < 0) break;
                const float4 g = geom[bk];
                const int mt = meta[bk];
                const bool sk = SPH && (mt & kSphereBit);
                const int pcomp = mt & kCompMask;
                const float t = best;
                ncoll++;
                const float px = fmaf(t, ux, sx), py = fmaf(t, uy, sy), pz = fmaf(t, uz, sz);
                const bool ins = bins_;
                bool transmit = false;
                if (permeable) {
                    u32x4 oc = philox10(wid, (uint32_t)i, kStreamColl | ((uint32_t)kc >> 2), A.inst, A.k0, A.k1);
                    const uint32_t sel = (uint32_t)kc & 3u;
                    const uint32_t wv = sel == 0 ? oc.x : sel == 1 ? oc.y : sel == 2 ? oc.z : oc.w;
                    const float U = unif(wv);
                    const float thr = ins ? A.thr_out[pcomp - 1] : A.thr_in[pcomp - 1];
                    transmit = U <= thr;
                    kc++;
                }
                if (transmit) {
                    const int nc = ins ? 0 : pcomp;
                    const float rem = ell - t;
                    const float f = rem / Lc;
                    const float ph = 1.0f - f;
                    E = fmaf((float)(i - li) + (ph - lf), A.rate[comp], E);
                    li = i;
                    lf = ph;
                    comp = nc;
                    Lc = A.compL[nc];
                    ell = f * Lc;
                    ntran++;
                } else {
                    float nx = px - g.x, ny = py - g.y, nz = 0.0f, kd;
                    if (sk) {
                        nz = pz - g.z;
                        const float inv = 1.0f / sqrtf(fmaf(nz, nz, fmaf(ny, ny, nx * nx)));
                        nx = nx * inv; ny = ny * inv; nz = nz * inv;
                        kd = fmaf(nz, uz, fmaf(ny, uy, nx * ux));
                    } else {
                        const float inv = 1.0f / sqrtf(fmaf(ny, ny, nx * nx));
                        nx = nx * inv; ny = ny * inv;
                        kd = fmaf(ny, uy, nx * ux);
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
                if (++nb >= cap) { ncap++; ell = 0.0f; break; }
            }
        }
        const float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
        dx = dx + (ex - x);
        dy = dy + (ey - y);
        dz = dz + (ez - z);
        x = wrap1(ex, a);
        y = wrap1(ey, b);
        z = wrap1(ez, c);
        // phase moments (kernelLattice.cuh:81-85): wave-uniform profile samples
        const float* g = A.prof + (size_t)i * P;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const float gp = g[p];
            if (AX & 1) m[p][0] = fmaf(gp, dx, m[p][0]);
            if (AX & 2) m[p][1] = fmaf(gp, dy, m[p][1]);
            if (AX & 4) m[p][2] = fmaf(gp, dz, m[p][2]);
        }
    }

    fs[(size_t)F_X * N + w] = x; fs[(size_t)F_Y * N + w] = y; fs[(size_t)F_Z * N + w] = z;
    fs[(size_t)F_DX * N + w] = dx; fs[(size_t)F_DY * N + w] = dy; fs[(size_t)F_DZ * N + w] = dz;
    fs[(size_t)F_E * N + w] = E; fs[(size_t)F_LF * N + w] = lf;
    is[(size_t)F_COMP * N + w] = comp; is[(size_t)F_LI * N + w] = li;
    st[(size_t)F_NCOLL * N + w] = ncoll; st[(size_t)F_NTRAN * N + w] = ntran;
    st[(size_t)F_NCAP * N + w] = ncap;
    is[(size_t)F_STEP * N + w] = A.i1;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        fs[(size_t)(F_MOM + 3 * p + 0) * N + w] = m[p][0];
        fs[(size_t)(F_MOM + 3 * p + 1) * N + w] = m[p][1];
        fs[(size_t)(F_MOM + 3 * p + 2) * N + w] = m[p][2];
    }
}

}  // namespace axx
