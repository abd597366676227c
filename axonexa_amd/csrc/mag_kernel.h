// mag_kernel.h — the magnetization (Bloch) walk: the walk of walk_kernel.h
// with a per-step fp64 magnetization update in place of the phase moments
// (updateWalkersMag, src/kernels/kernelMag.cuh:1-141).
//
// Field tables: for step i (t_i = i dt, i = 0 .. n_steps-1) with s_i
// substeps of h = dt / s_i the caller samples every field at
//   rotation: t_i + k h            (k < s_i)
//   RK4:      t_i + k h / 2        (k <= 2 s_i)
//   DOPRI:    t_i + (k + c_m) h    (k < s_i, c = 0, 1/5, 3/10, 4/5, 8/9, 1)
// as records (b0[3], J[3][3]) with B(r) = b0 + J (r - origin), r the
// walker's unwrapped position after step i (its start for i = 0).  off[i]
// is the first sample of step i; sample k of field j is the 12 doubles at
// tab + ((off[i] + k) * F + j) * 12.  All indices are wave-uniform, so the
// records come through scalar loads.
#pragma once
#include "mag_math.h"
#include "walk_kernel.h"

namespace axx {

constexpr int kMagRotate = 0, kMagRK4 = 1, kMagDopri = 2;
constexpr int kMagGroup = 8;       // fields carried per walk
constexpr int kMagBlock = 256;

typedef const __attribute__((address_space(4))) double* const_dptr;
typedef const __attribute__((address_space(4))) int* const_iptr;

struct MagArgs {
    const double* tab;
    const int* off;       // [n_steps + 1]
    const int* nsub;      // [n_steps]
    const double* iT1;    // [n_prims + 1], 0 = no longitudinal relaxation
    const double* iT2;    // [n_prims + 1]
    double dt, gamma, ox, oy, oz;
    double m0x, m0y, m0z;
    int F;                // fields of this group (= the kernel's MB)
    int relax;            // relaxation terms in RK4 / DOPRI
};

inline int mag_samples(int integ, int s) {
    return integ == kMagRotate ? s : (integ == kMagRK4 ? 2 * s + 1 : 6 * s);
}

// Magnetization of MB fields in double precision; word F_MOM + 6 j + 2 c
// holds the low half of component c of field j, the next word the high half.
template <int MB, int INTEG>
struct MagAcc {
    static constexpr bool kOneSite = true;
    const_dptr tab;
    const_iptr off;
    const_iptr nsub;
    const double* iT1;
    const double* iT2;
    double dt, gamma, ox, oy, oz;
    bool relax;
    float x0, y0, z0;
    d3 M[MB];

    __device__ __forceinline__ void init(const MagArgs& G) {
        tab = (const_dptr)G.tab;
        off = (const_iptr)G.off;
        nsub = (const_iptr)G.nsub;
        iT1 = G.iT1;
        iT2 = G.iT2;
        dt = G.dt; gamma = G.gamma;
        ox = G.ox; oy = G.oy; oz = G.oz;
        relax = G.relax != 0;
    }
    __device__ __forceinline__ void load(const float* fs, long long N, uint32_t wo) {
        x0 = (fs + F_X0 * N)[wo]; y0 = (fs + F_Y0 * N)[wo]; z0 = (fs + F_Z0 * N)[wo];
        const uint32_t* st = reinterpret_cast<const uint32_t*>(fs);
#pragma unroll
        for (int j = 0; j < MB; ++j) {
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const uint32_t lo = (st + (F_MOM + 6 * j + 2 * c) * N)[wo];
                const uint32_t hi = (st + (F_MOM + 6 * j + 2 * c + 1) * N)[wo];
                v[c] = __hiloint2double((int)hi, (int)lo);
            }
            M[j] = d3{v[0], v[1], v[2]};
        }
    }
    __device__ __forceinline__ void store(float* fs, long long N, uint32_t wo) const {
        uint32_t* st = reinterpret_cast<uint32_t*>(fs);
#pragma unroll
        for (int j = 0; j < MB; ++j) {
            const double v[3] = {M[j].x, M[j].y, M[j].z};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const unsigned long long b = (unsigned long long)__double_as_longlong(v[c]);
                (st + (F_MOM + 6 * j + 2 * c) * N)[wo] = (uint32_t)b;
                (st + (F_MOM + 6 * j + 2 * c + 1) * N)[wo] = (uint32_t)(b >> 32);
            }
        }
    }
    __device__ __forceinline__ void update(int i, float dx, float dy, float dz, int comp) {
        const double rx = ((double)x0 + (double)dx) - ox;
        const double ry = ((double)y0 + (double)dy) - oy;
        const double rz = ((double)z0 + (double)dz) - oz;
        const int s0 = off[i];
        const int ns = nsub[i];
        const double h = dt / (double)ns;
        double i1 = 0.0, i2 = 0.0;
        if (INTEG != kMagRotate && relax) {
            i1 = iT1[comp];
            i2 = iT2[comp];
        }
        if (INTEG == kMagRotate) {
            // substeps outside, fields inside: the MB rotations of a substep
            // are independent chains the compiler interleaves
            for (int k = 0; k < ns; ++k) {
#pragma unroll
                for (int j = 0; j < MB; ++j) {
                    const d3 B = field_at(tab + ((size_t)(s0 + k) * MB + j) * 12, rx, ry, rz);
                    rotate(M[j], B, gamma * h);
                }
            }
        } else {
            // one field at a time: the stages carry enough independent work
            // and interleaving fields would exhaust the registers
#pragma unroll
            for (int j = 0; j < MB; ++j) {
                for (int k = 0; k < ns; ++k) {
                    if (INTEG == kMagRK4) {
                        const int q = s0 + 2 * k;
                        const d3 B0 = field_at(tab + ((size_t)q * MB + j) * 12, rx, ry, rz);
                        const d3 Bh = field_at(tab + ((size_t)(q + 1) * MB + j) * 12, rx, ry, rz);
                        const d3 B1 = field_at(tab + ((size_t)(q + 2) * MB + j) * 12, rx, ry, rz);
                        rk4(M[j], B0, Bh, B1, gamma, h, relax, i1, i2);
                    } else {
                        d3 B[6];
#pragma unroll
                        for (int c = 0; c < 6; ++c)
                            B[c] = field_at(tab + ((size_t)(s0 + 6 * k + c) * MB + j) * 12, rx, ry, rz);
                        dopri5(M[j], B, gamma, h, relax, i1, i2);
                    }
                }
            }
        }
    }
    __device__ __forceinline__ void at_start(float dx, float dy, float dz, int comp) { update(0, dx, dy, dz, comp); }
    __device__ __forceinline__ void step(int i, float dx, float dy, float dz, int comp) { update(i, dx, dy, dz, comp); }
};

// Register budget as waves per SIMD.  Four (at most 128 VGPRs): up to four
// fields the kernel fits by itself; with more the compiler would take
// 129-169 and drop to three waves, which measured slower than the few spills
// of the budget (8 fields: rotation +1 %, RK4 +16 %, Dormand-Prince +39 %).
// Five (96 VGPRs, 16-36 B of spills) for one or two fields with rotation or
// RK4: +2-4 % (Dormand-Prince, whose stages hold more live values, loses 3 %).
template <int MB, int INTEG>
constexpr int mag_waves() {
    return (MB <= 2 && INTEG != kMagDopri) ? 5 : 4;
}

// ALT: the axis-sign and Gaussian step generators compiled in (isotropic
// runs take the instantiation without them: 1-3 % faster).
template <int MB, int INTEG, bool GEN, bool REJ, bool ALT>
__global__ __launch_bounds__(kMagBlock) __attribute__((amdgpu_waves_per_eu(mag_waves<MB, INTEG>()))) void
mag_walk_kernel(WalkArgs A, MagArgs G, WalkQueue Q) {
    extern __shared__ __align__(16) unsigned char lds[];
    // spheres and planes (GEN) always take lookup groups of two
    walk_tasks<MagAcc<MB, INTEG>, GEN, REJ, ALT, GEN ? 1 : 0>(A, Q, lds, [&]() {
        MagAcc<MB, INTEG> acc;
        acc.init(G);
        return acc;
    });
}

typedef void (*mag_fn)(WalkArgs, MagArgs, WalkQueue);
mag_fn pick_mag_rotate(int mb, bool gen, bool rej, bool alt);
mag_fn pick_mag_rk4(int mb, bool gen, bool rej, bool alt);
mag_fn pick_mag_dopri(int mb, bool gen, bool rej, bool alt);

}  // namespace axx
