/*
 * axx_oracle.c — CPU restatement of axonexa's random-walk phase simulator.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
 * check and bench.py's CPU-baseline leg may call it; the product library
 * (axonexa_amd/) never links or calls it.
 *
 * It restates, in plain C and fp32 arithmetic written out operation by
 * operation (built with -ffp-contract=off), the semantics of the reference's
 * lattice walk:
 *   kernelLattice.cuh:1-104           per-walker loop (init, step, phase, T2)
 *   boundaryCheck.cuh:34-64, :66-113  multi-reflection / permeable boundary
 *   lattice.cuh:31-247                periodic cell, wrap (lmod), region lookup
 *   cylinderXY.cuh:69-131, Sphere.cuh:89-149   inside / intersect / normal
 *   plane.cuh:18-39                   plane crossing, normal, no region
 *   kernelSetup.cuh:1-5               per-walker RNG (replaced: Philox4x32-10)
 *   phaseAcquisitionStream.cuh:171-189 + reduction.cuh:76-195   signal sums
 * with the deliberate, documented deviations of the build (counter-based
 * Philox, sphere-point picking without acos, segment form of the reflection
 * loop, side-tracked crossing test, true T2 time split, continued boundary
 * checks after a transmission, planes hit within plane_eps with a push off
 * the plane after reflecting).  Every function cites the reference lines it
 * follows.
 *
 * Pins (what checks this file, tests/): Philox known-answer vectors
 * (Random123) and rocRAND's engine; closed-form ray/circle and ray/sphere
 * roots in fp64; the exact free-diffusion sinc product; exp(-bD); Neuman /
 * GPD restricted limits; invariants (step length, reflection isometry,
 * lattice translation, uniform equilibrium under permeability 4*kappa/v);
 * the transmitted fraction of the boundary events from each side equals
 * 4*kappa/v of that side; the T2 exponent of every walker equals the one
 * rebuilt from its trajectory with fp64 crossing times; the ORC_PLANT
 * variants below fail those checks.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/axonexa.h"
#include "axx_oracle.h"

/* ORC_PLANT selects a deliberately wrong variant, built only by the tests
 * that show their checks catch it: 1 doubles the transmission probability,
 * 2 charges a whole crossing step's T2 decay to the far side (the
 * reference's by-value accumtime, defect D2), 3 uses the lattice speed for
 * the transmission threshold from both sides.  0 is the oracle. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Replaces cuRAND
 * XORWOW (kernelSetup.cuh:1-5, curand_uniform at kernelLattice.cuh:52-53). */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform in (0,1): odd multiples of 2^-24, exactly representable. */
static inline float unif(uint32_t w) { return (float)((w >> 8) | 1u) * 0x1p-24f; }
float orc_unif(uint32_t w) { return unif(w); }

/* Counter streams: ctr = (walker, index, stream<<24 | sub, instance). */
#define STREAM_STEP 0u
#define STREAM_COLL 1u
#define STREAM_INIT 2u
#define STREAM_GAUSS 3u

static inline void draw(uint32_t wid, uint32_t idx, uint32_t stream_sub, uint32_t inst,
                        const uint32_t key[2], uint32_t out[4]) {
    uint32_t ctr[4] = {wid, idx, stream_sub, inst};
    orc_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------ */
/* Hazard of a boundary event that transmits with probability p = 4 kappa / v
 * (boundaryCheck.cuh:94): h = -2 ln(1 - p), so that a budget drawn as
 * 2 Exp(1) = -2 ln u (u uniform) runs out at the event with probability p;
 * p >= 1: always. */
static float hazard(double p) { return p >= 1.0 ? INFINITY : (float)(-2.0 * log1p(-p)); }

/* sin/cos of 2 pi v 2^-24 for a 24-bit integer v.  Quarter-turn range
 * reduction in integers, then Taylor polynomials of sin(pi/2 r) and
 * cos(pi/2 r) on r in [-1/2, 1/2] (truncation < 2e-9).  Builds the azimuth
 * table below. */
#define S1  1.5707963267948966f
#define S3 -0.6459640975062462f
#define S5  0.07969262624616704f
#define S7 -0.004681754135318687f
#define S9  0.00016044118478735982f
#define C2 -1.2337005501361697f
#define C4  0.25366950790104802f
#define C6 -0.020863480763352960f
#define C8  0.00091926027483942659f
#define C10 -0.000025202042373060605f

void orc_sincos_u24(uint32_t v, float* s_out, float* c_out) {
    uint32_t q = (v + (1u << 21)) >> 22;         /* nearest quarter 0..4       */
    int32_t ri = (int32_t)v - (int32_t)(q << 22);
    float r = (float)ri * 0x1p-22f;              /* exact, in [-1/2, 1/2]      */
    float y = r * r;
    float ps = fmaf(y, S9, S7);
    ps = fmaf(y, ps, S5);
    ps = fmaf(y, ps, S3);
    ps = fmaf(y, ps, S1);
    float sp = r * ps;
    float pc = fmaf(y, C10, C8);
    pc = fmaf(y, pc, C6);
    pc = fmaf(y, pc, C4);
    pc = fmaf(y, pc, C2);
    float cp = fmaf(y, pc, 1.0f);
    switch (q & 3u) {
    case 0: *c_out = cp;  *s_out = sp;  break;
    case 1: *c_out = -sp; *s_out = cp;  break;
    case 2: *c_out = -cp; *s_out = -sp; break;
    default: *c_out = sp; *s_out = -cp; break;
    }
}

/* Azimuth table: entry j = (cos, sin) of 2 pi (j + 1/2) / 256 from the
 * polynomial above at v = (2j + 1) 2^15; filled when the library loads. */
static float az_cos[256], az_sin[256];
__attribute__((constructor)) static void az_init(void) {
    for (uint32_t j = 0; j < 256; ++j) orc_sincos_u24((2u * j + 1u) << 15, &az_sin[j], &az_cos[j]);
}

void orc_azimuth_table(float* c, float* s) {
    for (int j = 0; j < 256; ++j) { c[j] = az_cos[j]; s[j] = az_sin[j]; }
}

/* sin/cos of 2 pi u(w), u(w) = ((w >> 8) | 1) 2^-24, from the Philox word:
 * the table entry of the top 8 bits j of the 24-bit v, turned by
 * d = 2 pi (u - (j + 1/2)/256) = (f - 2^15) K1 (f = low 16 bits of v,
 * K1 = float(2 pi) 2^-24, |d| <= pi/256) with cos d = 1 - d^2/2 and
 * sin d = d (1 - d^2/6) (truncation < 1e-9) and the angle-sum formulas.
 * Replaces phi = 2*PI*curand_uniform; sin(phi), cos(phi)
 * (kernelLattice.cuh:52-57). */
void orc_sincos_turn(uint32_t w, float* s_out, float* c_out) {
    uint32_t v = (w >> 8) | 1u;
    uint32_t j = v >> 16;
    float d = fmaf((float)(v & 0xFFFFu), 0x1.921fb6p-22f, -0x1.921fb6p-7f);
    float d2 = d * d;
    float cd = fmaf(d2, -0.5f, 1.0f);
    float sd = d * fmaf(d2, -0.16666667f, 1.0f);
    *c_out = fmaf(-az_sin[j], sd, az_cos[j] * cd);
    *s_out = fmaf(az_cos[j], sd, az_sin[j] * cd);
}

void orc_sincos_turn_n(const uint32_t* w, float* s, float* c, int64_t n) {
    for (int64_t k = 0; k < n; ++k) orc_sincos_turn(w[k], &s[k], &c[k]);
}

/* -2 ln u(w) for Box-Muller (Gaussian steps, kernelWC.cuh:56-61): u = k 2^-24
 * with k odd, ln u = (e - 24) ln 2 + ln m for k = 2^e m, m in [sqrt(1/2),
 * sqrt(2)), ln m = 2 atanh(s), s = (m - 1) / (m + 1), |s| < 0.172, odd series
 * to s^9 (truncation < 4e-10).  Result in [1.19e-7, 33.3]. */
static inline float neg2log_unif(uint32_t w) {
    float x = (float)((w >> 8) | 1u);
    uint32_t bits;
    memcpy(&bits, &x, 4);
    int e = (int)(bits >> 23) - 127;
    uint32_t mb = (bits & 0x007FFFFFu) | 0x3F800000u;
    float m;
    memcpy(&m, &mb, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = fmaf(s2, 0.11111111f, 0.14285714f);
    p = fmaf(s2, p, 0.2f);
    p = fmaf(s2, p, 0.33333334f);
    p = fmaf(s2, p, 1.0f);
    float lnm = (2.0f * s) * p;
    float lnu = fmaf((float)(e - 24), 0.69314718f, lnm);
    return -2.0f * lnu;
}

void orc_neg2log_unif(const uint32_t* w, float* out, int64_t n) {
    for (int64_t k = 0; k < n; ++k) out[k] = neg2log_unif(w[k]);
}


/* ------------------------------------------------------------------ */
/* Magnetization (Bloch) path in double precision (blochdiff.cuh:7-210,
 * kernelMag.cuh:1-141); same operations in the same order as the kernel
 * (axonexa_amd/csrc/mag_math.h), so each walker's M matches bit for bit. */
typedef struct { double x, y, z; } d3;

/* sqrt and a residual test of its two neighbours (libm's sqrt is already
 * correctly rounded; the test then keeps it) */
static inline double sqrt_cr(double v) {
    double y = sqrt(v);
    int64_t b;
    memcpy(&b, &y, 8);
    int64_t bd = b - 1, bu = b + 1;
    double ydn, yup;
    memcpy(&ydn, &bd, 8);
    memcpy(&yup, &bu, 8);
    double rdn = fma(-ydn, y, v), rup = fma(-yup, y, v);
    y = rdn <= 0.0 ? ydn : y;
    y = rup > 0.0 ? yup : y;
    return y;
}

/* sin, cos: Cody-Waite reduction by pi/2 (three fused parts), Taylor
 * polynomials of sin to r^17 and cos to r^16 on |r| <= pi/4 */
static inline void sincos64(double x, double* s, double* c) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = fma(-k, 0x1.921fb54442d18p+0, x);
    r = fma(-k, 0x1.1a62633145c07p-54, r);
    r = fma(-k, -0x1.f1976b7ed8fbcp-110, r);
    const double z = r * r;
    double ps = 0x1.952c77030ad4ap-49;
    ps = fma(z, ps, -0x1.ae7f3e733b81fp-41);
    ps = fma(z, ps, 0x1.6124613a86d09p-33);
    ps = fma(z, ps, -0x1.ae64567f544e4p-26);
    ps = fma(z, ps, 0x1.71de3a556c734p-19);
    ps = fma(z, ps, -0x1.a01a01a01a01ap-13);
    ps = fma(z, ps, 0x1.1111111111111p-7);
    ps = fma(z, ps, -0x1.5555555555555p-3);
    const double sr = fma(r * z, ps, r);
    double pc = 0x1.ae7f3e733b81fp-45;
    pc = fma(z, pc, -0x1.93974a8c07c9dp-37);
    pc = fma(z, pc, 0x1.1eed8eff8d898p-29);
    pc = fma(z, pc, -0x1.27e4fb7789f5cp-22);
    pc = fma(z, pc, 0x1.a01a01a01a01ap-16);
    pc = fma(z, pc, -0x1.6c16c16c16c17p-10);
    pc = fma(z, pc, 0x1.5555555555555p-5);
    pc = fma(z, pc, -0.5);
    const double cr = fma(z, pc, 1.0);
    const int q = (int)k & 3;
    const double a = (q & 1) ? cr : sr;
    const double b = (q & 1) ? sr : cr;
    *s = (q & 2) ? -a : a;
    *c = (q == 1 || q == 2) ? -b : b;
}

void orc_sincos64(const double* x, double* s, double* c, int64_t n) {
    for (int64_t k = 0; k < n; ++k) sincos64(x[k], &s[k], &c[k]);
}

void orc_sqrt_cr(const double* x, double* y, int64_t n) {
    for (int64_t k = 0; k < n; ++k) y[k] = sqrt_cr(x[k]);
}

/* B = b0 + J r for one record (b0[3], J row major) */
static inline d3 field_at(const double* rec, double rx, double ry, double rz) {
    d3 B;
    B.x = fma(rec[5], rz, fma(rec[4], ry, fma(rec[3], rx, rec[0])));
    B.y = fma(rec[8], rz, fma(rec[7], ry, fma(rec[6], rx, rec[1])));
    B.z = fma(rec[11], rz, fma(rec[10], ry, fma(rec[9], rx, rec[2])));
    return B;
}

/* rotation about B by -gh |B| (updateMag_noT2_rotate, blochdiff.cuh:186-199) */
static inline void mrotate(d3* M, d3 B, double gh) {
    const double b2 = fma(B.z, B.z, fma(B.y, B.y, B.x * B.x));
    if (!(b2 > 0.0)) return;
    const double bn = sqrt_cr(b2);
    const double ib = 1.0 / bn;
    double s, c;
    sincos64(-(gh * bn), &s, &c);
    const double f = (fma(B.z, M->z, fma(B.y, M->y, B.x * M->x)) * ib) * ib;
    const double sb = s * ib;
    const double cx = f * B.x, cy = f * B.y, cz = f * B.z;
    const double wx = fma(B.y, M->z, -(B.z * M->y));
    const double wy = fma(B.z, M->x, -(B.x * M->z));
    const double wz = fma(B.x, M->y, -(B.y * M->x));
    M->x = fma(wx, sb, fma(M->x - cx, c, cx));
    M->y = fma(wy, sb, fma(M->y - cy, c, cy));
    M->z = fma(wz, sb, fma(M->z - cz, c, cz));
}

/* h (gamma M x B - relaxation) (bloch, blochdiff.cuh:7-22) */
static inline d3 bloch_h(d3 M, d3 B, double gam, double h, int relax, double i1, double i2) {
    d3 d;
    d.x = gam * fma(M.y, B.z, -(M.z * B.y));
    d.y = gam * fma(M.z, B.x, -(M.x * B.z));
    d.z = gam * fma(M.x, B.y, -(M.y * B.x));
    if (relax) {
        d.x = fma(-M.x, i2, d.x);
        d.y = fma(-M.y, i2, d.y);
        d.z = fma(-(M.z - 1.0), i1, d.z);
    }
    d.x = d.x * h; d.y = d.y * h; d.z = d.z * h;
    return d;
}

static inline d3 axpy(double a, d3 x, d3 y) {
    d3 r = {fma(a, x.x, y.x), fma(a, x.y, y.y), fma(a, x.z, y.z)};
    return r;
}

/* classical RK4 (updateMagRK4, blochdiff.cuh:24-67) */
static inline void mrk4(d3* M, d3 B0, d3 Bh, d3 B1, double gam, double h, int relax, double i1, double i2) {
    d3 k1 = bloch_h(*M, B0, gam, h, relax, i1, i2);
    d3 k2 = bloch_h(axpy(0.5, k1, *M), Bh, gam, h, relax, i1, i2);
    d3 k3 = bloch_h(axpy(0.5, k2, *M), Bh, gam, h, relax, i1, i2);
    d3 k4 = bloch_h(axpy(1.0, k3, *M), B1, gam, h, relax, i1, i2);
    d3 t = axpy(2.0, k3, axpy(2.0, k2, k1));
    d3 u = {t.x + k4.x, t.y + k4.y, t.z + k4.z};
    *M = axpy(1.0 / 6.0, u, *M);
}

/* Dormand-Prince fifth-order step (updateMagDORP, blochdiff.cuh:68-118),
 * with the published a63 = +46732/5247 (the reference has -46732/5247) */
static inline void mdopri5(d3* M, const d3* B, double gam, double h, int relax, double i1, double i2) {
    d3 k1 = bloch_h(*M, B[0], gam, h, relax, i1, i2);
    d3 k2 = bloch_h(axpy(1.0 / 5.0, k1, *M), B[1], gam, h, relax, i1, i2);
    d3 k3 = bloch_h(axpy(9.0 / 40.0, k2, axpy(3.0 / 40.0, k1, *M)), B[2], gam, h, relax, i1, i2);
    d3 k4 = bloch_h(axpy(32.0 / 9.0, k3, axpy(-56.0 / 15.0, k2, axpy(44.0 / 45.0, k1, *M))), B[3], gam, h,
                    relax, i1, i2);
    d3 k5 = bloch_h(axpy(-212.0 / 729.0, k4,
                         axpy(64448.0 / 6561.0, k3, axpy(-25360.0 / 2187.0, k2, axpy(19372.0 / 6561.0, k1, *M)))),
                    B[4], gam, h, relax, i1, i2);
    d3 k6 = bloch_h(axpy(-5103.0 / 18656.0, k5,
                         axpy(49.0 / 176.0, k4,
                              axpy(46732.0 / 5247.0, k3, axpy(-355.0 / 33.0, k2, axpy(9017.0 / 3168.0, k1, *M))))),
                    B[5], gam, h, relax, i1, i2);
    *M = axpy(11.0 / 84.0, k6,
              axpy(-2187.0 / 6784.0, k5,
                   axpy(125.0 / 192.0, k4, axpy(500.0 / 1113.0, k3, axpy(35.0 / 384.0, k1, *M)))));
}

/* samples per step for s substeps */
static inline int mag_samples(int integ, int s) {
    return integ == AXX_MAG_ROTATE ? s : (integ == AXX_MAG_RK4 ? 2 * s + 1 : 6 * s);
}

/* What a walk of one group of fields needs: fields g0 .. g0+F-1 of f. */
typedef struct {
    const axx_field* f;
    int g0, F;
    const int* off;       /* [n_steps + 1] first sample of each step */
    const double* iT1;    /* per compartment */
    const double* iT2;
    int relax;
} magjob_t;

/* M of every field after step i at unwrapped position r0 + d (kernelMag.cuh:93-136) */
static void mag_update(const magjob_t* G, d3* M, int i, float x0, float y0, float z0, float dx, float dy,
                       float dz, int comp) {
    const axx_field* f = G->f;
    const double rx = ((double)x0 + (double)dx) - f->origin[0];
    const double ry = ((double)y0 + (double)dy) - f->origin[1];
    const double rz = ((double)z0 + (double)dz) - f->origin[2];
    const int s0 = G->off[i];
    const int ns = f->substeps ? f->substeps[i] : 1;
    const double h = f->dt / (double)ns;
    double i1 = 0.0, i2 = 0.0;
    if (f->integrator != AXX_MAG_ROTATE && G->relax) { i1 = G->iT1[comp]; i2 = G->iT2[comp]; }
    const int NF = f->n_fields;
#define REC(k, j) (f->samples + ((size_t)(k) * NF + G->g0 + (j)) * 12)
    for (int j = 0; j < G->F; ++j) {
        d3 m = M[j];
        for (int k = 0; k < ns; ++k) {
            if (f->integrator == AXX_MAG_ROTATE) {
                mrotate(&m, field_at(REC(s0 + k, j), rx, ry, rz), f->gamma * h);
            } else if (f->integrator == AXX_MAG_RK4) {
                int q = s0 + 2 * k;
                mrk4(&m, field_at(REC(q, j), rx, ry, rz), field_at(REC(q + 1, j), rx, ry, rz),
                     field_at(REC(q + 2, j), rx, ry, rz), f->gamma, h, G->relax, i1, i2);
            } else {
                d3 B[6];
                for (int c = 0; c < 6; ++c) B[c] = field_at(REC(s0 + 6 * k + c, j), rx, ry, rz);
                mdopri5(&m, B, f->gamma, h, G->relax, i1, i2);
            }
        }
        M[j] = m;
    }
#undef REC
}

/* ------------------------------------------------------------------ */
/* Prepared substrate: primitives with their periodic images (what the
 * reference's two-image test plus correctEdges copies cover,
 * lattice.cuh:205-247 and RPSinitializer.h:131-186), per-compartment step
 * lengths and T2 rates, per-primitive transmission thresholds. */
typedef struct {
    float cx, cy, cz, R2;
    float R;     /* radius (float), R2 = R * R */
    int type;    /* AXX_PRIM_* */
    int comp;    /* compartment id = original index + 1 */
    int orig;
    float nx, ny, nz, d;   /* planes: unit normal and offset n . point */
} oprim;

typedef struct {
    float a, b, c;   /* cell the positions wrap in (unit cell of a repeating substrate) */
    float ca, cb;    /* the substrate's cell along x and y (placement) */
    int units;       /* unit cells in the substrate's cell */
    int n_img;
    oprim* img;
    int n_comp;
    float* L;      /* [n_comp] step length sqrt(6 D dt) */
    float* rate;   /* [n_comp] dt / T2 (0 if no decay)   */
    float* thr_out;/* [n_prims] hazard of p = 4 kappa / v_inside  */
    float* thr_in; /* [n_prims] hazard of p = 4 kappa / v_lattice */
    float Lmax;
    float H2c, H2s; /* squared half of the smallest cell side (xy; xyz) */
    float plane_eps;/* minimum distance from a plane after reflecting from it */
    int gmask;      /* steps per lookup group - 1: positions wrap after step i when i & gmask == 0 */
    int any_3d;     /* a sphere or plane is present (z matters) */
    /* optional cell bins (binned variant of the scan) */
    int nb[3];
    float binw_inv[3];
    int* bin_off;  /* [nbins+1] */
    int* bin_idx;
} oprep;

/* Projection of a box onto a unit normal: [min, max] of n . corner. */
static void box_proj(const double n[3], const double lo[3], const double hi[3], double* pmin, double* pmax) {
    *pmin = *pmax = 0.0;
    for (int a = 0; a < 3; ++a) {
        double u = n[a] * lo[a], v = n[a] * hi[a];
        *pmin += u < v ? u : v;
        *pmax += u < v ? v : u;
    }
}

static void unit_normal(const axx_substrate* s, int k, double n[3]) {
    double l = sqrt(s->normal[3 * k] * s->normal[3 * k] + s->normal[3 * k + 1] * s->normal[3 * k + 1] +
                    s->normal[3 * k + 2] * s->normal[3 * k + 2]);
    for (int a = 0; a < 3; ++a) n[a] = s->normal[3 * k + a] / l;
}

static int prim_reach_hits_box(const axx_substrate* s, int k, const double ck[3], double sx, double sy,
                               double sz, double lo[3], double hi[3], double reach) {
    /* Does primitive k at centre ck shifted by (sx,sy,sz), grown by `reach`,
     * touch the box? */
    double c[3] = {ck[0] + sx, ck[1] + sy, ck[2] + sz};
    if (s->type[k] == AXX_PRIM_PLANE) {
        double n[3], pmin, pmax;
        unit_normal(s, k, n);
        double d = n[0] * c[0] + n[1] * c[1] + n[2] * c[2];
        box_proj(n, lo, hi, &pmin, &pmax);
        return d >= pmin - reach && d <= pmax + reach;
    }
    double R = s->radius[k] + reach;
    int dims = s->type[k] == AXX_PRIM_SPHERE ? 3 : 2;
    double d2 = 0.0;
    for (int a = 0; a < dims; ++a) {
        double d = 0.0;
        if (c[a] < lo[a]) d = lo[a] - c[a];
        else if (c[a] > hi[a]) d = c[a] - hi[a];
        d2 += d * d;
    }
    return d2 <= R * R;
}

/* Boundary reachable from the box: the circle/sphere surface comes within
 * `reach` of some point of the box. */
static int prim_surface_near_box(const oprim* p, double lo[3], double hi[3], double reach) {
    if (p->type == AXX_PRIM_PLANE) {
        double n[3] = {p->nx, p->ny, p->nz}, pmin, pmax;
        box_proj(n, lo, hi, &pmin, &pmax);
        return (double)p->d >= pmin - reach && (double)p->d <= pmax + reach;
    }
    double c[3] = {p->cx, p->cy, p->cz};
    double R = sqrt((double)p->R2);
    int dims = p->type == AXX_PRIM_SPHERE ? 3 : 2;
    double dmin2 = 0.0, dmax2 = 0.0;
    for (int a = 0; a < dims; ++a) {
        double d = 0.0;
        if (c[a] < lo[a]) d = lo[a] - c[a];
        else if (c[a] > hi[a]) d = c[a] - hi[a];
        dmin2 += d * d;
        double e = fmax(fabs(c[a] - lo[a]), fabs(c[a] - hi[a]));
        dmax2 += e * e;
    }
    double dmin = sqrt(dmin2), dmax = sqrt(dmax2);
    return dmin <= R + reach && dmax >= R - reach;
}

static void oprep_free(oprep* P) {
    free(P->img); free(P->L); free(P->rate); free(P->thr_out); free(P->thr_in);
    free(P->bin_off); free(P->bin_idx);
    memset(P, 0, sizeof(*P));
}

int orc_validate(const axx_substrate* s, const axx_acq* q) {
    if (!s || !q) return AXX_ERR_INVALID;
    if (!(s->a > 0 && s->b > 0 && s->c > 0)) return AXX_ERR_INVALID;
    if (s->n_prims < 0 || s->n_prims > AXX_MAX_PRIMS) return AXX_ERR_LIMIT;
    if (!(q->dt > 0) || q->n_steps < 1) return AXX_ERR_INVALID;
    if (q->n_profiles < 1 || q->n_profiles > AXX_MAX_PROFILES) return AXX_ERR_LIMIT;
    if (!(s->D_lat >= 0)) return AXX_ERR_INVALID;
    double Lmax = sqrt(6.0 * s->D_lat * q->dt);
    for (int k = 0; k < s->n_prims; ++k) {
        if (s->type[k] == AXX_PRIM_PLANE) {
            if (!s->normal) return AXX_ERR_INVALID;
            double l2 = s->normal[3 * k] * s->normal[3 * k] + s->normal[3 * k + 1] * s->normal[3 * k + 1] +
                        s->normal[3 * k + 2] * s->normal[3 * k + 2];
            if (!(l2 > 0)) return AXX_ERR_INVALID;
            continue;
        }
        if (s->type[k] != AXX_PRIM_CYLINDER_XY && s->type[k] != AXX_PRIM_SPHERE) return AXX_ERR_INVALID;
        if (!(s->radius[k] > 0) || !(s->D[k] >= 0)) return AXX_ERR_INVALID;
        double R = s->radius[k];
        if (2 * R >= s->a || 2 * R >= s->b || (s->type[k] == AXX_PRIM_SPHERE && 2 * R >= s->c))
            return AXX_ERR_GEOMETRY;
        double L = sqrt(6.0 * s->D[k] * q->dt);
        if (L > Lmax) Lmax = L;
    }
    if (2 * Lmax >= s->a || 2 * Lmax >= s->b || 2 * Lmax >= s->c) return AXX_ERR_GEOMETRY;
    /* non-overlap under the minimum image convention (fixes D8,
     * RPSinitializer.h:191-219) */
    for (int i = 0; i < s->n_prims; ++i)
        for (int j = i + 1; j < s->n_prims; ++j) {
            if (s->type[i] == AXX_PRIM_PLANE || s->type[j] == AXX_PRIM_PLANE) continue;
            double d[3];
            int dims = (s->type[i] == AXX_PRIM_SPHERE && s->type[j] == AXX_PRIM_SPHERE) ? 3 : 2;
            double L3[3] = {s->a, s->b, s->c};
            double d2 = 0;
            for (int a = 0; a < dims; ++a) {
                d[a] = s->center[3 * i + a] - s->center[3 * j + a];
                d[a] -= L3[a] * nearbyint(d[a] / L3[a]);
                d2 += d[a] * d[a];
            }
            double rr = s->radius[i] + s->radius[j];
            if (d2 < rr * rr) return AXX_ERR_GEOMETRY;
        }
    return AXX_OK;
}

/* Longest step in units of sqrt(6 D dt): Gaussian steps are at most
 * |n| = sqrt(2 * 33.27) = 8.157 standard normals long (24-bit uniforms),
 * 4.71 such units. */
static double step_reach_factor(int step_mode) { return step_mode == AXX_STEP_GAUSSIAN ? 4.75 : 1.0; }

/* Steps between two wraps of the position into the cell (the product's
 * lookup group, capi.hip lookup_group, same rule and arithmetic): when steps
 * have a fixed length and every surface is a cylinder, eight if every
 * cylinder is at least sixteen step lengths from every other (periodic
 * images of the whole cell included), four if at least eight; two
 * otherwise.  The reference wraps after every step
 * (src/kernels/kernelLattice.cuh:96-99); where a position wraps changes no
 * geometry, only the fp32 rounding of the coordinates. */
static int lookup_group(const axx_substrate* s, double Lmax, double reach_factor) {
    if (reach_factor != 1.0) return 2;
    int n = s->n_prims;
    for (int k = 0; k < n; ++k)
        if (s->type[k] != AXX_PRIM_CYLINDER_XY) return 2;
    double lim4 = 8.0 * Lmax, lim8 = 16.0 * Lmax;
    int g = 8;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j)
            for (int sy = -1; sy <= 1; ++sy)
                for (int sx = -1; sx <= 1; ++sx) {
                    if (i == j && sx == 0 && sy == 0) continue;
                    double dx = s->center[3 * j] + sx * s->a - s->center[3 * i];
                    double dy = s->center[3 * j + 1] + sy * s->b - s->center[3 * i + 1];
                    double d2 = dx * dx + dy * dy;
                    double r4 = lim4 + s->radius[i] + s->radius[j];
                    double r8 = lim8 + s->radius[i] + s->radius[j];
                    if (d2 < r4 * r4) return 2;
                    if (d2 < r8 * r8) g = 4;
                }
    return g;
}

/* Walk a repeating substrate in its unit cell unless this is cleared (tests
 * and the reference-faithful CPU timing walk the whole cell). */
static int g_unit_cells = 1;
void orc_set_unit_cells(int on) { g_unit_cells = on; }

/* The primitives the walk sees and the cell they repeat in (restates the
 * product's host preparation): a substrate whose primitives map onto
 * themselves under a shift of a/mx along x and b/my along y is walked in its
 * unit cell (a/mx) x (b/my) x c; each translation class is represented by its
 * lowest-numbered member (compartment id = its index + 1), at its centre
 * reduced into the unit cell.  Otherwise the whole cell. */
typedef struct {
    double a, b, c;
    int mx, my, n;
    int* orig;        /* [n] substrate index */
    double* center;   /* [3 n] */
} oview;

static double pdist(double d, double L) { return fabs(d - L * nearbyint(d / L)); }

static int same_kind(const axx_substrate* s, int i, int j) {
    double ki = s->kappa ? s->kappa[i] : 0.0, kj = s->kappa ? s->kappa[j] : 0.0;
    return s->type[i] == s->type[j] && s->radius[i] == s->radius[j] && s->D[i] == s->D[j] &&
           (s->T2[i] == s->T2[j] || (isinf(s->T2[i]) && isinf(s->T2[j]))) && ki == kj;
}

static int shift_is_symmetry(const axx_substrate* s, int ax, double sh, double tol) {
    double L[3] = {s->a, s->b, s->c};
    for (int k = 0; k < s->n_prims; ++k) {
        int found = 0;
        for (int j = 0; j < s->n_prims && !found; ++j) {
            if (!same_kind(s, k, j)) continue;
            int ok = 1;
            for (int a = 0; a < 3 && ok; ++a)
                ok = pdist(s->center[3 * k + a] + (a == ax ? sh : 0.0) - s->center[3 * j + a], L[a]) <= tol;
            found = ok;
        }
        if (!found) return 0;
    }
    return 1;
}

static int periods(const axx_substrate* s, int ax, double tol) {
    double L = ax == 0 ? s->a : s->b;
    for (int m = s->n_prims; m >= 2; --m)
        if (s->n_prims % m == 0 && shift_is_symmetry(s, ax, L / m, tol)) return m;
    return 1;
}

static void oview_whole(const axx_substrate* s, oview* V) {
    V->a = s->a; V->b = s->b; V->c = s->c;
    V->mx = V->my = 1;
    V->n = s->n_prims;
    V->orig = malloc(sizeof(int) * (size_t)(s->n_prims + 1));
    V->center = malloc(sizeof(double) * 3 * (size_t)(s->n_prims + 1));
    for (int k = 0; k < s->n_prims; ++k) {
        V->orig[k] = k;
        for (int a = 0; a < 3; ++a) V->center[3 * k + a] = s->center[3 * k + a];
    }
}

static void oview_build(const axx_substrate* s, int allow, double reach, oview* V) {
    oview_whole(s, V);
    int n = s->n_prims;
    if (!allow || n < 2) return;
    for (int k = 0; k < n; ++k)
        if (s->type[k] == AXX_PRIM_PLANE) return;
    double tol = 1e-9 * fmax(fmax(s->a, s->b), s->c);
    int mx = periods(s, 0, tol), my = periods(s, 1, tol);
    if (mx * my == 1) return;
    double W = s->a / mx, H = s->b / my;
    if (!(2.0 * reach < fmin(W, H))) return;
    int* reps = malloc(sizeof(int) * (size_t)n);
    double* rc = malloc(sizeof(double) * 3 * (size_t)n);
    int nr = 0;
    for (int k = 0; k < n; ++k) {
        if (2.0 * s->radius[k] >= fmin(W, H)) { free(reps); free(rc); return; }
        double rx = s->center[3 * k] - W * floor(s->center[3 * k] / W);
        double ry = s->center[3 * k + 1] - H * floor(s->center[3 * k + 1] / H);
        if (W - rx <= tol) rx -= W;
        if (H - ry <= tol) ry -= H;
        double rz = s->center[3 * k + 2];
        int member = 0;
        for (int r = 0; r < nr && !member; ++r)
            member = same_kind(s, k, reps[r]) && fabs(rx - rc[3 * r]) <= tol && fabs(ry - rc[3 * r + 1]) <= tol &&
                     pdist(rz - rc[3 * r + 2], s->c) <= tol;
        if (member) continue;
        reps[nr] = k;
        rc[3 * nr] = rx; rc[3 * nr + 1] = ry; rc[3 * nr + 2] = rz;
        ++nr;
    }
    if ((long long)nr * mx * my != n) { free(reps); free(rc); return; }
    free(V->orig); free(V->center);
    V->a = W; V->b = H;
    V->mx = mx; V->my = my;
    V->n = nr;
    V->orig = reps;
    V->center = rc;
}

static int oprep_build_view(oprep* P, const axx_substrate* s, const axx_acq* q, double bin_width,
                            double reach_factor, int allow_unit_cell);

static int oprep_build(oprep* P, const axx_substrate* s, const axx_acq* q, double bin_width,
                       double reach_factor) {
    return oprep_build_view(P, s, q, bin_width, reach_factor, 0);
}

static int oprep_build_view(oprep* P, const axx_substrate* s, const axx_acq* q, double bin_width,
                            double reach_factor, int allow_unit_cell) {
    memset(P, 0, sizeof(*P));
    int rc = orc_validate(s, q);
    if (rc) return rc;
    P->plane_eps = (float)(64.0 * ldexp(1.0, -23) * fmax(fmax(s->a, s->b), s->c));
    int n = s->n_prims;
    P->n_comp = n + 1;
    P->L = malloc(sizeof(float) * (n + 1));
    P->rate = malloc(sizeof(float) * (n + 1));
    P->thr_out = malloc(sizeof(float) * (n + 1));
    P->thr_in = malloc(sizeof(float) * (n + 1));
    double dt = q->dt;
    double Lmax = sqrt(6.0 * s->D_lat * dt);
    P->L[0] = (float)sqrt(6.0 * s->D_lat * dt);
    P->rate[0] = (s->T2_lat > 0 && isfinite(s->T2_lat)) ? (float)(dt / s->T2_lat) : 0.0f;
    double v_lat = sqrt(6.0 * s->D_lat / dt);
    int any_plane = 0;
    for (int k = 0; k < n; ++k) {
        if (s->type[k] == AXX_PRIM_PLANE) {
            /* planes bound no compartment: nothing diffuses "in" them */
            P->L[k + 1] = 0.0f;
            P->rate[k + 1] = 0.0f;
            P->thr_out[k] = P->thr_in[k] = 0.0f;
            any_plane = 1;
            continue;
        }
        double L = sqrt(6.0 * s->D[k] * dt);
        if (L > Lmax) Lmax = L;
        P->L[k + 1] = (float)L;
        P->rate[k + 1] = (s->T2[k] > 0 && isfinite(s->T2[k])) ? (float)(dt / s->T2[k]) : 0.0f;
        double kap = s->kappa ? s->kappa[k] : 0.0;
        double v_in = sqrt(6.0 * s->D[k] / dt);
        P->thr_out[k] = kap > 0 ? hazard(4.0 * kap / v_in) : 0.0f;
        P->thr_in[k] = kap > 0 ? hazard(4.0 * kap / v_lat) : 0.0f;
    }
    P->Lmax = (float)Lmax;
    for (int k = 0; k < n; ++k) P->any_3d |= s->type[k] != AXX_PRIM_CYLINDER_XY;
    /* periodic images: every shift in {-1,0,1}^2 (x{-1,0,1} for spheres)
     * whose grown shape touches the cell, in (original, shift) order */
    /* positions wrap after every G-th step (G = lookup_group), so a step may
     * start up to G - 1 step lengths outside the cell: images within G step
     * lengths; a plane counts as hit within plane_eps of it, so its reach is
     * longer */
    int G = lookup_group(s, Lmax, reach_factor);
    P->gmask = G - 1;
    double reach = G * reach_factor * Lmax * 1.001 + 1e-6 * fmax(s->a, s->b) +
                   (any_plane ? 4.0 * P->plane_eps : 0.0);
    oview V;
    oview_build(s, allow_unit_cell, reach, &V);
    P->a = (float)V.a; P->b = (float)V.b; P->c = (float)V.c;
    P->ca = (float)s->a; P->cb = (float)s->b;
    P->units = V.mx * V.my;
    {
        double hc = 0.5 * fmin(V.a, V.b), hs = 0.5 * fmin(fmin(V.a, V.b), V.c);
        P->H2c = (float)(hc * hc);
        P->H2s = (float)(hs * hs);
    }
    double lo[3] = {0, 0, 0}, hi[3] = {V.a, V.b, V.c};
    int cap = V.n * 27 + 1;
    P->img = malloc(sizeof(oprim) * cap);
    int m = 0;
    for (int kk = 0; kk < V.n; ++kk) {
        int k = V.orig[kk];
        const double* ck = V.center + 3 * kk;
        int zr = s->type[k] != AXX_PRIM_CYLINDER_XY ? 1 : 0;
        int first = m;
        for (int kz = -zr; kz <= zr; ++kz)
            for (int ky = -1; ky <= 1; ++ky)
                for (int kx = -1; kx <= 1; ++kx) {
                    double sx = kx * V.a, sy = ky * V.b, sz = kz * V.c;
                    if (!prim_reach_hits_box(s, k, ck, sx, sy, sz, lo, hi, reach)) continue;
                    if (s->type[k] == AXX_PRIM_PLANE) {
                        /* shifts along the plane give the same plane */
                        double nn[3];
                        unit_normal(s, k, nn);
                        float dd = (float)(nn[0] * (ck[0] + sx) + nn[1] * (ck[1] + sy) + nn[2] * (ck[2] + sz));
                        int dup = 0;
                        for (int q = first; q < m; ++q) dup |= P->img[q].d == dd;
                        if (dup) continue;
                        oprim* p = &P->img[m++];
                        memset(p, 0, sizeof(*p));
                        p->type = AXX_PRIM_PLANE;
                        p->comp = k + 1;
                        p->orig = k;
                        p->nx = (float)nn[0]; p->ny = (float)nn[1]; p->nz = (float)nn[2];
                        p->d = dd;
                        continue;
                    }
                    oprim* p = &P->img[m++];
                    p->cx = (float)(ck[0] + sx);
                    p->cy = (float)(ck[1] + sy);
                    p->cz = (float)(ck[2] + sz);
                    float Rf = (float)s->radius[k];
                    p->R2 = Rf * Rf;
                    p->R = Rf;
                    p->type = s->type[k];
                    p->comp = k + 1;
                    p->orig = k;
                    p->nx = p->ny = p->nz = p->d = 0.0f;
                }
    }
    P->n_img = m;
    free(V.orig); free(V.center);
    /* cell bins (binned variant): surface within reach of the bin */
    if (bin_width > 0) {
        int any_sphere = 0;
        for (int k = 0; k < n; ++k) any_sphere |= (s->type[k] != AXX_PRIM_CYLINDER_XY);
        double Ls[3] = {V.a, V.b, V.c};
        for (int a = 0; a < 3; ++a) {
            int nb = (int)floor(Ls[a] / bin_width);
            if (nb < 1) nb = 1;
            if (nb > 1024) nb = 1024;
            if (a == 2 && !any_sphere) nb = 1;
            P->nb[a] = nb;
            P->binw_inv[a] = (float)(nb / Ls[a]);
        }
        int nbins = P->nb[0] * P->nb[1] * P->nb[2];
        P->bin_off = calloc((size_t)nbins + 1, sizeof(int));
        /* two passes over (image, touched bin): count, then fill in image order */
        int* fillp = calloc((size_t)nbins + 1, sizeof(int));
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = 0; i < m; ++i) {
                const oprim* pi = &P->img[i];
                double R = sqrt((double)pi->R2) + reach;
                double c3[3] = {pi->cx, pi->cy, pi->cz};
                int lo_b[3], hi_b[3];
                for (int a = 0; a < 3; ++a) {
                    if ((a == 2 && pi->type == AXX_PRIM_CYLINDER_XY) || pi->type == AXX_PRIM_PLANE) {
                        lo_b[a] = 0;
                        hi_b[a] = P->nb[a] - 1;
                        continue;
                    }
                    double bw = Ls[a] / P->nb[a];
                    lo_b[a] = (int)floor((c3[a] - R) / bw) - 1;
                    hi_b[a] = (int)floor((c3[a] + R) / bw) + 1;
                    if (lo_b[a] < 0) lo_b[a] = 0;
                    if (hi_b[a] > P->nb[a] - 1) hi_b[a] = P->nb[a] - 1;
                }
                for (int bz = lo_b[2]; bz <= hi_b[2]; ++bz)
                    for (int by = lo_b[1]; by <= hi_b[1]; ++by)
                        for (int bx = lo_b[0]; bx <= hi_b[0]; ++bx) {
                            double blo[3] = {bx * Ls[0] / P->nb[0], by * Ls[1] / P->nb[1], bz * Ls[2] / P->nb[2]};
                            double bhi[3] = {(bx + 1) * Ls[0] / P->nb[0], (by + 1) * Ls[1] / P->nb[1],
                                             (bz + 1) * Ls[2] / P->nb[2]};
                            if (!prim_surface_near_box(pi, blo, bhi, reach)) continue;
                            int b = (bz * P->nb[1] + by) * P->nb[0] + bx;
                            if (pass == 0) P->bin_off[b + 1]++;
                            else P->bin_idx[fillp[b]++] = i;
                        }
            }
            if (pass == 0) {
                for (int b = 0; b < nbins; ++b) P->bin_off[b + 1] += P->bin_off[b];
                P->bin_idx = malloc(sizeof(int) * ((size_t)P->bin_off[nbins] + 1));
                memcpy(fillp, P->bin_off, sizeof(int) * (size_t)nbins);
            }
        }
        free(fillp);
    }
    return AXX_OK;
}

/* ------------------------------------------------------------------ */
/* Geometry predicates (cylinderXY.cuh:69-80, Sphere.cuh:89-94). */
static inline int inside_at(const oprim* p, float x, float y, float z) {
    if (p->type == AXX_PRIM_PLANE) return 0;
    float ox = x - p->cx, oy = y - p->cy;
    if (p->type == AXX_PRIM_SPHERE) {
        float oz = z - p->cz;
        return fmaf(oz, oz, fmaf(oy, oy, ox * ox)) < p->R2;
    }
    return fmaf(oy, oy, ox * ox) < p->R2;
}

/* Region of a point (lattice.cuh:134-148): compartment of the primitive that
 * contains it, else the lattice (0).  Non-overlap makes the match unique. */
static int region_of(const oprep* P, float x, float y, float z) {
    for (int i = 0; i < P->n_img; ++i)
        if (inside_at(&P->img[i], x, y, z)) return P->img[i].comp;
    return 0;
}

/* Crossing test of the segment s + t*u, t in [0, ell], against one surface
 * (cylinderXY.cuh:82-122, Sphere.cuh:102-139 restated).  The walker's side
 * of the surface is taken from its tracked compartment, and the start point
 * is clamped onto that side (C' = min(C,0) inside, max(C,0) outside), which
 * replaces the reference's epsilon test (real_equal(root*|dr|, 0, eps)). */
/* Plane crossing (plane.cuh:18-28 restated): the walker's side is the side
 * of the step's start point (x0), and the segment start is clamped onto it.
 * The plane counts as hit when the segment comes closer than eps to it (not
 * only when it crosses), so that, with the push applied after a plane
 * reflection, every step ends at least about eps away on the walker's side and
 * the side taken from the next start point is the tracked one.  The reported
 * time is that of reaching the plane, at most ell. */
static inline int plane_test(const oprim* p, float x0, float y0, float z0, float sx, float sy,
                             float sz, float ux, float uy, float uz, float ell, float eps,
                             float* t_out) {
    float un = fmaf(p->nz, uz, fmaf(p->ny, uy, p->nx * ux));
    float sd = fmaf(p->nz, sz, fmaf(p->ny, sy, p->nx * sx)) - p->d;
    float s0 = fmaf(p->nz, z0, fmaf(p->ny, y0, p->nx * x0)) - p->d;
    float sg = s0 >= 0.0f ? 1.0f : -1.0f;
    float dd = sg * sd;
    dd = dd > 0.0f ? dd : 0.0f;     /* clamped distance of the segment start */
    float v = -(sg * un);           /* approach speed towards the plane */
    if (!(v > 0.0f)) return 0;
    if (!(fmaf(-ell, v, dd) < eps)) return 0;
    float t = dd / v;
    *t_out = t < ell ? t : ell;
    return 1;
}

static inline int hit_test(const oprim* p, int same_comp, float H2c, float H2s, float sx,
                           float sy, float sz, float ux, float uy, float uz, float Axy, float A3,
                           float ell, float* t_out, int* ins_out) {
    float ox = sx - p->cx, oy = sy - p->cy;
    float A, B, q, H2;
    if (p->type == AXX_PRIM_SPHERE) {
        float oz = sz - p->cz;
        A = A3;
        B = fmaf(oz, uz, fmaf(oy, uy, ox * ux));
        q = fmaf(oz, oz, fmaf(oy, oy, ox * ox));
        H2 = H2s;
    } else {
        A = Axy;
        B = fmaf(oy, uy, ox * ux);
        q = fmaf(oy, oy, ox * ox);
        H2 = H2c;
    }
    float C = q - p->R2;
    /* the walker is inside this image iff it is in the image's compartment
     * and within half a cell of its centre (other images of the same
     * primitive are at least half a cell away) */
    int inside = same_comp && q < H2;
    *ins_out = inside;
    if (!(A > 0.0f)) return 0;
    if (inside) {
        float Cp = C < 0.0f ? C : 0.0f;
        float fe = fmaf(fmaf(A, ell, B + B), ell, Cp);
        if (!(fe > 0.0f)) return 0;
        float disc = fmaf(B, B, -(A * Cp));
        float sq = sqrtf(disc);
        float t;
        if (B < 0.0f) {
            t = (sq - B) / A;
        } else {
            float den = B + sq;
            t = den > 0.0f ? (-Cp) / den : 0.0f;
        }
        *t_out = t < ell ? t : ell;
        return 1;
    }
    if (!(B < 0.0f)) return 0;
    float Cp = C > 0.0f ? C : 0.0f;
    if (Cp == 0.0f) {
        *t_out = 0.0f;
        return 1;
    }
    float fe = fmaf(fmaf(A, ell, B + B), ell, Cp);
    float disc = fmaf(B, B, -(A * Cp));
    if (!(fe < 0.0f) && !((-B) < A * ell && disc > 0.0f)) return 0;
    float sq = sqrtf(disc > 0.0f ? disc : 0.0f);
    float t = Cp / (sq - B);
    *t_out = t < ell ? t : ell;
    return 1;
}

/* lmod (lattice.cuh:66-88): strict comparisons, one cell at most. */
static inline float wrap1(float v, float a) {
    if (v > a) v = v - a;
    if (v < 0.0f) v = v + a;
    return v;
}

/* A placement coordinate of the substrate's cell taken into the unit cell of
 * width w: v - floor(v / w) w in one rounding, then wrapped as a position. */
static inline float unit_wrap(float v, float w) {
    float k = floorf(v / w);
    return wrap1(fmaf(-k, w, v), w);
}

static inline void cand_range(const oprep* P, int use_bins, float x, float y, float z,
                              const int** list, int* cnt, int* dense) {
    if (!use_bins) {
        *list = NULL;
        *cnt = P->n_img;
        *dense = 1;
        return;
    }
    int bx = (int)(x * P->binw_inv[0]);
    int by = (int)(y * P->binw_inv[1]);
    int bz = (int)(z * P->binw_inv[2]);
    if (bx > P->nb[0] - 1) bx = P->nb[0] - 1;
    if (by > P->nb[1] - 1) by = P->nb[1] - 1;
    if (bz > P->nb[2] - 1) bz = P->nb[2] - 1;
    if (bx < 0) bx = 0;
    if (by < 0) by = 0;
    if (bz < 0) bz = 0;
    int b = (bz * P->nb[1] + by) * P->nb[0] + bx;
    *list = P->bin_idx + P->bin_off[b];
    *cnt = P->bin_off[b + 1] - P->bin_off[b];
    *dense = 0;
}

/* ------------------------------------------------------------------ */
/* One walker over steps [i0, i1) (kernelLattice.cuh:46-99). */
typedef struct {
    const oprep* P;
    const axx_acq* q;
    const axx_run* run;
    uint32_t* st;     /* SoA state */
    int64_t N;
    int Pn;           /* profiles */
    int axes;         /* bit mask of axes carrying phase */
    int i0, i1;
    int use_bins;
    int64_t w_begin, w_end;
    const magjob_t* mag;   /* magnetization of a group of fields instead of phase moments */
    int err;
    /* boundary events at circles and spheres: [0] from inside the surface,
     * [1] of those transmitted, [2] from outside, [3] of those transmitted */
    int64_t ev[4];
} job_t;

#define FS(f) (((float*)J->st) + (size_t)(f) * J->N)
#define IS(f) (((int32_t*)J->st) + (size_t)(f) * J->N)

static int init_walker(const job_t* J, int64_t w, uint32_t wid, const uint32_t key[2],
                       float* x, float* y, float* z, float* x0, float* y0, float* z0, int* comp) {
    const oprep* P = J->P;
    const axx_run* R = J->run;
    uint32_t inst = R->instance;
    uint32_t o[4];
    /* placement in the substrate's cell (x0, y0, z0); the walk position is
     * its image in the unit cell when the walk runs in one */
    const int unit = P->ca != P->a || P->cb != P->b;
    if (R->init_mode == AXX_INIT_UNIFORM) {
        draw(wid, 0, STREAM_INIT << 24, inst, key, o);
        *x0 = unif(o[0]) * P->ca;
        *y0 = unif(o[1]) * P->cb;
        *z0 = unif(o[2]) * P->c;
        *x = unit ? unit_wrap(*x0, P->a) : *x0;
        *y = unit ? unit_wrap(*y0, P->b) : *y0;
        *z = *z0;
        *comp = region_of(P, *x, *y, *z);
        return 0;
    }
    if (R->init_mode == AXX_INIT_COMPARTMENT)
        for (int i = 0; i < P->n_img; ++i)
            if (P->img[i].comp == R->init_arg && P->img[i].type == AXX_PRIM_PLANE) return AXX_ERR_INVALID;
    for (uint32_t it = 0; it < (1u << 20); ++it) {
        draw(wid, it, STREAM_INIT << 24, inst, key, o);
        float px, py, pz;
        if (R->init_mode == AXX_INIT_COMPARTMENT) {
            float qx = unif(o[0]) * P->ca, qy = unif(o[1]) * P->cb;
            pz = unif(o[2]) * P->c;
            px = unit ? unit_wrap(qx, P->a) : qx;
            py = unit ? unit_wrap(qy, P->b) : qy;
            int cm = region_of(P, px, py, pz);
            if (cm != R->init_arg) continue;
            *x = px; *y = py; *z = pz; *comp = cm;
            *x0 = qx; *y0 = qy; *z0 = pz;
            return 0;
        } else { /* inside primitive k: its bounding cube (cylinderXY.cuh:176-183) */
            int k = R->init_arg;
            const oprim* pk = NULL;
            for (int i = 0; i < P->n_img; ++i)
                if (P->img[i].orig == k) { pk = &P->img[i]; break; }
            if (!pk || pk->type == AXX_PRIM_PLANE) return AXX_ERR_INVALID;
            float Rf = pk->R;
            px = pk->cx + (unif(o[0]) * 2.0f - 1.0f) * Rf;
            py = pk->cy + (unif(o[1]) * 2.0f - 1.0f) * Rf;
            pz = pk->cz + (unif(o[2]) * 2.0f - 1.0f) * Rf;
            if (!inside_at(pk, px, py, pz)) continue;
            px = wrap1(px, P->a); py = wrap1(py, P->b); pz = wrap1(pz, P->c);
            *x = px; *y = py; *z = pz;
            *x0 = px; *y0 = py; *z0 = pz;
            *comp = region_of(P, px, py, pz);
            return 0;
        }
    }
    (void)w;
    return AXX_ERR_INIT;
}

static void* job_run(void* arg) {
    job_t* J = (job_t*)arg;
    const oprep* P = J->P;
    const axx_acq* q = J->q;
    const axx_run* R = J->run;
    const int Pn = J->Pn;
    const int mode = R->boundary_mode;
    const int cap = R->bounce_cap > 0 ? R->bounce_cap : AXX_DEFAULT_BOUNCE_CAP;
    const uint32_t key[2] = {(uint32_t)R->seed, (uint32_t)(R->seed >> 32)};
    const uint32_t inst = R->instance;
    float m[AXX_MAX_PROFILES][3];
    d3 Mg[AXX_MAG_GROUP];
    const magjob_t* G = J->mag;

    for (int64_t w = J->w_begin; w < J->w_end; ++w) {
        uint32_t wid = R->walker_offset + (uint32_t)w;
        float x, y, z, dx, dy, dz, E, lf, bud;
        int comp, li;
        uint32_t ncoll, ntran, ncap;
        if (J->i0 == 0) {
            float px0, py0, pz0;
            int rc = init_walker(J, w, wid, key, &x, &y, &z, &px0, &py0, &pz0, &comp);
            if (rc) { J->err = rc; return NULL; }
            dx = dy = dz = 0.0f;
            E = 0.0f; lf = 0.0f; li = 1;
            ncoll = ntran = ncap = 0;
            {   /* budget of the first transmission (ordinal 0) */
                uint32_t ob[4];
                draw(wid, 0u, STREAM_COLL << 24, inst, key, ob);
                bud = neg2log_unif(ob[0]);
            }
            memset(m, 0, sizeof(m));
            FS(AXX_F_X0)[w] = px0; FS(AXX_F_Y0)[w] = py0; FS(AXX_F_Z0)[w] = pz0;
            IS(AXX_F_COMP0)[w] = comp;
            if (G) {
                /* M = m0, then the update at t = 0 at the start point */
                for (int j = 0; j < G->F; ++j) {
                    Mg[j].x = G->f->m0[0]; Mg[j].y = G->f->m0[1]; Mg[j].z = G->f->m0[2];
                }
                if (J->i1 > 0) mag_update(G, Mg, 0, px0, py0, pz0, 0.0f, 0.0f, 0.0f, comp);
            }
        } else {
            x = FS(AXX_F_X)[w]; y = FS(AXX_F_Y)[w]; z = FS(AXX_F_Z)[w];
            dx = FS(AXX_F_DX)[w]; dy = FS(AXX_F_DY)[w]; dz = FS(AXX_F_DZ)[w];
            E = FS(AXX_F_E)[w]; lf = FS(AXX_F_LF)[w];
            comp = IS(AXX_F_COMP)[w]; li = IS(AXX_F_LI)[w];
            ncoll = (uint32_t)IS(AXX_F_NCOLL)[w]; ntran = (uint32_t)IS(AXX_F_NTRAN)[w];
            ncap = (uint32_t)IS(AXX_F_NCAP)[w];
            bud = FS(AXX_F_BUD)[w];
            if (G) {
                const uint32_t* U = J->st;
                for (int j = 0; j < G->F; ++j) {
                    double v[3];
                    for (int c = 0; c < 3; ++c) {
                        uint64_t lo = U[(size_t)(AXX_F_MOM + 6 * j + 2 * c) * J->N + w];
                        uint64_t hi = U[(size_t)(AXX_F_MOM + 6 * j + 2 * c + 1) * J->N + w];
                        uint64_t b = lo | (hi << 32);
                        memcpy(&v[c], &b, 8);
                    }
                    Mg[j].x = v[0]; Mg[j].y = v[1]; Mg[j].z = v[2];
                }
            } else {
                for (int p = 0; p < Pn; ++p)
                    for (int a = 0; a < 3; ++a) m[p][a] = FS(AXX_F_MOM + 3 * p + a)[w];
            }
        }
        const float x0 = FS(AXX_F_X0)[w], y0 = FS(AXX_F_Y0)[w], z0 = FS(AXX_F_Z0)[w];
        int ibeg = J->i0 < 1 ? 1 : J->i0;
        uint32_t o[4] = {0, 0, 0, 0}, og[4] = {0, 0, 0, 0};
        for (int i = ibeg; i < J->i1; ++i) {
            /* draws: one Philox call per pair of steps (i-1)>>1 (Gaussian
             * steps: a second one on stream 3) */
            uint32_t h = (uint32_t)(i - 1) & 1u;
            if (h == 0 || i == ibeg) {
                draw(wid, (uint32_t)(i - 1) >> 1, STREAM_STEP << 24, inst, key, o);
                if (R->step_mode == AXX_STEP_GAUSSIAN)
                    draw(wid, (uint32_t)(i - 1) >> 1, STREAM_GAUSS << 24, inst, key, og);
            }
            uint32_t wa = o[2 * h], wb = o[2 * h + 1];
            float ux, uy, uz, gs = 1.0f;
            if (R->step_mode == AXX_STEP_ISOTROPIC) {
                /* cos(theta) = 2u - 1 in one rounding; sin^2 = 1 - c^2 in one rounding */
                float c = fmaf((float)((wb >> 8) | 1u), 0x1p-23f, -1.0f);
                float sth = sqrtf(fmaf(-c, c, 1.0f));
                float sphi, cphi;
                orc_sincos_turn(wa, &sphi, &cphi);
                ux = sth * cphi; uy = sth * sphi; uz = c;
            } else if (R->step_mode == AXX_STEP_AXES) {
                /* +-sqrt(2 D dt) per axis (kernelPhase.cuh:31-35, length kept) */
                const float k3 = 0.57735026f;
                ux = (wa & 1u) ? -k3 : k3;
                uy = (wa & 2u) ? -k3 : k3;
                uz = (wa & 4u) ? -k3 : k3;
            } else {
                /* N(0, 2 D dt) per axis by Box-Muller (kernelWC.cuh:56-67) */
                uint32_t wc = og[2 * h], wd = og[2 * h + 1];
                float r1 = sqrtf(neg2log_unif(wa)), r2 = sqrtf(neg2log_unif(wc));
                float s1, c1, s2, c2;
                orc_sincos_turn(wb, &s1, &c1);
                orc_sincos_turn(wd, &s2, &c2);
                float n1 = r1 * c1, n2 = r1 * s1, n3 = r2 * c2;
                float nn = sqrtf(fmaf(n3, n3, fmaf(n2, n2, n1 * n1)));
                float inv = 1.0f / nn;
                ux = n1 * inv; uy = n2 * inv; uz = n3 * inv;
                gs = nn * 0.57735026f;
            }
            float ell = gs * P->L[comp];
            int lastplane = -1;
            float sx = x, sy = y, sz = z;
            const int* list;
            int cnt, dense;
            cand_range(P, J->use_bins, x, y, z, &list, &cnt, &dense);
            if (mode == AXX_BOUNDARY_REJECTION) {
                float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
                int rej = 0;
                for (int j = 0; j < cnt; ++j) {
                    const oprim* pk = &P->img[dense ? j : list[j]];
                    float tp;
                    /* a plane rejects the step that would come within
                     * plane_eps of it (it bounds no region) */
                    if (pk->type == AXX_PRIM_PLANE
                            ? plane_test(pk, x, y, z, sx, sy, sz, ux, uy, uz, ell, P->plane_eps, &tp)
                            : inside_at(pk, sx, sy, sz) != inside_at(pk, ex, ey, ez)) {
                        rej = 1;
                        break;
                    }
                }
                if (rej) { ncoll++; ell = 0.0f; }
            } else {
                float Axy = fmaf(uy, uy, ux * ux);
                float A3 = fmaf(uz, uz, Axy);
                int nb = 0;
                int excl = -1;   /* image just left outward: a convex surface
                                  * (or a plane) cannot be hit again */
                lastplane = -1;
                for (;;) {
                    float best = INFINITY;
                    int bk = -1, bins_ = 0;
                    for (int j = 0; j < cnt; ++j) {
                        int k = dense ? j : list[j];
                        if (k == excl) continue;
                        const oprim* pk = &P->img[k];
                        float t;
                        int ins = 0, hit;
                        if (pk->type == AXX_PRIM_PLANE)
                            hit = plane_test(pk, x, y, z, sx, sy, sz, ux, uy, uz, ell, P->plane_eps, &t);
                        else
                            hit = hit_test(pk, comp == pk->comp, P->H2c, P->H2s, sx, sy, sz, ux, uy, uz,
                                           Axy, A3, ell, &t, &ins);
                        if (hit && t < best) {
                            best = t;
                            bk = k;
                            bins_ = ins;
                        }
                    }
                    if (bk < 0) break;
                    const oprim* pk = &P->img[bk];
                    float t = best;
                    ncoll++;
                    float px = fmaf(t, ux, sx), py = fmaf(t, uy, sy), pz = fmaf(t, uz, sz);
                    int inside = bins_;
                    int transmit = 0;
                    if (mode == AXX_BOUNDARY_PERMEABLE && pk->type != AXX_PRIM_PLANE) {
                        /* the budget (a 2 Exp(1) draw) pays this side's hazard
                         * h = -2 ln(1 - 4 kappa / v); the event transmits when
                         * it runs out: probability 4 kappa / v, independent of
                         * the events before (memoryless) */
                        float h = inside ? P->thr_out[pk->orig] : P->thr_in[pk->orig];
#if ORC_PLANT == 1
                        h = h + h;                       /* planted: twice the hazard (p -> 2p - p^2) */
#elif ORC_PLANT == 3
                        h = P->thr_in[pk->orig];         /* planted: lattice speed on both sides */
#endif
                        bud = bud - h;
                        transmit = bud <= 0.0f;
                        J->ev[inside ? 0 : 2]++;
                        J->ev[inside ? 1 : 3] += transmit;
                    }
                    lastplane = pk->type == AXX_PRIM_PLANE ? bk : -1;
                    /* left outward: outside reflection, transmission from
                     * inside, or a plane reflection */
                    excl = (pk->type == AXX_PRIM_PLANE || inside == transmit) ? bk : -1;
                    if (transmit) {
                        int nc = inside ? 0 : pk->comp;
                        float rem = ell - t;
                        float f = rem / (gs * P->L[comp]);
                        float ph = 1.0f - f;
#if ORC_PLANT == 2
                        ph = 0.0f;   /* planted: the whole step charged to the far side (defect D2) */
#endif
                        E = fmaf((float)(i - li) + (ph - lf), P->rate[comp], E);
                        li = i;
                        lf = ph;
                        comp = nc;
                        ell = f * (gs * P->L[nc]);
                        ntran++;
#if ORC_PLANT == 4
                        /* planted: the budget is not redrawn after a transmission */
#else
                        {   /* the next budget: transmission ordinal ntran */
                            uint32_t ob[4];
#if ORC_PLANT == 5
                            draw(wid, 0u, STREAM_COLL << 24, inst, key, ob);   /* planted: always budget 0 */
#else
                            draw(wid, ntran, STREAM_COLL << 24, inst, key, ob);
#endif
                            bud = neg2log_unif(ob[0]);
                        }
#endif
                    } else {
                        float nx = px - pk->cx, ny = py - pk->cy, nz = 0.0f, kd;
                        if (pk->type == AXX_PRIM_PLANE) {
                            nx = pk->nx; ny = pk->ny; nz = pk->nz;
                            kd = fmaf(nz, uz, fmaf(ny, uy, nx * ux));
                        } else if (pk->type == AXX_PRIM_SPHERE) {
                            nz = pz - pk->cz;
                            float inv = 1.0f / pk->R;
                            nx = nx * inv; ny = ny * inv; nz = nz * inv;
                            kd = fmaf(nz, uz, fmaf(ny, uy, nx * ux));
                        } else {
                            float inv = 1.0f / pk->R;
                            nx = nx * inv; ny = ny * inv;
                            kd = fmaf(ny, uy, nx * ux);
                        }
                        float k2 = kd + kd;
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
            float ex = fmaf(ell, ux, sx), ey = fmaf(ell, uy, sy), ez = fmaf(ell, uz, sz);
            if (lastplane >= 0) {
                /* after a plane reflection keep at least plane_eps from the
                 * plane on the walker's side, so rounding cannot put the next
                 * step's start on the far side */
                const oprim* pl = &P->img[lastplane];
                float s0 = fmaf(pl->nz, z, fmaf(pl->ny, y, pl->nx * x)) - pl->d;
                float se = fmaf(pl->nz, ez, fmaf(pl->ny, ey, pl->nx * ex)) - pl->d;
                float sg = s0 >= 0.0f ? 1.0f : -1.0f;
                float gap = P->plane_eps - sg * se;
                if (gap > 0.0f) {
                    ex = fmaf(sg * gap, pl->nx, ex);
                    ey = fmaf(sg * gap, pl->ny, ey);
                    ez = fmaf(sg * gap, pl->nz, ez);
                }
            }
            dx = dx + (ex - x);
            dy = dy + (ey - y);
            dz = dz + (ez - z);
            /* wrap after the last step of each lookup group */
            if ((i & P->gmask) == 0) {
                x = wrap1(ex, P->a);
                y = wrap1(ey, P->b);
                /* with cylinders only, z enters no geometry: it is not wrapped */
                z = P->any_3d ? wrap1(ez, P->c) : ez;
            } else {
                x = ex;
                y = ey;
                z = ez;
            }
            if (G) {
                mag_update(G, Mg, i, x0, y0, z0, dx, dy, dz, comp);
                continue;
            }
            /* phase moments (kernelLattice.cuh:81-85, factored by profile) */
            const float* g = q->profiles + (size_t)i * Pn;
            for (int p = 0; p < Pn; ++p) {
                float gp = g[p];
                if (J->axes & 1) m[p][0] = fmaf(gp, dx, m[p][0]);
                if (J->axes & 2) m[p][1] = fmaf(gp, dy, m[p][1]);
                if (J->axes & 4) m[p][2] = fmaf(gp, dz, m[p][2]);
            }
        }
        FS(AXX_F_X)[w] = x; FS(AXX_F_Y)[w] = y; FS(AXX_F_Z)[w] = z;
        FS(AXX_F_DX)[w] = dx; FS(AXX_F_DY)[w] = dy; FS(AXX_F_DZ)[w] = dz;
        FS(AXX_F_E)[w] = E; FS(AXX_F_LF)[w] = lf;
        IS(AXX_F_COMP)[w] = comp; IS(AXX_F_LI)[w] = li;
        IS(AXX_F_NCOLL)[w] = (int32_t)ncoll; IS(AXX_F_NTRAN)[w] = (int32_t)ntran;
        IS(AXX_F_NCAP)[w] = (int32_t)ncap;
        IS(AXX_F_STEP)[w] = J->i1;
        FS(AXX_F_BUD)[w] = bud;
        if (G) {
            uint32_t* U = J->st;
            for (int j = 0; j < G->F; ++j) {
                const double v[3] = {Mg[j].x, Mg[j].y, Mg[j].z};
                for (int c = 0; c < 3; ++c) {
                    uint64_t b;
                    memcpy(&b, &v[c], 8);
                    U[(size_t)(AXX_F_MOM + 6 * j + 2 * c) * J->N + w] = (uint32_t)b;
                    U[(size_t)(AXX_F_MOM + 6 * j + 2 * c + 1) * J->N + w] = (uint32_t)(b >> 32);
                }
            }
        } else {
            for (int p = 0; p < Pn; ++p)
                for (int a = 0; a < 3; ++a) FS(AXX_F_MOM + 3 * p + a)[w] = m[p][a];
        }
    }
    return NULL;
}

/* Axes that carry phase, widened to the smallest of {x}, {x,y}, {x,z},
 * {x,y,z} that contains them (the set of moment axes the walk keeps). */
int orc_axes_mask(const axx_acq* q) {
    int mask = 0;
    for (int j = 0; j < q->n_waveforms; ++j)
        for (int a = 0; a < 3; ++a)
            if (q->wf_dir[3 * j + a] != 0.0 && q->wf_amplitude[j] != 0.0) mask |= 1 << a;
    static const int widen[8] = {1, 1, 3, 3, 5, 5, 7, 7};
    return widen[mask];
}

static int walk_common(const axx_substrate* s, const axx_acq* q, const axx_run* run, int step_begin,
                       int step_end, uint32_t* state, int nthreads, double bin_width, const magjob_t* mag,
                       int64_t* ev_out) {
    if (!s || !q || !run || !state) return AXX_ERR_INVALID;
    if (run->n_walkers < 1) return AXX_ERR_INVALID;
    if (step_begin < 0 || step_end > q->n_steps || step_begin > step_end) return AXX_ERR_INVALID;
    if (run->boundary_mode < 0 || run->boundary_mode > 2) return AXX_ERR_INVALID;
    if (run->step_mode < 0 || run->step_mode > 2) return AXX_ERR_INVALID;
    oprep P;
    const int allow_unit = g_unit_cells && (run->init_mode == AXX_INIT_UNIFORM ||
                                            (run->init_mode == AXX_INIT_COMPARTMENT && run->init_arg == 0));
    int rc = oprep_build_view(&P, s, q, bin_width, step_reach_factor(run->step_mode), allow_unit);
    if (!rc && run->step_mode == AXX_STEP_GAUSSIAN) {
        double need = 2.0 * step_reach_factor(AXX_STEP_GAUSSIAN) * (double)P.Lmax;
        if (need >= s->a || need >= s->b || need >= s->c) rc = AXX_ERR_GEOMETRY;
    }
    if (rc) { oprep_free(&P); return rc; }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads > run->n_walkers) nthreads = (int)run->n_walkers;
    job_t* jobs = calloc(nthreads, sizeof(job_t));
    pthread_t* th = calloc(nthreads, sizeof(pthread_t));
    int axes = orc_axes_mask(q);
    for (int t = 0; t < nthreads; ++t) {
        job_t* J = &jobs[t];
        J->P = &P; J->q = q; J->run = run; J->st = state; J->N = run->n_walkers;
        J->Pn = q->n_profiles; J->axes = axes; J->i0 = step_begin; J->i1 = step_end;
        J->use_bins = bin_width > 0;
        J->mag = mag;
        J->w_begin = run->n_walkers * t / nthreads;
        J->w_end = run->n_walkers * (t + 1) / nthreads;
        if (nthreads > 1) pthread_create(&th[t], NULL, job_run, J);
        else job_run(J);
    }
    if (nthreads > 1)
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    for (int t = 0; t < nthreads; ++t) {
        if (jobs[t].err) rc = jobs[t].err;
        if (ev_out)
            for (int e = 0; e < 4; ++e) ev_out[e] += jobs[t].ev[e];
    }
    free(jobs); free(th);
    oprep_free(&P);
    return rc;
}

int orc_walk(const axx_substrate* s, const axx_acq* q, const axx_run* run, int step_begin,
             int step_end, uint32_t* state, int nthreads, double bin_width) {
    return walk_common(s, q, run, step_begin, step_end, state, nthreads, bin_width, NULL, NULL);
}

/* orc_walk that also adds the walk's boundary events at circles and
 * spheres to ev[4]: {from inside, transmitted from inside, from outside,
 * transmitted from outside} (the decision of boundaryCheck.cuh:94-111). */
int orc_walk_events(const axx_substrate* s, const axx_acq* q, const axx_run* run, int step_begin,
                    int step_end, uint32_t* state, int nthreads, double bin_width, int64_t ev[4]) {
    return walk_common(s, q, run, step_begin, step_end, state, nthreads, bin_width, NULL, ev);
}

/* Magnetization walk of fields [group*AXX_MAG_GROUP, +F) (updateWalkersMag,
 * kernelMag.cuh:1-141): the walk above with M per field in place of the
 * phase moments; state holds AXX_F_MOM + 6 F words per walker. */
int orc_walk_mag(const axx_substrate* s, const axx_field* f, const axx_run* run, int group, int step_begin,
                 int step_end, uint32_t* state, int nthreads, double bin_width) {
    if (!s || !f || !run || !state || !f->samples) return AXX_ERR_INVALID;
    if (f->n_steps < 1 || f->n_fields < 1 || f->integrator < 0 || f->integrator > 2) return AXX_ERR_INVALID;
    if (group < 0 || group * AXX_MAG_GROUP >= f->n_fields) return AXX_ERR_INVALID;
    magjob_t G;
    G.f = f;
    G.g0 = group * AXX_MAG_GROUP;
    G.F = f->n_fields - G.g0 < AXX_MAG_GROUP ? f->n_fields - G.g0 : AXX_MAG_GROUP;
    int* off = malloc(((size_t)f->n_steps + 1) * sizeof(int));
    double* iT1 = calloc((size_t)s->n_prims + 1, sizeof(double));
    double* iT2 = calloc((size_t)s->n_prims + 1, sizeof(double));
    int64_t tot = 0;
    for (int i = 0; i < f->n_steps; ++i) {
        off[i] = (int)tot;
        tot += mag_samples(f->integrator, f->substeps ? f->substeps[i] : 1);
    }
    off[f->n_steps] = (int)tot;
    for (int k = 0; k <= s->n_prims; ++k) {
        double t1 = f->T1 ? f->T1[k] : INFINITY;
        iT1[k] = (t1 > 0 && isfinite(t1)) ? 1.0 / t1 : 0.0;
        double t2 = k == 0 ? s->T2_lat : s->T2[k - 1];
        iT2[k] = (t2 > 0 && isfinite(t2)) ? 1.0 / t2 : 0.0;
    }
    G.off = off; G.iT1 = iT1; G.iT2 = iT2;
    G.relax = run->relaxation != 0 && f->integrator != AXX_MAG_ROTATE;
    /* the walk core sees a one-profile acquisition (its moments unused) */
    float* prof = calloc((size_t)f->n_steps, sizeof(float));
    int wfp = 0;
    double amp = 0.0, dir[3] = {1.0, 0.0, 0.0};
    axx_acq q;
    memset(&q, 0, sizeof(q));
    q.n_steps = f->n_steps; q.dt = f->dt; q.gamma = f->gamma;
    q.n_profiles = 1; q.profiles = prof;
    q.n_waveforms = 1; q.wf_profile = &wfp; q.wf_amplitude = &amp; q.wf_dir = dir;
    int rc = walk_common(s, &q, run, step_begin, step_end, state, nthreads, bin_width, &G, NULL);
    free(off); free(iT1); free(iT2); free(prof);
    return rc;
}

/* Sums over walkers of Mx, My, Mz for the F fields of a group's state. */
void orc_reduce_mag(const uint32_t* state, int64_t n_walkers, int F, double* out) {
    for (int j = 0; j < F; ++j)
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int64_t w = 0; w < n_walkers; ++w) {
                uint64_t lo = state[(size_t)(AXX_F_MOM + 6 * j + 2 * c) * n_walkers + w];
                uint64_t hi = state[(size_t)(AXX_F_MOM + 6 * j + 2 * c + 1) * n_walkers + w];
                uint64_t b = lo | (hi << 32);
                double v;
                memcpy(&v, &b, 8);
                acc += v;
            }
            out[j * 3 + c] = acc;
        }
}

/* ------------------------------------------------------------------ */
/* Signal sums in fp64 (phaseAcquisitionStream.cuh:171-189, misc.cuh:21-39):
 *   phi_j = gamma*dt*amp_j * sum_a dir_ja * (m_{p(j),a} + r0_a * S_p),
 *   S_p = sum_i profile_p(t_i);  weight = exp(-(E + tail*rate)). */
/* T2 exponent of walker w at the end of an n_steps walk: the stays summed
 * so far plus the last one, (n - li - lf) steps in the compartment it is in
 * (kernelLattice.cuh:87-96 with the true split; T2factor = -exponent). */
static double t2_exponent(const oprep* P, const uint32_t* state, int64_t N, int n, int64_t w) {
    const float* F = (const float*)state;
    const int32_t* I = (const int32_t*)state;
    const int comp = I[AXX_F_COMP * N + w];
    const int li = I[AXX_F_LI * N + w];
    return (double)F[AXX_F_E * N + w] + ((double)(n - li) - (double)F[AXX_F_LF * N + w]) * (double)P->rate[comp];
}

int orc_exponents(const axx_substrate* s, const axx_acq* q, const axx_run* run, const uint32_t* state,
                  double* out) {
    if (!s || !q || !run || !state || !out) return AXX_ERR_INVALID;
    oprep P;
    int rc = oprep_build(&P, s, q, 0.0, 1.0);
    if (rc) { oprep_free(&P); return rc; }
    for (int64_t w = 0; w < run->n_walkers; ++w) out[w] = t2_exponent(&P, state, run->n_walkers, q->n_steps, w);
    oprep_free(&P);
    return AXX_OK;
}

int orc_reduce(const axx_substrate* s, const axx_acq* q, const axx_run* run,
               const uint32_t* state, axx_result* out) {
    if (!s || !q || !run || !state || !out) return AXX_ERR_INVALID;
    oprep P;
    int rc = oprep_build(&P, s, q, 0.0, 1.0);
    if (rc) { oprep_free(&P); return rc; }
    int64_t N = run->n_walkers;
    int M = q->n_waveforms, Pn = q->n_profiles, n = q->n_steps;
    double* S = calloc(Pn, sizeof(double));
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < Pn; ++p) S[p] += (double)q->profiles[(size_t)i * Pn + p];
    double* acc = calloc((size_t)6 * M, sizeof(double));
    const float* F = (const float*)state;
    const int32_t* I = (const int32_t*)state;
    int64_t nintra = 0;
    uint64_t coll = 0, tran = 0, capn = 0;
    for (int64_t w = 0; w < N; ++w) {
        double Ef = t2_exponent(&P, state, N, n, w);
        double wt = run->relaxation ? exp(-Ef) : 1.0;
        int intra = I[AXX_F_COMP0 * N + w] != 0;
        nintra += intra;
        coll += (uint32_t)I[AXX_F_NCOLL * N + w];
        tran += (uint32_t)I[AXX_F_NTRAN * N + w];
        capn += (uint32_t)I[AXX_F_NCAP * N + w];
        double r0[3] = {F[AXX_F_X0 * N + w], F[AXX_F_Y0 * N + w], F[AXX_F_Z0 * N + w]};
        for (int j = 0; j < M; ++j) {
            int p = q->wf_profile[j];
            double ph = 0.0;
            for (int a = 0; a < 3; ++a) {
                double d = q->wf_dir[3 * j + a];
                if (d == 0.0) continue;
                ph += d * ((double)F[(AXX_F_MOM + 3 * p + a) * N + w] + r0[a] * S[p]);
            }
            double phi = q->gamma * q->dt * q->wf_amplitude[j] * ph;
            double cr = cos(phi) * wt, sr = sin(phi) * wt;
            acc[6 * j + 0] += cr;
            acc[6 * j + 1] += sr;
            acc[6 * j + 2] += cr * cr;
            acc[6 * j + 3] += sr * sr;
            if (intra) { acc[6 * j + 4] += cr; acc[6 * j + 5] += sr; }
        }
    }
    for (int j = 0; j < M; ++j) {
        if (out->sum_cos) out->sum_cos[j] = acc[6 * j + 0];
        if (out->sum_sin) out->sum_sin[j] = acc[6 * j + 1];
        if (out->sum_cos2) out->sum_cos2[j] = acc[6 * j + 2];
        if (out->sum_sin2) out->sum_sin2[j] = acc[6 * j + 3];
        if (out->intra_cos) out->intra_cos[j] = acc[6 * j + 4];
        if (out->intra_sin) out->intra_sin[j] = acc[6 * j + 5];
    }
    out->n_walkers = N;
    out->n_intra = nintra;
    out->collisions = coll;
    out->transmissions = tran;
    out->cap_hits = capn;
    free(S); free(acc);
    oprep_free(&P);
    return AXX_OK;
}

/* Unit cells of the substrate's cell that a walk of `run` runs in one of. */
int orc_unit_cells(const axx_substrate* s, const axx_acq* q, const axx_run* run) {
    if (!s || !q || !run) return AXX_ERR_INVALID;
    oprep P;
    const int allow_unit = g_unit_cells && (run->init_mode == AXX_INIT_UNIFORM ||
                                            (run->init_mode == AXX_INIT_COMPARTMENT && run->init_arg == 0));
    int rc = oprep_build_view(&P, s, q, 0.0, step_reach_factor(run->step_mode), allow_unit);
    int u = rc ? rc : P.units;
    oprep_free(&P);
    return u;
}

/* Expanded primitive table (for tests of the image construction). */
int orc_image_count(const axx_substrate* s, const axx_acq* q) {
    oprep P;
    int rc = oprep_build(&P, s, q, 0.0, 1.0);
    int n = rc ? rc : P.n_img;
    oprep_free(&P);
    return n;
}

int orc_hit_test(int type, float cx, float cy, float cz, float R, int inside, const float s3[3],
                 const float u3[3], float ell, float* t) {
    oprim p = {cx, cy, cz, R * R, R, type, 1, 0, 0.0f, 0.0f, 0.0f, 0.0f};
    float Axy = fmaf(u3[1], u3[1], u3[0] * u3[0]);
    float A3 = fmaf(u3[2], u3[2], Axy);
    int ins;
    return hit_test(&p, inside, INFINITY, INFINITY, s3[0], s3[1], s3[2], u3[0], u3[1], u3[2], Axy,
                    A3, ell, t, &ins);
}
