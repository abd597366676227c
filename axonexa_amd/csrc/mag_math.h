// mag_math.h — fp64 magnetization updates of the Bloch path.
//
// The lab-frame precession angle gamma |B| dt is dominated by the static
// field B0 while the signal lives in the small gradient part of B, so the
// magnetization is integrated in double precision (as the reference does with
// USE_DOUBLE).  Every operation is written out (fma where intended, the
// library is built with -ffp-contract=off); sqrt and '/' are the correctly
// rounded IEEE operations, and sine/cosine are this file's own, so the CPU
// oracle reproduces each walker's magnetization bit for bit.
//
// Reference counterparts: updateMag_noT2_rotate, bloch, updateMagRK4,
// updateMagDORP (src/bloch/blochdiff.cuh:7-210).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axx {

struct d3 {
    double x, y, z;
};

// Correctly rounded square root: sqrt() and a residual test of its two
// neighbours (the result is the correctly rounded root whenever sqrt() is
// within one ulp of it).
__device__ __forceinline__ double sqrt_cr(double v) {
    double y = sqrt(v);
    const double ydn = __longlong_as_double(__double_as_longlong(y) - 1);
    const double yup = __longlong_as_double(__double_as_longlong(y) + 1);
    const double rdn = fma(-ydn, y, v);
    const double rup = fma(-yup, y, v);
    y = rdn <= 0.0 ? ydn : y;
    y = rup > 0.0 ? yup : y;
    return y;
}

// sin and cos of x: k = nearest integer to x 2/pi, r = x - k pi/2 with a
// three-part split of pi/2 (Cody-Waite, fused), Taylor polynomials of sin
// (to r^17) and cos (to r^16) on |r| <= pi/4, quadrant from k mod 4.
// Accurate to about one ulp for |x| < 2^20.
__device__ __forceinline__ void sincos64(double x, double& s, double& c) {
    const double k = rint(x * 0x1.45f306dc9c883p-1);
    double r = fma(-k, 0x1.921fb54442d18p+0, x);
    r = fma(-k, 0x1.1a62633145c07p-54, r);
    r = fma(-k, -0x1.f1976b7ed8fbcp-110, r);
    const double z = r * r;
    double ps = 0x1.952c77030ad4ap-49;        //  1/17!
    ps = fma(z, ps, -0x1.ae7f3e733b81fp-41);  // -1/15!
    ps = fma(z, ps, 0x1.6124613a86d09p-33);   //  1/13!
    ps = fma(z, ps, -0x1.ae64567f544e4p-26);  // -1/11!
    ps = fma(z, ps, 0x1.71de3a556c734p-19);   //  1/9!
    ps = fma(z, ps, -0x1.a01a01a01a01ap-13);  // -1/7!
    ps = fma(z, ps, 0x1.1111111111111p-7);    //  1/5!
    ps = fma(z, ps, -0x1.5555555555555p-3);   // -1/3!
    const double sr = fma(r * z, ps, r);
    double pc = 0x1.ae7f3e733b81fp-45;        //  1/16!
    pc = fma(z, pc, -0x1.93974a8c07c9dp-37);  // -1/14!
    pc = fma(z, pc, 0x1.1eed8eff8d898p-29);   //  1/12!
    pc = fma(z, pc, -0x1.27e4fb7789f5cp-22);  // -1/10!
    pc = fma(z, pc, 0x1.a01a01a01a01ap-16);   //  1/8!
    pc = fma(z, pc, -0x1.6c16c16c16c17p-10);  // -1/6!
    pc = fma(z, pc, 0x1.5555555555555p-5);    //  1/4!
    pc = fma(z, pc, -0.5);
    const double cr = fma(z, pc, 1.0);
    const int q = (int)k & 3;
    const double a = (q & 1) ? cr : sr;
    const double b = (q & 1) ? sr : cr;
    s = (q & 2) ? -a : a;
    c = (q == 1 || q == 2) ? -b : b;
}

// B = b0 + J r for one field sample rec = (b0[3], J[3][3] row major).
template <class PTR>
__device__ __forceinline__ d3 field_at(PTR rec, double rx, double ry, double rz) {
    d3 B;
    B.x = fma(rec[5], rz, fma(rec[4], ry, fma(rec[3], rx, rec[0])));
    B.y = fma(rec[8], rz, fma(rec[7], ry, fma(rec[6], rx, rec[1])));
    B.z = fma(rec[11], rz, fma(rec[10], ry, fma(rec[9], rx, rec[2])));
    return B;
}

// Rotation of M about B by the angle -gh |B| (updateMag_noT2_rotate,
// blochdiff.cuh:186-199): with ib = 1/|B| and C = B (B.M) ib^2 the part
// along B, M' = C + (M - C) cos + (B x M) sin ib.  No rotation when B = 0
// (the reference divides by |B| there).  Branch-free, so the rotations of
// several fields interleave.
__device__ __forceinline__ void rotate(d3& M, const d3 B, double gh) {
    const double b2 = fma(B.z, B.z, fma(B.y, B.y, B.x * B.x));
    const bool on = b2 > 0.0;
    const double bn = sqrt_cr(on ? b2 : 1.0);
    const double ib = 1.0 / bn;
    double s, c;
    sincos64(-(gh * bn), s, c);
    const double f = (fma(B.z, M.z, fma(B.y, M.y, B.x * M.x)) * ib) * ib;
    const double sb = s * ib;
    const double cx = f * B.x, cy = f * B.y, cz = f * B.z;
    const double wx = fma(B.y, M.z, -(B.z * M.y));
    const double wy = fma(B.z, M.x, -(B.x * M.z));
    const double wz = fma(B.x, M.y, -(B.y * M.x));
    const double nx = fma(wx, sb, fma(M.x - cx, c, cx));
    const double ny = fma(wy, sb, fma(M.y - cy, c, cy));
    const double nz = fma(wz, sb, fma(M.z - cz, c, cz));
    M.x = on ? nx : M.x;
    M.y = on ? ny : M.y;
    M.z = on ? nz : M.z;
}

// h dM/dt with dM/dt = gamma M x B, minus (Mx/T2, My/T2, (Mz - 1)/T1) when
// relaxing (bloch, blochdiff.cuh:7-22); i1, i2 = 1/T1, 1/T2.
__device__ __forceinline__ d3 bloch_h(const d3 M, const d3 B, double gam, double h, bool relax, double i1,
                                      double i2) {
    d3 d;
    d.x = gam * fma(M.y, B.z, -(M.z * B.y));
    d.y = gam * fma(M.z, B.x, -(M.x * B.z));
    d.z = gam * fma(M.x, B.y, -(M.y * B.x));
    if (relax) {
        d.x = fma(-M.x, i2, d.x);
        d.y = fma(-M.y, i2, d.y);
        d.z = fma(-(M.z - 1.0), i1, d.z);
    }
    d.x = d.x * h;
    d.y = d.y * h;
    d.z = d.z * h;
    return d;
}

__device__ __forceinline__ d3 axpy(double a, const d3 x, const d3 y) {
    return d3{fma(a, x.x, y.x), fma(a, x.y, y.y), fma(a, x.z, y.z)};
}

// Classical fourth-order Runge-Kutta step of length h with the field at the
// start, middle and end of the step (updateMagRK4, blochdiff.cuh:24-67).
__device__ __forceinline__ void rk4(d3& M, const d3 B0, const d3 Bh, const d3 B1, double gam, double h, bool relax,
                                    double i1, double i2) {
    const d3 k1 = bloch_h(M, B0, gam, h, relax, i1, i2);
    const d3 k2 = bloch_h(axpy(0.5, k1, M), Bh, gam, h, relax, i1, i2);
    const d3 k3 = bloch_h(axpy(0.5, k2, M), Bh, gam, h, relax, i1, i2);
    const d3 k4 = bloch_h(axpy(1.0, k3, M), B1, gam, h, relax, i1, i2);
    const d3 t = axpy(2.0, k3, axpy(2.0, k2, k1));
    M = axpy(1.0 / 6.0, d3{t.x + k4.x, t.y + k4.y, t.z + k4.z}, M);
}

// Dormand-Prince stage nodes (the fifth-order solution of the embedded pair
// of updateMagDORP, blochdiff.cuh:68-118; fixed step, no error control).
constexpr double kDP_c[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0};

// One Dormand-Prince step of length h, B[c] the field at t + kDP_c[c] h.
// Row 6 uses a63 = +46732/5247 (the published tableau; its row then sums
// to c6 = 1); blochdiff.cuh:92-96 has -46732/5247.
__device__ __forceinline__ void dopri5(d3& M, const d3* B, double gam, double h, bool relax, double i1, double i2) {
    const d3 k1 = bloch_h(M, B[0], gam, h, relax, i1, i2);
    const d3 k2 = bloch_h(axpy(1.0 / 5.0, k1, M), B[1], gam, h, relax, i1, i2);
    const d3 k3 = bloch_h(axpy(9.0 / 40.0, k2, axpy(3.0 / 40.0, k1, M)), B[2], gam, h, relax, i1, i2);
    const d3 k4 = bloch_h(axpy(32.0 / 9.0, k3, axpy(-56.0 / 15.0, k2, axpy(44.0 / 45.0, k1, M))), B[3], gam, h,
                          relax, i1, i2);
    const d3 k5 = bloch_h(axpy(-212.0 / 729.0, k4,
                               axpy(64448.0 / 6561.0, k3, axpy(-25360.0 / 2187.0, k2, axpy(19372.0 / 6561.0, k1, M)))),
                          B[4], gam, h, relax, i1, i2);
    const d3 k6 = bloch_h(
        axpy(-5103.0 / 18656.0, k5,
             axpy(49.0 / 176.0, k4, axpy(46732.0 / 5247.0, k3, axpy(-355.0 / 33.0, k2, axpy(9017.0 / 3168.0, k1, M))))),
        B[5], gam, h, relax, i1, i2);
    M = axpy(11.0 / 84.0, k6,
             axpy(-2187.0 / 6784.0, k5, axpy(125.0 / 192.0, k4, axpy(500.0 / 1113.0, k3, axpy(35.0 / 384.0, k1, M)))));
}

}  // namespace axx
