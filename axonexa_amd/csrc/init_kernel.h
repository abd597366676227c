// init_kernel.h — initial walker placement.
#pragma once
#include "walk_kernel.h"

namespace axx {

// Images of A staged into LDS as geom[n_img] then meta[n_img]
// (20 * n_img bytes).
__device__ __forceinline__ void stage_images(const WalkArgs& A, unsigned char* lds, float4*& geom,
                                             int*& meta) {
    geom = reinterpret_cast<float4*>(lds);
    meta = reinterpret_cast<int*>(geom + A.n_img);
    for (int i = threadIdx.x; i < A.n_img; i += blockDim.x) {
        geom[i] = A.g_geom[i];
        meta[i] = A.g_meta[i];
    }
    __syncthreads();
}

// Initial placement of walker w (kernelLattice.cuh:29-37, lattice.cuh:31-58,
// cylinderXY.cuh:176-183), all draws on the init stream; writes the full
// walker state (n_mom moment fields zeroed).
__device__ __forceinline__ void init_lane(const WalkArgs& A, const float4* geom, const int* meta,
                                          long long w, int n_mom) {
    const long long N = A.N;
    uint32_t* st = A.state;
    float* fs = reinterpret_cast<float*>(st);
    int* is = reinterpret_cast<int*>(st);
    const uint32_t wid = A.woff + (uint32_t)w;
    bool done = false;
    float x = 0.0f, y = 0.0f, z = 0.0f;
    int comp = 0;
    // placement in the substrate's cell (x0, y0, z0); the walk position is
    // its image in the unit cell when the walk runs in one (ca != a, cb != b)
    const bool unit = A.ca != A.a || A.cb != A.b;
    float x0 = 0.0f, y0 = 0.0f, z0 = 0.0f;
    if (A.init_mode == 0) {
        u32x4 o = philox10(wid, 0u, kStreamInit, A.inst, A.k0, A.k1);
        x0 = unif(o.x) * A.ca;
        y0 = unif(o.y) * A.cb;
        z0 = unif(o.z) * A.c;
        x = unit ? unit_wrap(x0, A.a) : x0;
        y = unit ? unit_wrap(y0, A.b) : y0;
        z = z0;
        comp = region_of(geom, meta, A.n_img, x, y, z);
        done = true;
    } else {
        for (uint32_t it = 0; it < (1u << 20) && !done; ++it) {
            u32x4 o = philox10(wid, it, kStreamInit, A.inst, A.k0, A.k1);
            if (A.init_mode == 1) {
                float qx = unif(o.x) * A.ca, qy = unif(o.y) * A.cb, pz = unif(o.z) * A.c;
                float px = unit ? unit_wrap(qx, A.a) : qx, py = unit ? unit_wrap(qy, A.b) : qy;
                int cm = region_of(geom, meta, A.n_img, px, py, pz);
                if (cm != A.init_arg) continue;
                x = px; y = py; z = pz; comp = cm;
                x0 = qx; y0 = qy; z0 = pz;
                done = true;
            } else {
                float4 gk = geom[A.init_img];
                bool sk = (meta[A.init_img] & kSphereBit) != 0;
                float Rf = gk.w;
                float px = gk.x + (unif(o.x) * 2.0f - 1.0f) * Rf;
                float py = gk.y + (unif(o.y) * 2.0f - 1.0f) * Rf;
                float pz = gk.z + (unif(o.z) * 2.0f - 1.0f) * Rf;
                if (!inside_at(gk, sk, px, py, pz)) continue;
                px = wrap1(px, A.a); py = wrap1(py, A.b); pz = wrap1(pz, A.c);
                x = px; y = py; z = pz;
                x0 = px; y0 = py; z0 = pz;
                comp = region_of(geom, meta, A.n_img, px, py, pz);
                done = true;
            }
        }
    }
    if (!done) atomicOr(A.err, 1);
    fs[(size_t)F_X * N + w] = x;
    fs[(size_t)F_Y * N + w] = y;
    fs[(size_t)F_Z * N + w] = z;
    fs[(size_t)F_DX * N + w] = 0.0f;
    fs[(size_t)F_DY * N + w] = 0.0f;
    fs[(size_t)F_DZ * N + w] = 0.0f;
    fs[(size_t)F_X0 * N + w] = x0;
    fs[(size_t)F_Y0 * N + w] = y0;
    fs[(size_t)F_Z0 * N + w] = z0;
    fs[(size_t)F_E * N + w] = 0.0f;
    fs[(size_t)F_LF * N + w] = 0.0f;
    is[(size_t)F_COMP * N + w] = comp;
    is[(size_t)F_COMP0 * N + w] = comp;
    is[(size_t)F_LI * N + w] = 1;
    st[(size_t)F_NCOLL * N + w] = 0u;
    st[(size_t)F_NTRAN * N + w] = 0u;
    st[(size_t)F_NCAP * N + w] = 0u;
    is[(size_t)F_STEP * N + w] = 1;
    // transmission budget for the walk's first transmission (ordinal 0)
    fs[(size_t)F_BUD * N + w] = neg2log_unif(philox10(wid, 0u, kStreamColl, A.inst, A.k0, A.k1).x);
    for (int f = 0; f < n_mom; ++f) fs[(size_t)(F_MOM + f) * N + w] = 0.0f;
}

static __global__ __launch_bounds__(1024) void init_kernel(WalkArgs A, int n_mom) {
    extern __shared__ __align__(16) unsigned char lds[];
    float4* geom;
    int* meta;
    stage_images(A, lds, geom, meta);
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= A.N) return;
    init_lane(A, geom, meta, w, n_mom);
}

}  // namespace axx
