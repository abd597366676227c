// mag_batch_kernel.h — many magnetization runs in one persistent launch:
// a grid sized to the chip takes (run, walker range) tiles from an atomic
// counter, restages the run's substrate tables in LDS when the run changes
// and walks the tile's walkers with that run's fields.  All runs of a batch
// have the same field count MB and integrator; geometry class, boundary
// policy and step generator are per run (a wave-uniform switch), or, when
// every run shares them with isotropic steps, fixed in the instantiation.
#pragma once
#include "mag_kernel.h"

namespace axx {

struct MagBatchArgs {
    const WalkArgs* __restrict__ args;   // [n_runs], steps [0, n_steps)
    const MagArgs* __restrict__ margs;   // [n_runs]
    const int* __restrict__ cls;         // [n_runs] general geometry | rejection << 1
    const int4* __restrict__ tiles;      // [n_tiles] (run, w0, w1, 0)
    int n_tiles;
    unsigned int* __restrict__ next;     // tile counter, zeroed before the launch
};

// CLS: -1 = runs of mixed geometry classes / boundary policies / step
// generators (a wave-uniform switch over four walk variants, every step
// generator compiled in); 0..3 = every run of the batch has class CLS
// (general geometry | rejection << 1) and isotropic steps: one walk variant.
//
// Waves per SIMD the register allocation must allow.  Mixed: left to itself
// the compiler takes 200-310 VGPRs (one or two waves); three waves (168
// VGPRs, spills outside the step loop) measured faster for every field count
// and integrator except Dormand-Prince with one or two fields, which keeps
// its own allocation.  Uniform: four waves (128 VGPRs): rotation, 2 fields
// +9 %, 8 fields +43 %; RK4 +14 %; Dormand-Prince +30-37 % over the mixed
// kernel on the same runs.
template <int MB, int INTEG, int CLS>
constexpr int mag_batch_waves() {
    return CLS >= 0 ? 4 : ((INTEG == kMagDopri && MB <= 2) ? 1 : 3);
}

template <int MB, int INTEG, int CLS>
__global__ __launch_bounds__(kMagBlock) __attribute__((amdgpu_waves_per_eu(mag_batch_waves<MB, INTEG, CLS>())))
void mag_batch_walk_kernel(MagBatchArgs B) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_tile;
    int staged = -1;
    LdsTables T{};
    for (;;) {
        if (threadIdx.x == 0) s_tile = (int)atomicAdd(B.next, 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane(s_tile);
        __syncthreads();
        if (t >= B.n_tiles) break;
        const int4 tile = B.tiles[t];
        const int i = __builtin_amdgcn_readfirstlane(tile.x);
        const int w0 = __builtin_amdgcn_readfirstlane(tile.y);
        const int w1 = __builtin_amdgcn_readfirstlane(tile.z);
        const WalkArgs A = B.args[i];
        const int c = CLS >= 0 ? CLS : __builtin_amdgcn_readfirstlane(B.cls[i]);
        if (i != staged) {
            // every lane finished the previous tile (barriers above)
            if constexpr (CLS >= 0)
                T = stage_tables<(CLS & 1) != 0>(A, lds);
            else
                T = (c & 1) ? stage_tables<true>(A, lds) : stage_tables<false>(A, lds);
            staged = i;
        }
        const MagArgs G = B.margs[i];
        for (int w = w0 + (int)threadIdx.x; w < w1; w += (int)blockDim.x) {
            MagAcc<MB, INTEG> acc;
            acc.init(G);
            if constexpr (CLS >= 0) {
                // spheres and planes (class bit 0) take lookup groups of two
                walk_lane<MagAcc<MB, INTEG>, (CLS & 1) != 0, (CLS & 2) != 0, false, (CLS & 1) != 0 ? 1 : 0>(
                    A, T, acc, (uint32_t)w);
            } else {
                switch (c) {
                case 0: walk_lane<MagAcc<MB, INTEG>, false, false, true, 0>(A, T, acc, (uint32_t)w); break;
                case 1: walk_lane<MagAcc<MB, INTEG>, true, false, true, 1>(A, T, acc, (uint32_t)w); break;
                case 2: walk_lane<MagAcc<MB, INTEG>, false, true, true, 0>(A, T, acc, (uint32_t)w); break;
                default: walk_lane<MagAcc<MB, INTEG>, true, true, true, 1>(A, T, acc, (uint32_t)w); break;
                }
            }
        }
    }
}

typedef void (*mag_batch_fn)(MagBatchArgs);
mag_batch_fn pick_mag_batch_rotate(int mb, int cls);
mag_batch_fn pick_mag_batch_rk4(int mb, int cls);
mag_batch_fn pick_mag_batch_dopri(int mb, int cls);

}  // namespace axx
