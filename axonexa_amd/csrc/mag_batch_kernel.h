// mag_batch_kernel.h — many magnetization runs in one persistent launch:
// a grid sized to the chip takes (run, walker range) tiles from an atomic
// counter, restages the run's substrate tables in LDS when the run changes
// and walks the tile's walkers with that run's fields.  All runs of a batch
// have the same field count MB and integrator; geometry class and boundary
// policy are per run (a wave-uniform switch).
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

// Waves per SIMD the register allocation must allow.  The kernel inlines
// four walk variants (geometry class x boundary policy) with every step
// generator; left to itself the compiler takes 200-310 VGPRs (one or two
// waves).  Three waves (168 VGPRs, spills outside the step loop) measured
// faster for every field count and integrator except Dormand-Prince with one
// or two fields (rotation, 2 fields: +14 %; 8 fields: +43 %; Dormand-Prince,
// 8 fields: 2.2x; Dormand-Prince, 2 fields: -9 % at three, so two there).
template <int MB, int INTEG>
constexpr int mag_batch_waves() {
    return (INTEG == kMagDopri && MB <= 2) ? 2 : 3;
}

template <int MB, int INTEG>
__global__ __launch_bounds__(kMagBlock) __attribute__((amdgpu_waves_per_eu(mag_batch_waves<MB, INTEG>())))
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
        const int c = __builtin_amdgcn_readfirstlane(B.cls[i]);
        if (i != staged) {
            // every lane finished the previous tile (barriers above)
            T = (c & 1) ? stage_tables<true>(A, lds) : stage_tables<false>(A, lds);
            staged = i;
        }
        const MagArgs G = B.margs[i];
        for (int w = w0 + (int)threadIdx.x; w < w1; w += (int)blockDim.x) {
            MagAcc<MB, INTEG> acc;
            acc.init(G);
            switch (c) {
            case 0: walk_lane<MagAcc<MB, INTEG>, false, false, true>(A, T, acc, (uint32_t)w); break;
            case 1: walk_lane<MagAcc<MB, INTEG>, true, false, true>(A, T, acc, (uint32_t)w); break;
            case 2: walk_lane<MagAcc<MB, INTEG>, false, true, true>(A, T, acc, (uint32_t)w); break;
            default: walk_lane<MagAcc<MB, INTEG>, true, true, true>(A, T, acc, (uint32_t)w); break;
            }
        }
    }
}

typedef void (*mag_batch_fn)(MagBatchArgs);
mag_batch_fn pick_mag_batch_rotate(int mb);
mag_batch_fn pick_mag_batch_rk4(int mb);
mag_batch_fn pick_mag_batch_dopri(int mb);

}  // namespace axx
