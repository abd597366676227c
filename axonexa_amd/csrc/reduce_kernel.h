// reduce_kernel.h — fp64 signal sums from the walker state.
//
// Reference: transformAndSum<COS_UNARY/SIN_UNARY> and
// transformAndSumTwoVectors<COS_EXP_BINARY/SIN_EXP_BINARY>
// (src/util/reduction.cuh:76-195, src/util/cudaVector.cu:143-172), one launch
// per waveform and transform with a racy zero-then-atomicAdd (defect D7).
// Here: one launch covers every waveform; each block reduces a fixed walker
// chunk for a group of waveforms with a wavefront butterfly then LDS, writes
// its fp64 partial to a fixed slot, and a second kernel adds the partials in
// chunk order — deterministic for any launch shape.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/axonexa.h"

namespace axx {

constexpr int kRedThreads = 256;
constexpr int kRedJB = 8;        // waveforms per block
constexpr int kRedVals = 6;      // cos, sin, cos^2, sin^2, intra cos, intra sin

struct RedArgs {
    const uint32_t* __restrict__ state;
    long long N;
    int P;                 // profile stride of the moments
    int M;
    int n_steps;
    int relaxation;
    int chunk;             // walkers per block
    int n_chunks;
    const float* __restrict__ rate;      // [n_comp]
    const int* __restrict__ wf_p;        // [M]
    const int* __restrict__ wf_ord;      // [M] waveforms in profile order: group gy reduces
                                         // wf_ord[8 gy .. 8 gy + 7]
    int ax;                              // moment axes stored (bit a: axis a)
    const double* __restrict__ wf_dir;   // [M][3]
    const double* __restrict__ wf_scale; // [M] gamma*dt*amplitude
    const double* __restrict__ S;        // [P] sum_i profile_p(t_i)
    double* __restrict__ partial;        // [M][n_chunks][6]
    unsigned long long* __restrict__ counters;  // [4]: coll, tran, cap, intra
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One block of kRedThreads reduces walker chunk chunk_id for waveform group
// gy (waveforms wf_ord[gy*kRedJB ...]); group 0 also adds the counters.  The
// waveforms of a group are taken in profile order, so a walker's moments are
// loaded once per profile of the group rather than once per waveform; each
// waveform's phase, sum and partial slot are the same whatever the grouping.
__device__ __forceinline__ void reduce_chunk(const RedArgs& R, const int chunk_id, const int gy) {
    __shared__ double sh[kRedThreads / 64][kRedJB * kRedVals];
    __shared__ unsigned long long shc[kRedThreads / 64][4];
    const int j0 = gy * kRedJB;
    const long long N = R.N;
    const float* fs = reinterpret_cast<const float*>(R.state);
    const int* is = reinterpret_cast<const int*>(R.state);
    double acc[kRedJB][kRedVals];
#pragma unroll
    for (int jj = 0; jj < kRedJB; ++jj)
#pragma unroll
        for (int v = 0; v < kRedVals; ++v) acc[jj][v] = 0.0;
    unsigned long long cc[4] = {0ull, 0ull, 0ull, 0ull};
    const long long wbeg = (long long)chunk_id * R.chunk;
    long long wend = wbeg + R.chunk;
    if (wend > N) wend = N;
    for (long long w = wbeg + threadIdx.x; w < wend; w += kRedThreads) {
        const int comp = is[AXX_F_COMP * N + w];
        const int li = is[AXX_F_LI * N + w];
        const double Ef = (double)fs[AXX_F_E * N + w] +
                          ((double)(R.n_steps - li) - (double)fs[AXX_F_LF * N + w]) * (double)R.rate[comp];
        const double wt = R.relaxation ? exp(-Ef) : 1.0;
        const bool intra = is[AXX_F_COMP0 * N + w] != 0;
        const double r0[3] = {(double)fs[AXX_F_X0 * N + w], (double)fs[AXX_F_Y0 * N + w], (double)fs[AXX_F_Z0 * N + w]};
        if (gy == 0) {
            cc[0] += R.state[AXX_F_NCOLL * N + w];
            cc[1] += R.state[AXX_F_NTRAN * N + w];
            cc[2] += R.state[AXX_F_NCAP * N + w];
            cc[3] += intra ? 1ull : 0ull;
        }
        // the moments of profile pm (axes not stored are zero in the state)
        int pm = -1;
        float mo[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int jj = 0; jj < kRedJB; ++jj) {
            if (j0 + jj >= R.M) break;
            const int j = R.wf_ord[j0 + jj];
            const int p = R.wf_p[j];
            if (p != pm) {
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    if ((R.ax >> a) & 1) mo[a] = fs[(AXX_F_MOM + 3 * p + a) * N + w];
                pm = p;
            }
            double ph = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double d = R.wf_dir[3 * j + a];
                if (d == 0.0) continue;
                ph += d * ((double)mo[a] + r0[a] * R.S[p]);
            }
            const double phi = R.wf_scale[j] * ph;
            double sv, cv;
            sincos(phi, &sv, &cv);
            const double cr = cv * wt, sr = sv * wt;
            acc[jj][0] += cr;
            acc[jj][1] += sr;
            acc[jj][2] += cr * cr;
            acc[jj][3] += sr * sr;
            if (intra) {
                acc[jj][4] += cr;
                acc[jj][5] += sr;
            }
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int jj = 0; jj < kRedJB; ++jj)
#pragma unroll
        for (int v = 0; v < kRedVals; ++v) {
            const double s = wave_sum(acc[jj][v]);
            if (lane == 0) sh[wave][jj * kRedVals + v] = s;
        }
    if (gy == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned long long s = wave_sum_u64(cc[k]);
            if (lane == 0) shc[wave][k] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < kRedJB * kRedVals) {
        const int jj = threadIdx.x / kRedVals, v = threadIdx.x % kRedVals;
        if (j0 + jj < R.M) {
            const int j = R.wf_ord[j0 + jj];
            double s = 0.0;
            for (int wv = 0; wv < kRedThreads / 64; ++wv) s += sh[wv][threadIdx.x];
            R.partial[((size_t)j * R.n_chunks + chunk_id) * kRedVals + v] = s;
        }
    }
    if (gy == 0 && threadIdx.x < 4) {
        unsigned long long s = 0;
        for (int wv = 0; wv < kRedThreads / 64; ++wv) s += shc[wv][threadIdx.x];
        atomicAdd(&R.counters[threadIdx.x], s);
    }
}

static __global__ __launch_bounds__(kRedThreads) void reduce_kernel(RedArgs R) {
    reduce_chunk(R, blockIdx.x, blockIdx.y);
}

// Sum the per-chunk partials in chunk order: one thread per (waveform, value).
static __global__ void finalize_kernel(const double* __restrict__ partial, int M, int n_chunks,
                                double* __restrict__ out /*[M][6]*/) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * kRedVals) return;
    const int j = t / kRedVals, v = t % kRedVals;
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += partial[((size_t)j * n_chunks + c) * kRedVals + v];
    out[t] = s;
}

static __global__ void zero_u64(unsigned long long* p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

}  // namespace axx
