// mag_internal.h — host objects and shared device code of the magnetization
// path (mag.hip: single runs; mag_batch.hip: many runs in one launch).
#pragma once
#include "internal.h"
#include "mag_kernel.h"

namespace axx {

constexpr int kMagChunk = 4096;
constexpr int kMagRedThreads = 256;
// LDS for the substrate tables of one 256-thread workgroup: four workgroups
// (one wave per SIMD each) share a CU's 160 KB.
constexpr int kMagLdsBudget = 40 * 1024;
constexpr int kMagTaskSteps = 128;   // automatic steps per walk task of a magnetization run

// Sums of Mx, My, Mz of field j over the walkers of chunk c, in a fixed order
// (strided per thread, then a fixed tree), into partial[(j * n_chunks + c) * 3];
// with cnt, field 0 also adds the event counters.  One block of
// kMagRedThreads threads.
__device__ __forceinline__ void mag_reduce_block(const uint32_t* __restrict__ state, long long N, int n_chunks,
                                                 int c, int j, double* __restrict__ partial,
                                                 unsigned long long* __restrict__ cnt) {
    __shared__ double sh[3][kMagRedThreads];
    __shared__ unsigned long long shc[3][kMagRedThreads];
    const int t = threadIdx.x;
    const long long w0 = (long long)c * kMagChunk;
    const long long w1 = w0 + kMagChunk < N ? w0 + kMagChunk : N;
    double s[3] = {0.0, 0.0, 0.0};
    unsigned long long e[3] = {0ull, 0ull, 0ull};
    for (long long w = w0 + t; w < w1; w += kMagRedThreads) {
        for (int k = 0; k < 3; ++k) {
            const uint32_t lo = state[(size_t)(F_MOM + 6 * j + 2 * k) * N + w];
            const uint32_t hi = state[(size_t)(F_MOM + 6 * j + 2 * k + 1) * N + w];
            s[k] += __hiloint2double((int)hi, (int)lo);
        }
        if (j == 0 && cnt) {
            e[0] += state[(size_t)F_NCOLL * N + w];
            e[1] += state[(size_t)F_NTRAN * N + w];
            e[2] += state[(size_t)F_NCAP * N + w];
        }
    }
    for (int k = 0; k < 3; ++k) {
        sh[k][t] = s[k];
        shc[k][t] = e[k];
    }
    __syncthreads();
    for (int h = kMagRedThreads / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < 3; ++k) {
                sh[k][t] += sh[k][t + h];
                shc[k][t] += shc[k][t + h];
            }
        __syncthreads();
    }
    if (t < 3) {
        partial[((size_t)j * n_chunks + c) * 3 + t] = sh[t][0];
        if (j == 0 && cnt) atomicAdd(&cnt[t], shc[t][0]);
    }
}

// Chunk-ordered total of field j, component k over n_chunks partials.
__device__ __forceinline__ double mag_chunk_total(const double* __restrict__ partial, int n_chunks, int j, int k) {
    double s = 0.0;
    for (int c = 0; c < n_chunks; ++c) s += partial[((size_t)j * n_chunks + c) * 3 + k];
    return s;
}

}  // namespace axx

struct axx_mag {
    axx_ctx* ctx = nullptr;
    int n_fields = 0, integ = 0, n_steps = 0;
    long long N = 0;
    std::vector<axx_plan*> plans;    // one per group
    std::vector<int> gF, gMB;
    std::vector<double*> d_tab;
    int* d_off = nullptr;
    int* d_nsub = nullptr;
    double* d_iT1 = nullptr;
    double* d_iT2 = nullptr;
    double* d_part = nullptr;        // [group][F][n_chunks][3]
    double* d_out = nullptr;         // [n_fields][3]
    unsigned long long* d_cnt = nullptr;
    MagArgs base{};
    int n_chunks = 0;
    int next_step = 0;
    bool reduced = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};


void mag_free(axx_mag* m);
// Builds a magnetization run (one plan per group of fields); lds_budget 0:
// the plan default.  Returns an error code; *out is set on success only.
int mag_create_internal(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                        int lds_budget, axx_mag** out);
