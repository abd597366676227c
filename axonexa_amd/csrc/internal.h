// internal.h — host-side objects shared by the C ABI translation units
// (capi.hip: contexts and plans; batch.hip: batched runs).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/axonexa.h"
#include "reduce_kernel.h"
#include "walk_kernel.h"

using namespace axx;

struct axx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    size_t lds_max = 65536;
    unsigned char* scratch = nullptr;   // this device's scratch array (allocation-free runs)
    std::string err;
};

// Bump allocator over the scratch array of an allocation-free run.
struct DevArena {
    unsigned char* base;
    size_t size, used;
};

struct axx_plan {
    axx_ctx* ctx = nullptr;
    // host copies
    int n_steps = 0, P = 0, Pt = 0, M = 0, ax = 7, n_prims = 0;
    bool gen = false;       // spheres or planes present: general kernel path
    long long N = 0;
    axx_run run{};
    double dt = 0;
    std::vector<double> S, wf_scale, wf_dir;
    std::vector<int> wf_p;
    std::vector<int> wf_ord;   // waveforms ordered by profile (stable): the reduction's grouping
    // geometry sizes
    int n_img = 0, n_bins = 0, n_cand = 0, nb[3] = {1, 1, 1};
    float ib[3] = {1, 1, 1}, ib4[2] = {4, 4};
    size_t init_lds = 0;
    double bin_width = 0;
    size_t lds_bytes = 0;
    int init_img = 0;
    int block = 256, wpt = 1;
    int next_step = 0;
    bool reduced = false;
    bool arena = false;     // device buffers are pieces of the scratch array (not freed)
    // device buffers
    float4* d_geom = nullptr;
    int* d_meta = nullptr;
    float4* d_cyl = nullptr;
    float* d_inv = nullptr;
    float2* d_thr = nullptr;
    uint4* d_bins = nullptr;
    uint16_t* d_cand = nullptr;
    float* d_L = nullptr;
    float* d_rate = nullptr;
    float* d_prof = nullptr;
    uint32_t* d_state = nullptr;
    int* d_err = nullptr;
    int* d_wfp = nullptr;
    int* d_wford = nullptr;
    double* d_wfdir = nullptr;
    double* d_wfscale = nullptr;
    double* d_S = nullptr;
    double* d_partial = nullptr;
    double* d_out = nullptr;
    unsigned long long* d_cnt = nullptr;
    int n_chunks = 0, chunk = 0;
    unsigned int* d_queue = nullptr;   // walk task counter + chunks finished per group of 64 walkers
    int sched = 0;                     // steps per walk task: 0 automatic, -1 the whole walk, > 0 that many
    int auto_task = 0;                 // automatic steps per task once chunks pay off (0: the walk's rule)
    int last_task_steps = 0;           // steps per task of the last walk launch
    int last_grid = 0;                 // workgroups of the last walk launch
    float cell[3] = {0, 0, 0};      // cell the positions wrap in (unit cell of a repeating substrate)
    float whole[2] = {0, 0};        // the substrate's cell along x and y (placement)
    int units = 1;                  // unit cells in the substrate's cell
    float H2c = 0, H2s = 0, plane_eps = 0;
    int gmask = 1;          // steps per cell-bin lookup - 1 (1, 3 or 7)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

static int fail(axx_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

// Every int entry point of the C ABI turns a C++ exception (in practice a
// failed host allocation for an oversized input) into an error code instead
// of letting it cross the C boundary.
static inline int api_exception() {
    try {
        throw;
    } catch (const std::bad_alloc&) {
        return AXX_ERR_NOMEM;
    } catch (...) {
        return AXX_ERR_INVALID;
    }
}

#define HIP_TRY(ctx, call)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((ctx), AXX_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)


// Plan internals used by the batched runs.  min_pt: smallest padded
// profile width to allocate moments for (1 for a plain plan).
int plan_create_impl(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q, const axx_run* r,
                     const axx_tuning* tune, axx_plan* p, int min_pt, DevArena* arena = nullptr);
void plan_free(axx_plan* p);
WalkArgs plan_walk_args(const axx_plan* p, int i0, int i1);
int plan_tasks(axx_plan* p, const void* fn, int block, int i0, int i1, WalkQueue* Q, unsigned* blocks);
RedArgs plan_red_args(const axx_plan* p);
