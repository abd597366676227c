// mag_batch.hip — C ABI of batched magnetization runs: every acquisition of
// a magAcquisitionStream<F> (src/acquisition/magAcquisitionStream.cuh:
// 105-342, which runs them two at a time on CUDA streams) in one persistent
// launch.  Each run keeps its own plan, seed, instance id and field table,
// so its walkers and sums are bitwise those of its own axx_run_mag.
#include "mag_internal.h"
#include "init_kernel.h"
#include "mag_batch_kernel.h"

namespace axx {

struct MagRed {
    const uint32_t* state;
    long long N;
    int n_chunks;
    double* partial;
    double* out;
    unsigned long long* cnt;
};

// Tile counter and every run's event counters to zero (a kernel, so the
// sequence can be captured or replayed without memset nodes).
static __global__ void mag_batch_zero_kernel(unsigned int* next, const MagRed* __restrict__ red, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *next = 0u;
    if (t < 4 * n) red[t >> 2].cnt[t & 3] = 0ull;
}

// Placement and M = m0, one block per tile.
static __global__ __launch_bounds__(kMagBlock) void mag_batch_init_kernel(const WalkArgs* __restrict__ args,
                                                                          const MagArgs* __restrict__ margs,
                                                                          int mb, const int4* __restrict__ tiles) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int4 tile = tiles[blockIdx.x];
    const int i = __builtin_amdgcn_readfirstlane(tile.x);
    const WalkArgs A = args[i];
    const MagArgs G = margs[i];
    float4* geom;
    int* meta;
    stage_images(A, lds, geom, meta);
    const double v[3] = {G.m0x, G.m0y, G.m0z};
    for (int w = tile.y + (int)threadIdx.x; w < tile.z; w += (int)blockDim.x) {
        init_lane(A, geom, meta, w, 0);
        for (int j = 0; j < mb; ++j)
            for (int c = 0; c < 3; ++c) {
                const unsigned long long b = (unsigned long long)__double_as_longlong(v[c]);
                A.state[(size_t)(F_MOM + 6 * j + 2 * c) * A.N + w] = (uint32_t)b;
                A.state[(size_t)(F_MOM + 6 * j + 2 * c + 1) * A.N + w] = (uint32_t)(b >> 32);
            }
    }
}

// One block per (run, chunk) and field: the single-run reduction's blocks.
static __global__ __launch_bounds__(kMagRedThreads) void mag_batch_reduce_kernel(const MagRed* __restrict__ red,
                                                                                 const int2* __restrict__ blocks) {
    const int2 b = blocks[blockIdx.x];
    const MagRed R = red[__builtin_amdgcn_readfirstlane(b.x)];
    mag_reduce_block(R.state, R.N, R.n_chunks, __builtin_amdgcn_readfirstlane(b.y), blockIdx.y, R.partial, R.cnt);
}

// Chunk-ordered totals, one block per run.
static __global__ void mag_batch_finalize_kernel(const MagRed* __restrict__ red, int F) {
    const MagRed R = red[blockIdx.x];
    for (int t = threadIdx.x; t < F * 3; t += blockDim.x) R.out[t] = mag_chunk_total(R.partial, R.n_chunks, t / 3, t % 3);
}

static mag_batch_fn pick_batch(int integ, int mb, int cls) {
    if (integ == kMagRotate) return pick_mag_batch_rotate(mb, cls);
    if (integ == kMagRK4) return pick_mag_batch_rk4(mb, cls);
    return pick_mag_batch_dopri(mb, cls);
}

}  // namespace axx

struct axx_mag_batch {
    axx_ctx* ctx = nullptr;
    std::vector<axx_mag*> runs;
    int F = 0, integ = 0;
    WalkArgs* d_args = nullptr;
    MagArgs* d_margs = nullptr;
    int* d_cls = nullptr;
    int4* d_tiles = nullptr;
    unsigned int* d_next = nullptr;
    MagRed* d_red = nullptr;
    int2* d_red_blocks = nullptr;
    int n_tiles = 0, n_red = 0, grid = 0;
    int cls = -1;   // the class all runs share with isotropic steps, else -1
    size_t lds_walk = 0, lds_init = 0;
    bool ran = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void mag_batch_free(axx_mag_batch* b) {
    if (!b) return;
    void* bufs[] = {b->d_args, b->d_margs, b->d_cls, b->d_tiles, b->d_next, b->d_red, b->d_red_blocks};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (axx_mag* m : b->runs) mag_free(m);
    delete b;
}

template <typename T>
int put(axx_ctx* ctx, T** dst, const std::vector<T>& v) {
    HIP_TRY(ctx, hipMalloc((void**)dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
    if (!v.empty()) HIP_TRY(ctx, hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return AXX_OK;
}

}  // namespace

extern "C" void axx_mag_batch_destroy(axx_mag_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    mag_batch_free(b);
}

extern "C" int axx_mag_batch_create(axx_ctx* ctx, int n_inst, const axx_mag_instance* inst, axx_mag_batch** out) {
    axx_mag_batch* b = nullptr;   // freed by the handler if construction throws
    try {
        if (!ctx || !inst || !out || n_inst < 1) return fail(ctx, AXX_ERR_INVALID, "mag batch: bad arguments");
        *out = nullptr;
        for (int i = 0; i < n_inst; ++i) {
            const axx_mag_instance& I = inst[i];
            if (!I.substrate || !I.field || !I.run) return fail(ctx, AXX_ERR_INVALID, "run %d: null descriptor", i);
            if (I.field->n_fields < 1 || I.field->n_fields > AXX_MAG_GROUP)
                return fail(ctx, AXX_ERR_LIMIT, "run %d: batched runs carry 1 to %d fields", i, AXX_MAG_GROUP);
            if (I.field->n_fields != inst[0].field->n_fields || I.field->integrator != inst[0].field->integrator)
                return fail(ctx, AXX_ERR_INVALID, "run %d: every run of a batch needs the same field count and "
                                                  "integrator", i);
            if (I.run->n_walkers >= (1ll << 31)) return fail(ctx, AXX_ERR_LIMIT, "run %d: n_walkers >= 2^31", i);
        }
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        b = new (std::nothrow) axx_mag_batch();
        if (!b) return AXX_ERR_NOMEM;
        b->ctx = ctx;
        b->F = inst[0].field->n_fields;
        b->integ = inst[0].field->integrator;
        b->runs.reserve((size_t)n_inst);
        std::vector<WalkArgs> args;
        std::vector<MagArgs> margs;
        std::vector<int> cls;
        std::vector<MagRed> red;
        std::vector<int2> rblocks;
        struct Tile {
            double cost;
            int run, w0, w1;
        };
        std::vector<Tile> tiles;
        for (int i = 0; i < n_inst; ++i) {
            axx_mag* m = nullptr;
            const int rc = mag_create_internal(ctx, inst[i].substrate, inst[i].field, inst[i].run, kMagLdsBudget, &m);
            if (rc) {
                std::string msg = "run " + std::to_string(i) + ": " + ctx->err;
                mag_batch_free(b);
                return fail(ctx, rc, "%s", msg.c_str());
            }
            b->runs.push_back(m);   // capacity reserved: does not throw
            axx_plan* p = m->plans[0];
            args.push_back(plan_walk_args(p, 0, p->n_steps));
            MagArgs G = m->base;
            G.tab = m->d_tab[0];
            G.F = m->gF[0];
            margs.push_back(G);
            cls.push_back((p->gen ? 1 : 0) | (p->run.boundary_mode == AXX_BOUNDARY_REJECTION ? 2 : 0));
            const int ci = p->run.step_mode == AXX_STEP_ISOTROPIC ? cls.back() : -1;
            b->cls = i == 0 ? ci : (ci == b->cls ? ci : -1);
            red.push_back(MagRed{p->d_state, p->N, m->n_chunks, m->d_part, m->d_out, m->d_cnt});
            for (int c = 0; c < m->n_chunks; ++c) rblocks.push_back(make_int2(i, c));
            for (long long w0 = 0; w0 < p->N; w0 += kMagBlock) {
                const long long w1 = std::min<long long>(p->N, w0 + kMagBlock);
                tiles.push_back({(double)(w1 - w0) * p->n_steps, i, (int)w0, (int)w1});
            }
            b->lds_walk = std::max(b->lds_walk, p->lds_bytes);
            b->lds_init = std::max(b->lds_init, p->init_lds);
        }
        // longest tiles first; equal costs keep run then walker order
        std::stable_sort(tiles.begin(), tiles.end(), [](const Tile& x, const Tile& y) { return x.cost > y.cost; });
        std::vector<int4> ht(tiles.size());
        for (size_t k = 0; k < tiles.size(); ++k) ht[k] = make_int4(tiles[k].run, tiles[k].w0, tiles[k].w1, 0);
        b->n_tiles = (int)ht.size();
        b->n_red = (int)rblocks.size();
        int rc;
        if ((rc = put(ctx, &b->d_args, args)) || (rc = put(ctx, &b->d_margs, margs)) || (rc = put(ctx, &b->d_cls, cls)) ||
            (rc = put(ctx, &b->d_tiles, ht)) || (rc = put(ctx, &b->d_red, red)) ||
            (rc = put(ctx, &b->d_red_blocks, rblocks))) {
            mag_batch_free(b);
            return rc;
        }
        if (hipMalloc((void**)&b->d_next, sizeof(unsigned int)) != hipSuccess ||
            hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
            mag_batch_free(b);
            return fail(ctx, AXX_ERR_DEVICE, "mag batch: device allocation");
        }
        if (b->lds_walk > ctx->lds_max || b->lds_init > ctx->lds_max) {
            mag_batch_free(b);
            return fail(ctx, AXX_ERR_LIMIT, "mag batch: substrate tables exceed LDS");
        }
        int per_cu = 0;
        const mag_batch_fn fn = pick_batch(b->integ, b->F, b->cls);
        HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kMagBlock, b->lds_walk));
        if (per_cu < 1) {
            mag_batch_free(b);
            return fail(ctx, AXX_ERR_LIMIT, "mag batch kernel does not fit a CU");
        }
        b->grid = std::max(1, std::min(per_cu * std::max(ctx->n_cu, 1), b->n_tiles));
        *out = b;
        return AXX_OK;
    } catch (...) {
        if (b) mag_batch_free(b);
        if (out) *out = nullptr;
        return api_exception();
    }
}

extern "C" int axx_mag_batch_run(axx_mag_batch* b) {
    try {
        if (!b) return AXX_ERR_INVALID;
        axx_ctx* ctx = b->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const int n = (int)b->runs.size();
        hipLaunchKernelGGL(mag_batch_zero_kernel, dim3((4 * n + 255) / 256), dim3(256), 0, ctx->stream, b->d_next,
                           (const MagRed*)b->d_red, n);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(mag_batch_init_kernel, dim3(b->n_tiles), dim3(kMagBlock), b->lds_init, ctx->stream,
                           (const WalkArgs*)b->d_args, (const MagArgs*)b->d_margs, b->F, (const int4*)b->d_tiles);
        HIP_TRY(ctx, hipGetLastError());
        MagBatchArgs B{b->d_args, b->d_margs, b->d_cls, b->d_tiles, b->n_tiles, b->d_next};
        hipLaunchKernelGGL(pick_batch(b->integ, b->F, b->cls), dim3(b->grid), dim3(kMagBlock), b->lds_walk, ctx->stream, B);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(mag_batch_reduce_kernel, dim3(b->n_red, b->F), dim3(kMagRedThreads), 0, ctx->stream,
                           (const MagRed*)b->d_red, (const int2*)b->d_red_blocks);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(mag_batch_finalize_kernel, dim3(n), dim3(64), 0, ctx->stream, (const MagRed*)b->d_red,
                           b->F);
        HIP_TRY(ctx, hipGetLastError());
        for (axx_mag* m : b->runs) {
            m->next_step = m->n_steps;
            m->reduced = true;
        }
        b->ran = true;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_batch_time(axx_mag_batch* b, float* ms) {
    try {
        if (!b || !ms) return AXX_ERR_INVALID;
        axx_ctx* ctx = b->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipEventRecord(b->ev0, ctx->stream));
        const int rc = axx_mag_batch_run(b);
        if (rc) return rc;
        HIP_TRY(ctx, hipEventRecord(b->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(b->ev1));
        HIP_TRY(ctx, hipEventElapsedTime(ms, b->ev0, b->ev1));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_batch_results(axx_mag_batch* b, int i, axx_mag_result* out) {
    try {
        if (!b || !out || i < 0 || i >= (int)b->runs.size())
            return fail(b ? b->ctx : nullptr, AXX_ERR_INVALID, "mag batch results: bad index");
        if (!b->ran) return fail(b->ctx, AXX_ERR_STATE, "mag batch results requested before a run");
        return axx_mag_results(b->runs[i], out);
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_batch_get_state(axx_mag_batch* b, int i, uint32_t* host_state) {
    try {
        if (!b || !host_state || i < 0 || i >= (int)b->runs.size())
            return fail(b ? b->ctx : nullptr, AXX_ERR_INVALID, "mag batch state: bad index");
        return axx_mag_get_state(b->runs[i], 0, host_state);
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_run_mag_batch(axx_ctx* ctx, int n_inst, const axx_mag_instance* inst, axx_mag_result* out) {
    try {
        if (!out) return fail(ctx, AXX_ERR_INVALID, "null results");
        axx_mag_batch* b = nullptr;
        int rc = axx_mag_batch_create(ctx, n_inst, inst, &b);
        if (rc) return rc;
        rc = axx_mag_batch_run(b);
        for (int i = 0; !rc && i < n_inst; ++i) rc = axx_mag_batch_results(b, i, &out[i]);
        axx_mag_batch_destroy(b);
        return rc;
    } catch (...) {
        return api_exception();
    }
}
