// mag.hip — C ABI of the magnetization (Bloch) path: magAcquisitionStream<F>
// ::runAcquisition (src/acquisition/magAcquisitionStream.cuh:37-103) over
// the walk core of the phase path.  One plan per group of up to eight fields
// (its own walker state, same walkers); field tables, substep offsets and
// per-compartment relaxation rates on the device; deterministic fp64 sums.
#include "mag_internal.h"
#include "init_kernel.h"

namespace axx {

// M = m0 for every field slot of every walker (after placement).
static __global__ void mag_fill_kernel(uint32_t* __restrict__ state, long long N, int mb, double m0x, double m0y,
                                       double m0z) {
    const long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= N) return;
    const double v[3] = {m0x, m0y, m0z};
    for (int j = 0; j < mb; ++j)
        for (int c = 0; c < 3; ++c) {
            const unsigned long long b = (unsigned long long)__double_as_longlong(v[c]);
            state[(size_t)(F_MOM + 6 * j + 2 * c) * N + w] = (uint32_t)b;
            state[(size_t)(F_MOM + 6 * j + 2 * c + 1) * N + w] = (uint32_t)(b >> 32);
        }
}

static __global__ void mag_zero_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// Block (chunk c, field j): see mag_reduce_block.
static __global__ __launch_bounds__(kMagRedThreads) void mag_reduce_kernel(const uint32_t* __restrict__ state,
                                                                           long long N, int n_chunks,
                                                                           double* __restrict__ partial,
                                                                           unsigned long long* __restrict__ cnt) {
    mag_reduce_block(state, N, n_chunks, blockIdx.x, blockIdx.y, partial, cnt);
}

// Chunk-ordered totals: out[j * 3 + k].
static __global__ void mag_finalize_kernel(const double* __restrict__ partial, int F, int n_chunks,
                                           double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= F * 3) return;
    out[t] = mag_chunk_total(partial, n_chunks, t / 3, t % 3);
}

}  // namespace axx

void mag_free(axx_mag* m) {
    if (!m) return;
    for (axx_plan* p : m->plans) plan_free(p);
    for (double* t : m->d_tab)
        if (t) (void)hipFree(t);
    void* bufs[] = {m->d_off, m->d_nsub, m->d_iT1, m->d_iT2, m->d_part, m->d_out, m->d_cnt};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    delete m;
}

template <class T>
static int upload(axx_ctx* ctx, T** d, const T* h, size_t n) {
    if (hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        return fail(ctx, AXX_ERR_NOMEM, "device allocation (%zu bytes)", n * sizeof(T));
    if (h && n) HIP_TRY(ctx, hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice));
    return AXX_OK;
}

static int mag_offsets(int integ, int n_steps, const int* sub, std::vector<int>& off, std::vector<int>& ns,
                       long long& total) {
    off.assign((size_t)n_steps + 1, 0);
    ns.assign((size_t)n_steps, 1);
    total = 0;
    for (int i = 0; i < n_steps; ++i) {
        const int s = sub ? sub[i] : 1;
        if (s < 1 || s > 4096) return AXX_ERR_INVALID;
        ns[i] = s;
        off[i] = (int)total;
        total += mag_samples(integ, s);
        if (total > 0x7fffffffLL) return AXX_ERR_LIMIT;
    }
    off[n_steps] = (int)total;
    return AXX_OK;
}

extern "C" int axx_mag_sample_count(int integrator, int n_steps, const int* substeps, int64_t* n_samples) {
    try {
        if (!n_samples || n_steps < 1 || integrator < 0 || integrator > 2) return AXX_ERR_INVALID;
        std::vector<int> off, ns;
        long long total = 0;
        const int rc = mag_offsets(integrator, n_steps, substeps, off, ns, total);
        if (rc) return rc;
        *n_samples = total;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

int mag_create_internal(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                        int lds_budget, axx_mag** out) {
    axx_mag* m = nullptr;   // freed by the handler if construction throws
    axx_plan* p_new = nullptr;
    try {
        if (!ctx || !s || !f || !r || !out) return fail(ctx, AXX_ERR_INVALID, "null argument");
        *out = nullptr;
        if (s->n_prims < 0 || s->n_prims > AXX_MAX_PRIMS)
            return fail(ctx, AXX_ERR_INVALID, "n_prims must be in [0, %d]", AXX_MAX_PRIMS);
        if (f->n_fields > 4096) return fail(ctx, AXX_ERR_LIMIT, "field: n_fields > 4096");
        if (f->n_steps < 1 || !(f->dt > 0) || !std::isfinite(f->dt) || !std::isfinite(f->gamma))
            return fail(ctx, AXX_ERR_INVALID, "field: n_steps, dt, gamma");
        if (f->n_fields < 1) return fail(ctx, AXX_ERR_INVALID, "field: n_fields < 1");
        if (f->integrator < 0 || f->integrator > 2) return fail(ctx, AXX_ERR_INVALID, "field: integrator");
        if (!f->samples) return fail(ctx, AXX_ERR_INVALID, "field: samples missing");
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(f->origin[a]) || !std::isfinite(f->m0[a]))
                return fail(ctx, AXX_ERR_INVALID, "field: non-finite origin or m0");
        std::vector<int> off, ns;
        long long total = 0;
        int rc = mag_offsets(f->integrator, f->n_steps, f->substeps, off, ns, total);
        if (rc == AXX_ERR_INVALID) return fail(ctx, rc, "field: substeps must be in [1, 4096]");
        if (rc) return fail(ctx, rc, "field: more than 2^31 samples");
        const size_t n_rec = (size_t)total * f->n_fields * 12;
        for (size_t k = 0; k < n_rec; ++k)
            if (!std::isfinite(f->samples[k])) return fail(ctx, AXX_ERR_INVALID, "field: non-finite sample");
        std::vector<double> iT1((size_t)s->n_prims + 1, 0.0), iT2((size_t)s->n_prims + 1, 0.0);
        for (int k = 0; k <= s->n_prims; ++k) {
            if (f->T1) {
                const double t1 = f->T1[k];
                if (std::isnan(t1) || t1 <= 0) return fail(ctx, AXX_ERR_INVALID, "field: T1 must be > 0");
                iT1[k] = std::isfinite(t1) ? 1.0 / t1 : 0.0;
            }
            const double t2 = k == 0 ? s->T2_lat : (s->T2 ? s->T2[k - 1] : INFINITY);
            iT2[k] = (t2 > 0 && std::isfinite(t2)) ? 1.0 / t2 : 0.0;
        }

        m = new (std::nothrow) axx_mag();
        if (!m) return AXX_ERR_NOMEM;
        m->ctx = ctx;
        m->n_fields = f->n_fields;
        m->integ = f->integrator;
        m->n_steps = f->n_steps;
        m->N = r->n_walkers;
        // the walk core sees a one-profile acquisition whose moment area holds
        // 6 words per field slot
        std::vector<float> prof((size_t)f->n_steps, 0.0f);
        int wfp = 0;
        double amp = 0.0, dir[3] = {1.0, 0.0, 0.0};
        axx_acq q{};
        q.n_steps = f->n_steps; q.dt = f->dt; q.gamma = f->gamma;
        q.n_profiles = 1; q.profiles = prof.data();
        q.n_waveforms = 1; q.wf_profile = &wfp; q.wf_amplitude = &amp; q.wf_dir = dir;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const int n_groups = (f->n_fields + kMagGroup - 1) / kMagGroup;
        m->plans.reserve(n_groups);
        m->gF.reserve(n_groups);
        m->gMB.reserve(n_groups);
        m->d_tab.reserve(n_groups);
        for (int g = 0; g < n_groups; ++g) {
            const int F = std::min(kMagGroup, f->n_fields - g * kMagGroup);
            const int mb = F;
            axx_plan* p = new (std::nothrow) axx_plan();
            if (!p) {
                mag_free(m);
                return AXX_ERR_NOMEM;
            }
            p->ctx = ctx;
            p_new = p;
            axx_tuning t{};
            t.block = kMagBlock;
            t.lds_budget = lds_budget;
            rc = plan_create_impl(ctx, s, &q, r, &t, p, 2 * mb);
            if (rc) {
                plan_free(p);
                mag_free(m);
                return rc;
            }
            // tasks of 128 steps once the walker groups fill the resident
            // waves twice (low-field workload, 917,504 walkers x 1100 steps,
            // field-updates/s against whole walks: rotation 2 fields +8 %,
            // RK4 2 fields +6 %, Dormand-Prince 2 fields +14 %, RK4 8 fields
            // +4 %, rotation 4 and 8 fields -1 and -2 %)
            p->auto_task = kMagTaskSteps;
            m->plans.push_back(p);   // capacity reserved: does not throw
            p_new = nullptr;
            m->gF.push_back(F);
            m->gMB.push_back(mb);
            std::vector<double> tab((size_t)total * F * 12);
            for (long long k = 0; k < total; ++k)
                for (int j = 0; j < F; ++j)
                    std::memcpy(&tab[((size_t)k * F + j) * 12],
                                &f->samples[((size_t)k * f->n_fields + g * kMagGroup + j) * 12], 12 * sizeof(double));
            double* d = nullptr;
            if ((rc = upload(ctx, &d, tab.data(), tab.size()))) {
                mag_free(m);
                return rc;
            }
            m->d_tab.push_back(d);
        }
        m->n_chunks = (int)((m->N + kMagChunk - 1) / kMagChunk);
        if ((rc = upload(ctx, &m->d_off, off.data(), off.size())) || (rc = upload(ctx, &m->d_nsub, ns.data(), ns.size())) ||
            (rc = upload(ctx, &m->d_iT1, iT1.data(), iT1.size())) || (rc = upload(ctx, &m->d_iT2, iT2.data(), iT2.size())) ||
            (rc = upload<double>(ctx, &m->d_part, nullptr, (size_t)n_groups * kMagGroup * m->n_chunks * 3)) ||
            (rc = upload<double>(ctx, &m->d_out, nullptr, (size_t)f->n_fields * 3)) ||
            (rc = upload<unsigned long long>(ctx, &m->d_cnt, nullptr, 4))) {
            mag_free(m);
            return rc;
        }
        if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess) {
            mag_free(m);
            return fail(ctx, AXX_ERR_DEVICE, "event creation");
        }
        MagArgs& G = m->base;
        G.off = m->d_off; G.nsub = m->d_nsub; G.iT1 = m->d_iT1; G.iT2 = m->d_iT2;
        G.dt = f->dt; G.gamma = f->gamma;
        G.ox = f->origin[0]; G.oy = f->origin[1]; G.oz = f->origin[2];
        G.m0x = f->m0[0]; G.m0y = f->m0[1]; G.m0z = f->m0[2];
        G.relax = r->relaxation != 0 && f->integrator != AXX_MAG_ROTATE;
        *out = m;
        return AXX_OK;
    } catch (...) {
        if (p_new) plan_free(p_new);
        if (m && !(out && *out == m)) mag_free(m);
        if (out) *out = nullptr;
        return api_exception();
    }
}

extern "C" int axx_mag_create(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                              axx_mag** out) {
    return mag_create_internal(ctx, s, f, r, kMagLdsBudget, out);
}

static mag_fn pick_mag(int integ, int mb, bool gen, bool rej, bool alt) {
    if (integ == kMagRotate) return pick_mag_rotate(mb, gen, rej, alt);
    if (integ == kMagRK4) return pick_mag_rk4(mb, gen, rej, alt);
    return pick_mag_dopri(mb, gen, rej, alt);
}

extern "C" int axx_mag_set_schedule(axx_mag* m, int task_steps) {
    try {
        if (!m) return AXX_ERR_INVALID;
        if (task_steps < -1) return fail(m->ctx, AXX_ERR_INVALID, "task_steps must be -1, 0 or positive");
        for (axx_plan* p : m->plans) p->sched = task_steps;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_walk(axx_mag* m, int step_begin, int step_end) {
    try {
        if (!m) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        if (step_begin < 0 || step_end > m->n_steps || step_begin > step_end)
            return fail(ctx, AXX_ERR_INVALID, "step range [%d, %d) outside [0, %d)", step_begin, step_end, m->n_steps);
        if (step_begin != 0 && step_begin != m->next_step)
            return fail(ctx, AXX_ERR_STATE, "walk must restart at 0 or continue at step %d", m->next_step);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        for (size_t g = 0; g < m->plans.size(); ++g) {
            axx_plan* p = m->plans[g];
            WalkArgs A = plan_walk_args(p, step_begin, step_end);
            MagArgs G = m->base;
            G.tab = m->d_tab[g];
            G.F = m->gF[g];
            const long long grid = (p->N + kMagBlock - 1) / kMagBlock;
            if (grid > 0x7fffffff) return fail(ctx, AXX_ERR_LIMIT, "grid too large");
            if (step_begin == 0) {
                hipLaunchKernelGGL(init_kernel, dim3((unsigned)grid), dim3(kMagBlock), p->init_lds, ctx->stream, A, 0);
                HIP_TRY(ctx, hipGetLastError());
                hipLaunchKernelGGL(mag_fill_kernel, dim3((unsigned)grid), dim3(kMagBlock), 0, ctx->stream, p->d_state,
                                   p->N, m->gMB[g], G.m0x, G.m0y, G.m0z);
                HIP_TRY(ctx, hipGetLastError());
            }
            if (step_begin < step_end) {
                mag_fn fn = pick_mag(m->integ, m->gMB[g], p->gen, p->run.boundary_mode == AXX_BOUNDARY_REJECTION,
                                     p->run.step_mode != AXX_STEP_ISOTROPIC);
                WalkQueue Q;
                unsigned blocks = 0;
                int rc = plan_tasks(p, (const void*)fn, kMagBlock, A.i0, A.i1, &Q, &blocks);
                if (rc) return rc;
                hipLaunchKernelGGL(fn, dim3(blocks), dim3(kMagBlock), p->lds_bytes, ctx->stream, A, G, Q);
                HIP_TRY(ctx, hipGetLastError());
            }
        }
        m->next_step = step_end;
        m->reduced = false;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_time_walk(axx_mag* m, int step_begin, int step_end, float* ms) {
    try {
        if (!m || !ms) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        int rc;
        if (step_begin == 0) {
            // placement and the rotation at t = 0 first, outside the timed region
            if ((rc = axx_mag_walk(m, 0, 1))) return rc;
            step_begin = 1;
        }
        HIP_TRY(ctx, hipEventRecord(m->ev0, ctx->stream));
        if ((rc = axx_mag_walk(m, step_begin, step_end))) return rc;
        HIP_TRY(ctx, hipEventRecord(m->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(m->ev1));
        HIP_TRY(ctx, hipEventElapsedTime(ms, m->ev0, m->ev1));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

static int mag_reduce(axx_mag* m) {
    axx_ctx* ctx = m->ctx;
    hipLaunchKernelGGL(mag_zero_kernel, dim3(1), dim3(64), 0, ctx->stream, m->d_cnt, 4);
    for (size_t g = 0; g < m->plans.size(); ++g) {
        axx_plan* p = m->plans[g];
        const int F = m->gF[g];
        double* part = m->d_part + (size_t)g * kMagGroup * m->n_chunks * 3;
        hipLaunchKernelGGL(mag_reduce_kernel, dim3(m->n_chunks, F), dim3(kMagRedThreads), 0, ctx->stream,
                           (const uint32_t*)p->d_state, p->N, m->n_chunks, part,
                           g == 0 ? m->d_cnt : (unsigned long long*)nullptr);
        HIP_TRY(ctx, hipGetLastError());
        hipLaunchKernelGGL(mag_finalize_kernel, dim3((F * 3 + 63) / 64), dim3(64), 0, ctx->stream,
                           (const double*)part, F, m->n_chunks, m->d_out + (size_t)g * kMagGroup * 3);
        HIP_TRY(ctx, hipGetLastError());
    }
    m->reduced = true;
    return AXX_OK;
}

extern "C" int axx_mag_results(axx_mag* m, axx_mag_result* out) {
    try {
        if (!m || !out) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        if (m->next_step != m->n_steps)
            return fail(ctx, AXX_ERR_STATE, "walk not finished (at step %d of %d)", m->next_step, m->n_steps);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        int rc;
        if (!m->reduced && (rc = mag_reduce(m))) return rc;
        std::vector<double> h((size_t)m->n_fields * 3);
        unsigned long long cnt[4];
        HIP_TRY(ctx, hipMemcpyAsync(h.data(), m->d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(cnt, m->d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        int err = 0;
        for (axx_plan* p : m->plans) {
            int e = 0;
            HIP_TRY(ctx, hipMemcpyAsync(&e, p->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            err |= e;
        }
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (err) return fail(ctx, AXX_ERR_INIT, "rejection initialisation did not finish for some walkers");
        for (int j = 0; j < m->n_fields; ++j) {
            if (out->mx) out->mx[j] = h[(size_t)j * 3 + 0];
            if (out->my) out->my[j] = h[(size_t)j * 3 + 1];
            if (out->mz) out->mz[j] = h[(size_t)j * 3 + 2];
        }
        out->n_walkers = m->N;
        out->collisions = cnt[0];
        out->transmissions = cnt[1];
        out->cap_hits = cnt[2];
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mag_get_state(axx_mag* m, int group, uint32_t* host) {
    try {
        if (!m || !host || group < 0 || group >= (int)m->plans.size())
            return fail(m ? m->ctx : nullptr, AXX_ERR_INVALID, "mag state: bad group");
        axx_ctx* ctx = m->ctx;
        axx_plan* p = m->plans[group];
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const size_t nf = (size_t)F_MOM + 6 * (size_t)m->gF[group];
        HIP_TRY(ctx, hipMemcpy(host, p->d_state, nf * (size_t)p->N * sizeof(uint32_t), hipMemcpyDeviceToHost));
        int err = 0;
        HIP_TRY(ctx, hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost));
        if (err) return fail(ctx, AXX_ERR_INIT, "rejection initialisation did not finish for some walkers");
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" void axx_mag_destroy(axx_mag* m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    mag_free(m);
}

extern "C" int axx_run_mag(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                           axx_mag_result* out) {
    try {
        if (!out) return fail(ctx, AXX_ERR_INVALID, "null result");
        axx_mag* m = nullptr;
        int rc = axx_mag_create(ctx, s, f, r, &m);
        if (rc) return rc;
        rc = axx_mag_walk(m, 0, f->n_steps);
        if (!rc) rc = axx_mag_results(m, out);
        axx_mag_destroy(m);
        return rc;
    } catch (...) {
        return api_exception();
    }
}
