// capi.hip — host side of the C ABI declared in include/axonexa.h.
//
// Replaces the reference's host orchestration of the lattice walk,
// phaseAcquisitionStream<F>::runAcquisitionLattice
// (src/acquisition/phaseAcquisitionStream.cuh:126-190): validation, the
// periodic images and cell bins of the substrate, per-compartment tables,
// device allocation, kernel launches on the context stream and the
// device-to-host copy of the signal sums.  Errors are returned as codes, never
// aborts (the reference's safe_cuda resets the device and asserts,
// src/util/misc.cuh:5-11).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"
#include "init_kernel.h"

// Scratch of the allocation-free entry point axx_run_phase_noalloc: every
// device buffer of such a run (substrate tables, profile table, walker state,
// reduction partials) is a piece of this array.  Each GPU the library runs on
// holds its own copy until the process ends; host code takes its address on
// the current device from hipGetSymbolAddress (axx_create), never from the
// name.  Runs that use it on one GPU at the same time race.
static __device__ __align__(256) unsigned char scratch[AXX_NOALLOC_SCRATCH_BYTES];

extern "C" const char* axx_version(void) { return "axonexa_amd 0.1.0 (gfx950)"; }

extern "C" const char* axx_strerror(int code) {
    switch (code) {
    case AXX_OK: return "ok";
    case AXX_ERR_INVALID: return "invalid argument";
    case AXX_ERR_DEVICE: return "HIP runtime error";
    case AXX_ERR_NOMEM: return "out of memory";
    case AXX_ERR_LIMIT: return "size beyond a compiled limit";
    case AXX_ERR_GEOMETRY: return "invalid geometry (overlap or step >= half cell)";
    case AXX_ERR_INIT: return "rejection initialisation did not finish";
    case AXX_ERR_NODEVICE: return "no HIP device";
    case AXX_ERR_STATE: return "call out of order";
    default: return "unknown error";
    }
}

extern "C" const char* axx_last_error(const axx_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

extern "C" int axx_create(axx_ctx** out, int device) {
    try {
        if (!out) return AXX_ERR_INVALID;
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return AXX_ERR_NODEVICE;
        if (device < 0 || device >= n) return AXX_ERR_INVALID;
        axx_ctx* c = new (std::nothrow) axx_ctx();
        if (!c) return AXX_ERR_NOMEM;
        c->device = device;
        if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return AXX_ERR_DEVICE;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
            c->n_cu = prop.multiProcessorCount;
            c->lds_max = prop.sharedMemPerBlock;
        }
        // this device's copy of the scratch array (allocation-free runs)
        void* sp = nullptr;
        c->scratch = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess
                         ? static_cast<unsigned char*>(sp) : nullptr;
        *out = c;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" void axx_destroy(axx_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

// ---------------------------------------------------------------------------
// Substrate preparation

struct HostPrep {
    int group = 2;               // steps per cell-bin lookup (lookup_group)
    std::vector<float4> geom;
    std::vector<int> meta;
    std::vector<int> orig;
    std::vector<float> L, rate, thr_out, thr_in;
    std::vector<float4> cyl;
    std::vector<float> inv;      // 1 / R per image (float division), 0 for planes
    std::vector<float2> thr;     // per image: (thr_out, thr_in) of its primitive, 0 for planes
    std::vector<uint4> bins;
    std::vector<uint16_t> cand;
    int nb[3] = {1, 1, 1};
    float ib[3] = {1, 1, 1};
    float ib4[2] = {4, 4};
    double bin_width = 0;
    double cell[3] = {0, 0, 0};  // the cell positions wrap in (a unit cell of a repeating substrate)
    int units = 1;               // unit cells in the substrate's cell
    double Lmax = 0;
    float plane_eps = 0;
    bool gen = false;
};

// Unit normal of plane k (normalised in double).
static void unit_normal(const axx_substrate* s, int k, double n[3]) {
    const double* v = s->normal + 3 * k;
    const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int a = 0; a < 3; ++a) n[a] = v[a] / l;
}

// [min, max] of n . p over the box [lo, hi].
static void box_proj(const double n[3], const double lo[3], const double hi[3], double& pmin, double& pmax) {
    pmin = pmax = 0;
    for (int a = 0; a < 3; ++a) {
        const double u = n[a] * lo[a], v = n[a] * hi[a];
        pmin += std::min(u, v);
        pmax += std::max(u, v);
    }
}

static int validate(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q) {
    if (!(s->a > 0 && s->b > 0 && s->c > 0) || !std::isfinite(s->a + s->b + s->c))
        return fail(ctx, AXX_ERR_INVALID, "cell lengths must be finite and > 0");
    if (s->n_prims < 0) return fail(ctx, AXX_ERR_INVALID, "n_prims < 0");
    if (s->n_prims > AXX_MAX_PRIMS) return fail(ctx, AXX_ERR_LIMIT, "n_prims > %d", AXX_MAX_PRIMS);
    if (s->n_prims > 0 && (!s->type || !s->center || !s->radius || !s->D || !s->T2))
        return fail(ctx, AXX_ERR_INVALID, "primitive arrays missing");
    if (!(q->dt > 0) || q->n_steps < 1) return fail(ctx, AXX_ERR_INVALID, "dt and n_steps must be > 0");
    if (q->n_profiles < 1 || q->n_profiles > AXX_MAX_PROFILES)
        return fail(ctx, AXX_ERR_LIMIT, "n_profiles must be in [1, %d]", AXX_MAX_PROFILES);
    if (!q->profiles) return fail(ctx, AXX_ERR_INVALID, "profiles missing");
    if (q->n_waveforms < 1 || !q->wf_profile || !q->wf_amplitude || !q->wf_dir)
        return fail(ctx, AXX_ERR_INVALID, "waveforms missing");
    for (int j = 0; j < q->n_waveforms; ++j) {
        if (q->wf_profile[j] < 0 || q->wf_profile[j] >= q->n_profiles)
            return fail(ctx, AXX_ERR_INVALID, "waveform %d: profile index out of range", j);
        if (!std::isfinite(q->wf_amplitude[j] + q->wf_dir[3 * j] + q->wf_dir[3 * j + 1] + q->wf_dir[3 * j + 2]))
            return fail(ctx, AXX_ERR_INVALID, "waveform %d: non-finite amplitude or direction", j);
    }
    if (!std::isfinite(q->dt) || !std::isfinite(q->gamma)) return fail(ctx, AXX_ERR_INVALID, "dt / gamma");
    for (long long i = 0; i < (long long)q->n_steps * q->n_profiles; ++i)
        if (!std::isfinite(q->profiles[i])) return fail(ctx, AXX_ERR_INVALID, "non-finite profile sample");
    if (!(s->D_lat >= 0) || !std::isfinite(s->D_lat) || std::isnan(s->T2_lat))
        return fail(ctx, AXX_ERR_INVALID, "D_lat / T2_lat");
    double Lmax = std::sqrt(6.0 * s->D_lat * q->dt);
    for (int k = 0; k < s->n_prims; ++k) {
        if (!std::isfinite(s->center[3 * k] + s->center[3 * k + 1] + s->center[3 * k + 2]))
            return fail(ctx, AXX_ERR_INVALID, "primitive %d: non-finite centre", k);
        if (std::isnan(s->T2[k]) || (s->kappa && !(s->kappa[k] >= 0 && std::isfinite(s->kappa[k]))))
            return fail(ctx, AXX_ERR_INVALID, "primitive %d: T2 / permeability", k);
        if (s->type[k] == AXX_PRIM_PLANE) {
            if (!s->normal) return fail(ctx, AXX_ERR_INVALID, "plane normals missing");
            const double* v = s->normal + 3 * k;
            if (!(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 0) || !std::isfinite(v[0] + v[1] + v[2]))
                return fail(ctx, AXX_ERR_INVALID, "primitive %d: plane normal", k);
            continue;
        }
        if (s->type[k] != AXX_PRIM_CYLINDER_XY && s->type[k] != AXX_PRIM_SPHERE)
            return fail(ctx, AXX_ERR_INVALID, "primitive %d: unknown type", k);
        if (!(s->radius[k] > 0) || !(s->D[k] >= 0))
            return fail(ctx, AXX_ERR_INVALID, "primitive %d: radius/D", k);
        const double R = s->radius[k];
        if (2 * R >= s->a || 2 * R >= s->b || (s->type[k] == AXX_PRIM_SPHERE && 2 * R >= s->c))
            return fail(ctx, AXX_ERR_GEOMETRY, "primitive %d larger than half the cell", k);
        Lmax = std::max(Lmax, std::sqrt(6.0 * s->D[k] * q->dt));
    }
    if (2 * Lmax >= s->a || 2 * Lmax >= s->b || 2 * Lmax >= s->c)
        return fail(ctx, AXX_ERR_GEOMETRY, "step length %.3g >= half the cell", Lmax);
    // non-overlap under the minimum image (RPSinitializer.h:191-219, with
    // every image shift)
    const double Ls[3] = {s->a, s->b, s->c};
    for (int i = 0; i < s->n_prims; ++i)
        for (int j = i + 1; j < s->n_prims; ++j) {
            if (s->type[i] == AXX_PRIM_PLANE || s->type[j] == AXX_PRIM_PLANE) continue;
            const int dims = (s->type[i] == AXX_PRIM_SPHERE && s->type[j] == AXX_PRIM_SPHERE) ? 3 : 2;
            double d2 = 0;
            for (int a = 0; a < dims; ++a) {
                double d = s->center[3 * i + a] - s->center[3 * j + a];
                d -= Ls[a] * std::nearbyint(d / Ls[a]);
                d2 += d * d;
            }
            const double rr = s->radius[i] + s->radius[j];
            if (d2 < rr * rr) return fail(ctx, AXX_ERR_GEOMETRY, "primitives %d and %d overlap", i, j);
        }
    return AXX_OK;
}

// The primitives the walk sees and the cell they repeat in.  A substrate
// whose primitives map onto themselves under a shift of a/mx along x and b/my
// along y (a hexagonal lattice built as 20 x 25 copies of its two-cylinder
// unit cell, examples/hexPhase.cu:108-120) is walked in its unit cell
// (a/mx) x (b/my) x c: the unit cell's periodic images are the whole
// substrate, so the walk is the same physics with tables 1/(mx my) the size.
// Each class of translated primitives is represented by its lowest-numbered
// member, whose compartment id (index + 1) every walker inside any member
// carries.  No counterpart in the reference, which scans the whole basis
// (lattice.cuh:184-247).
struct GeoView {
    double a = 0, b = 0, c = 0;      // cell the positions wrap in
    int mx = 1, my = 1;              // unit cells along x and y of the whole cell
    std::vector<int> orig;           // substrate index of each primitive of the view
    std::vector<double> center;      // its centre in the view's cell, 3 per primitive
};

// |d| up to a multiple of L
static double pdist(double d, double L) { return std::fabs(d - L * std::nearbyint(d / L)); }

static bool same_kind(const axx_substrate* s, int i, int j) {
    const double ki = s->kappa ? s->kappa[i] : 0.0, kj = s->kappa ? s->kappa[j] : 0.0;
    return s->type[i] == s->type[j] && s->radius[i] == s->radius[j] && s->D[i] == s->D[j] &&
           (s->T2[i] == s->T2[j] || (std::isinf(s->T2[i]) && std::isinf(s->T2[j]))) && ki == kj;
}

// Does the shift by sh along axis ax map every primitive onto one of the same kind?
static bool shift_maps_onto_itself(const axx_substrate* s, int ax, double sh, double tol) {
    const double L[3] = {s->a, s->b, s->c};
    for (int k = 0; k < s->n_prims; ++k) {
        bool found = false;
        for (int j = 0; j < s->n_prims && !found; ++j) {
            if (!same_kind(s, k, j)) continue;
            bool ok = true;
            for (int a = 0; a < 3 && ok; ++a)
                ok = pdist(s->center[3 * k + a] + (a == ax ? sh : 0.0) - s->center[3 * j + a], L[a]) <= tol;
            found = ok;
        }
        if (!found) return false;
    }
    return true;
}

// Largest m dividing the primitive count whose shift L/m along ax is a symmetry.
static int period_count(const axx_substrate* s, int ax, double tol) {
    const double L = ax == 0 ? s->a : s->b;
    for (int m = s->n_prims; m >= 2; --m)
        if (s->n_prims % m == 0 && shift_maps_onto_itself(s, ax, L / m, tol)) return m;
    return 1;
}

static void whole_cell_view(const axx_substrate* s, GeoView& V) {
    V.a = s->a; V.b = s->b; V.c = s->c;
    V.mx = V.my = 1;
    V.orig.resize(s->n_prims);
    V.center.assign(s->center, s->center + 3 * (size_t)s->n_prims);
    for (int k = 0; k < s->n_prims; ++k) V.orig[k] = k;
}

// The unit-cell view when the substrate repeats (no planes) and the unit cell
// is at least twice every radius and twice the reach of a lookup group
// (reach), else the whole cell.
static void unit_cell_view(const axx_substrate* s, bool allow, double reach, GeoView& V) {
    whole_cell_view(s, V);
    const int n = s->n_prims;
    if (!allow || n < 2) return;
    for (int k = 0; k < n; ++k)
        if (s->type[k] == AXX_PRIM_PLANE) return;
    const double tol = 1e-9 * std::max(std::max(s->a, s->b), s->c);
    const int mx = period_count(s, 0, tol), my = period_count(s, 1, tol);
    if (mx * my == 1) return;
    const double W = s->a / mx, H = s->b / my;
    if (!(2.0 * reach < std::min(W, H))) return;
    std::vector<int> reps;
    std::vector<double> rc;
    for (int k = 0; k < n; ++k) {
        if (2.0 * s->radius[k] >= std::min(W, H)) return;
        double rx = s->center[3 * k] - W * std::floor(s->center[3 * k] / W);
        double ry = s->center[3 * k + 1] - H * std::floor(s->center[3 * k + 1] / H);
        if (W - rx <= tol) rx -= W;
        if (H - ry <= tol) ry -= H;
        const double rz = s->center[3 * k + 2];
        bool member = false;
        for (size_t r = 0; r < reps.size() && !member; ++r)
            member = same_kind(s, k, reps[r]) && std::fabs(rx - rc[3 * r]) <= tol &&
                     std::fabs(ry - rc[3 * r + 1]) <= tol && pdist(rz - rc[3 * r + 2], s->c) <= tol;
        if (member) continue;
        reps.push_back(k);
        rc.push_back(rx); rc.push_back(ry); rc.push_back(rz);
    }
    if ((long long)reps.size() * mx * my != n) return;
    V.a = W; V.b = H;
    V.mx = mx; V.my = my;
    V.orig = reps;
    V.center = rc;
}

static bool reach_hits_cell(const axx_substrate* s, const GeoView& V, int kk, const double sh[3], double reach) {
    const int k = V.orig[kk];
    const double lo[3] = {0, 0, 0}, hi[3] = {V.a, V.b, V.c};
    if (s->type[k] == AXX_PRIM_PLANE) {
        double n[3], pmin, pmax;
        unit_normal(s, k, n);
        double d = 0;
        for (int a = 0; a < 3; ++a) d += n[a] * (V.center[3 * kk + a] + sh[a]);
        box_proj(n, lo, hi, pmin, pmax);
        return d >= pmin - reach && d <= pmax + reach;
    }
    const int dims = s->type[k] == AXX_PRIM_SPHERE ? 3 : 2;
    const double R = s->radius[k] + reach;
    double d2 = 0;
    for (int a = 0; a < dims; ++a) {
        const double c = V.center[3 * kk + a] + sh[a];
        double d = 0;
        if (c < lo[a]) d = lo[a] - c;
        else if (c > hi[a]) d = c - hi[a];
        d2 += d * d;
    }
    return d2 <= R * R;
}

static bool surface_near_box(const float4& g, int meta, const double lo[3], const double hi[3], double reach) {
    if (meta & kPlaneBit) {
        const double n[3] = {g.x, g.y, g.z};
        double pmin, pmax;
        box_proj(n, lo, hi, pmin, pmax);
        return (double)g.w >= pmin - reach && (double)g.w <= pmax + reach;
    }
    const bool sph = (meta & kSphereBit) != 0;
    const double c[3] = {g.x, g.y, g.z};
    const double R = (double)g.w;
    const int dims = sph ? 3 : 2;
    double dmin2 = 0, dmax2 = 0;
    for (int a = 0; a < dims; ++a) {
        double d = 0;
        if (c[a] < lo[a]) d = lo[a] - c[a];
        else if (c[a] > hi[a]) d = c[a] - hi[a];
        dmin2 += d * d;
        const double e = std::max(std::fabs(c[a] - lo[a]), std::fabs(c[a] - hi[a]));
        dmax2 += e * e;
    }
    return std::sqrt(dmin2) <= R + reach && std::sqrt(dmax2) >= R - reach;
}

// LDS tables: the azimuth table (2 KB), bins (16 B), images (16 B),
// per-image transmission hazards (8 B), meta words (4 B, general geometry),
// inverse radii (4 B), candidates (2 B).
static size_t lds_size(size_t n_img, size_t n_bins, size_t n_cand, bool gen) {
    return kAzimuthEntries * sizeof(float2) + n_bins * 16 + n_img * (gen ? 32 : 28) + n_cand * 2;
}

// Images, compartments and bins.  The float32 image centres are
// (float)(centre + shift) computed in double; compartments use
// L = (float)sqrt(6 D dt), rate = (float)(dt / T2), transmission hazards
// (float)(-2 ln(1 - 4 kappa / v)).
// Hazard of a boundary event that transmits with probability p = 4 kappa / v
// (boundaryCheck.cuh:94): h = -2 ln(1 - p), so that a budget drawn as
// 2 Exp(1) = -2 ln u runs out at this event with probability p; p >= 1
// always transmits.
static float transmission_hazard(double p) {
    return p >= 1.0 ? INFINITY : (float)(-2.0 * std::log1p(-p));
}

// Longest step of a step generator in units of sqrt(6 D dt): Gaussian steps
// have |n| <= sqrt(2 * 33.27) = 8.157 standard normals (24-bit uniforms),
// i.e. at most 4.71 such units.
static double step_reach_factor(int step_mode) { return step_mode == AXX_STEP_GAUSSIAN ? 4.75 : 1.0; }

// Steps per cell-bin lookup; positions wrap into the cell after the last
// step of each group.  When steps have a fixed length and every surface is a
// cylinder: eight if every cylinder is at least sixteen step lengths from
// every other (periodic images included), four if at least eight -- the
// longer reach then keeps the candidate lists short, and one lookup serves
// four or two pairs.  Two otherwise (close surfaces, spheres, planes or
// Gaussian steps, where the longer lists cost more than the saved lookups).
// The oracle applies the same rule with the same arithmetic.
static int lookup_group(const axx_substrate* s, double Lmax, double reach_factor) {
    if (reach_factor != 1.0) return 2;
    const int n = s->n_prims;
    for (int k = 0; k < n; ++k)
        if (s->type[k] != AXX_PRIM_CYLINDER_XY) return 2;
    const double lim4 = 8.0 * Lmax, lim8 = 16.0 * Lmax;
    int g = 8;
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j)
            for (int sy = -1; sy <= 1; ++sy)
                for (int sx = -1; sx <= 1; ++sx) {
                    if (i == j && sx == 0 && sy == 0) continue;
                    const double dx = s->center[3 * j] + sx * s->a - s->center[3 * i];
                    const double dy = s->center[3 * j + 1] + sy * s->b - s->center[3 * i + 1];
                    const double d2 = dx * dx + dy * dy;
                    const double r4 = lim4 + s->radius[i] + s->radius[j];
                    const double r8 = lim8 + s->radius[i] + s->radius[j];
                    if (d2 < r4 * r4) return 2;
                    if (d2 < r8 * r8) g = 4;
                }
    return g;
}

static int prepare(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q, size_t lds_budget,
                   double bin_width_req, int step_mode, bool allow_unit_cell, HostPrep& H) {
    const int n = s->n_prims;
    const double dt = q->dt;
    H.L.resize(n + 1);
    H.rate.resize(n + 1);
    H.thr_out.assign(std::max(n, 1), 0.0f);
    H.thr_in.assign(std::max(n, 1), 0.0f);
    double Lmax = std::sqrt(6.0 * s->D_lat * dt);
    H.L[0] = (float)std::sqrt(6.0 * s->D_lat * dt);
    H.rate[0] = (s->T2_lat > 0 && std::isfinite(s->T2_lat)) ? (float)(dt / s->T2_lat) : 0.0f;
    const double v_lat = std::sqrt(6.0 * s->D_lat / dt);
    H.plane_eps = (float)(64.0 * std::ldexp(1.0, -23) * std::max(std::max(s->a, s->b), s->c));
    bool any_plane = false;
    for (int k = 0; k < n; ++k) {
        if (s->type[k] == AXX_PRIM_PLANE) {
            // planes bound no compartment
            H.L[k + 1] = 0.0f;
            H.rate[k + 1] = 0.0f;
            any_plane = true;
            H.gen = true;
            continue;
        }
        const double L = std::sqrt(6.0 * s->D[k] * dt);
        Lmax = std::max(Lmax, L);
        H.L[k + 1] = (float)L;
        H.rate[k + 1] = (s->T2[k] > 0 && std::isfinite(s->T2[k])) ? (float)(dt / s->T2[k]) : 0.0f;
        const double kap = s->kappa ? s->kappa[k] : 0.0;
        const double v_in = std::sqrt(6.0 * s->D[k] / dt);
        H.thr_out[k] = kap > 0 ? transmission_hazard(4.0 * kap / v_in) : 0.0f;
        H.thr_in[k] = kap > 0 ? transmission_hazard(4.0 * kap / v_lat) : 0.0f;
        if (s->type[k] == AXX_PRIM_SPHERE) H.gen = true;
    }
    H.Lmax = Lmax;
    H.group = lookup_group(s, Lmax, step_reach_factor(step_mode));
    // reach of a lookup group (positions wrap at the group's end, so a
    // lookup's start may lie up to group - 1 steps outside the cell, with the
    // rest of the group still to walk); a plane counts as hit within
    // plane_eps of it, so its reach is longer
    const double reach = H.group * step_reach_factor(step_mode) * Lmax * 1.001 + 1e-6 * std::max(s->a, s->b) +
                         (any_plane ? 4.0 * H.plane_eps : 0.0);
    GeoView V;
    unit_cell_view(s, allow_unit_cell, reach, V);
    H.cell[0] = V.a; H.cell[1] = V.b; H.cell[2] = V.c;
    H.units = V.mx * V.my;
    for (int kk = 0; kk < (int)V.orig.size(); ++kk) {
        const int k = V.orig[kk];
        const double* ck = &V.center[3 * (size_t)kk];
        const int zr = s->type[k] != AXX_PRIM_CYLINDER_XY ? 1 : 0;
        const size_t first = H.geom.size();
        for (int kz = -zr; kz <= zr; ++kz)
            for (int ky = -1; ky <= 1; ++ky)
                for (int kx = -1; kx <= 1; ++kx) {
                    const double sh[3] = {kx * V.a, ky * V.b, kz * V.c};
                    if (!reach_hits_cell(s, V, kk, sh, reach)) continue;
                    if (s->type[k] == AXX_PRIM_PLANE) {
                        // shifts along the plane give the same plane: one image
                        // per distinct float offset
                        double nn[3];
                        unit_normal(s, k, nn);
                        const float dd = (float)(nn[0] * (ck[0] + sh[0]) + nn[1] * (ck[1] + sh[1]) +
                                                 nn[2] * (ck[2] + sh[2]));
                        bool dup = false;
                        for (size_t q2 = first; q2 < H.geom.size(); ++q2) dup |= H.geom[q2].w == dd;
                        if (dup) continue;
                        H.geom.push_back(make_float4((float)nn[0], (float)nn[1], (float)nn[2], dd));
                        H.meta.push_back((k + 1) | kPlaneBit);
                        H.orig.push_back(k);
                        continue;
                    }
                    const float Rf = (float)s->radius[k];
                    H.geom.push_back(make_float4((float)(ck[0] + sh[0]), (float)(ck[1] + sh[1]),
                                                 (float)(ck[2] + sh[2]), Rf));
                    H.meta.push_back((k + 1) | (s->type[k] == AXX_PRIM_SPHERE ? kSphereBit : 0));
                    H.orig.push_back(k);
                }
    }
    const int n_img = (int)H.geom.size();
    if (n_img > 65535) return fail(ctx, AXX_ERR_LIMIT, "too many primitive images (%d)", n_img);
    for (int i = 0; i < n_img; ++i) {
        float mf;
        std::memcpy(&mf, &H.meta[i], sizeof(mf));
        H.cyl.push_back(make_float4(H.geom[i].x, H.geom[i].y, H.geom[i].w, mf));
        const float Rf = H.geom[i].w;
        H.inv.push_back((H.meta[i] & kPlaneBit) ? 0.0f : 1.0f / Rf);
        const int ci = H.meta[i] & kCompMask;
        H.thr.push_back((H.meta[i] & kPlaneBit) || ci < 1 ? make_float2(0.0f, 0.0f)
                                                           : make_float2(H.thr_out[ci - 1], H.thr_in[ci - 1]));
    }
    // bin width: at least 4 step lengths; widen until the tables fit in LDS.
    // Each candidate carries a 4 x 4 mask of the sub-bins (xy) whose
    // rectangle its surface comes within reach of.
    const double Ls[3] = {V.a, V.b, V.c};
    double w = bin_width_req > 0 ? bin_width_req : std::max(4.0 * Lmax, 1e-9);
    for (int attempt = 0; attempt < 64; ++attempt) {
        int nb[3];
        for (int a = 0; a < 3; ++a) {
            int v = (int)std::floor(Ls[a] / w);
            nb[a] = std::min(std::max(v, 1), 2048);
        }
        if (!H.gen) nb[2] = 1;
        const long long nbins = (long long)nb[0] * nb[1] * nb[2];
        if (nbins > (1 << 22) || lds_size(n_img, nbins, 0, H.gen) > lds_budget) {
            if (bin_width_req > 0) return fail(ctx, AXX_ERR_LIMIT, "bin tables exceed LDS");
            w *= 1.25;
            continue;
        }
        // per-bin lists, filled image by image (ascending image index) over
        // the bins that the image's grown bounding box touches
        std::vector<std::vector<uint32_t>> lists(nbins);
        const double bw[3] = {Ls[0] / nb[0], Ls[1] / nb[1], Ls[2] / nb[2]};
        for (int i = 0; i < n_img; ++i) {
            const int mi = H.meta[i];
            const bool sk = (mi & kSphereBit) != 0, pk = (mi & kPlaneBit) != 0;
            const double R = (double)H.geom[i].w + reach;
            const double c3[3] = {H.geom[i].x, H.geom[i].y, H.geom[i].z};
            int lo_b[3], hi_b[3];
            for (int a = 0; a < 3; ++a) {
                if (pk || (a == 2 && !sk)) { lo_b[a] = 0; hi_b[a] = nb[a] - 1; continue; }
                lo_b[a] = std::max(0, (int)std::floor((c3[a] - R) / bw[a]) - 1);
                hi_b[a] = std::min(nb[a] - 1, (int)std::floor((c3[a] + R) / bw[a]) + 1);
            }
            for (int bz = lo_b[2]; bz <= hi_b[2]; ++bz)
                for (int by = lo_b[1]; by <= hi_b[1]; ++by)
                    for (int bx = lo_b[0]; bx <= hi_b[0]; ++bx) {
                        const double lo[3] = {bx * bw[0], by * bw[1], bz * bw[2]};
                        const double hi[3] = {(bx + 1) * bw[0], (by + 1) * bw[1], (bz + 1) * bw[2]};
                        if (!surface_near_box(H.geom[i], mi, lo, hi, reach)) continue;
                        uint32_t mask = 0;
                        for (int ly = 0; ly < 4; ++ly)
                            for (int lx = 0; lx < 4; ++lx) {
                                const double slo[3] = {lo[0] + lx * bw[0] / 4, lo[1] + ly * bw[1] / 4, lo[2]};
                                const double shi[3] = {lo[0] + (lx + 1) * bw[0] / 4, lo[1] + (ly + 1) * bw[1] / 4,
                                                       hi[2]};
                                if (surface_near_box(H.geom[i], mi, slo, shi, reach)) mask |= 1u << (lx + 4 * ly);
                            }
                        if (mask) lists[(bz * nb[1] + by) * nb[0] + bx].push_back((uint32_t)i | (mask << 16));
                    }
        }
        // bin record: {offset << 8 | count, nibbles of sub-bins 0-7, 8-15, 0};
        // nibble = position of the only surface within reach, 7 none, 8 several
        std::vector<uint4> bins(nbins);
        std::vector<uint16_t> cand;
        bool ok = true;
        for (long long bi = 0; bi < nbins && ok; ++bi) {
            const size_t off = cand.size();
            const size_t cnt = lists[bi].size();
            if (cnt > 255 || off >= (1u << 24)) { ok = false; break; }
            uint32_t nib[2] = {0, 0};
            for (int sb = 0; sb < 16; ++sb) {
                int nrel = 0, pos = 0;
                for (size_t j = 0; j < cnt; ++j)
                    if ((lists[bi][j] >> (16 + sb)) & 1u) { if (!nrel) pos = (int)j; ++nrel; }
                uint32_t v = nrel == 0 ? kNibNone : (nrel == 1 && pos < 7 ? (uint32_t)pos : kNibMany);
                nib[sb >> 3] |= v << ((sb & 7) * 4);
            }
            for (size_t j = 0; j < cnt; ++j) cand.push_back((uint16_t)(lists[bi][j] & 0xFFFFu));
            bins[bi] = make_uint4((uint32_t)(off << 8) | (uint32_t)cnt, nib[0], nib[1], 0u);
        }
        if (!ok || lds_size(n_img, nbins, cand.size(), H.gen) > lds_budget) {
            if (bin_width_req > 0) return fail(ctx, AXX_ERR_LIMIT, "bin tables exceed LDS");
            if (!ok) { w *= 0.8; if (w < Lmax) return fail(ctx, AXX_ERR_LIMIT, "cannot bin substrate"); continue; }
            w *= 1.25;
            continue;
        }
        for (int a = 0; a < 3; ++a) {
            H.nb[a] = nb[a];
            H.ib[a] = (float)(nb[a] / Ls[a]);
        }
        H.ib4[0] = (float)(4.0 * nb[0] / Ls[0]);
        H.ib4[1] = (float)(4.0 * nb[1] / Ls[1]);
        H.bins.swap(bins);
        H.cand.swap(cand);
        if (H.cand.empty()) H.cand.push_back(0);
        H.bin_width = w;
        return AXX_OK;
    }
    return fail(ctx, AXX_ERR_LIMIT, "substrate tables do not fit in LDS");
}

// ---------------------------------------------------------------------------
// Kernel dispatch over (profile width, axes, general geometry, rejection) instantiations

// (gen implies groups of two: lookup_group gives four to cylinders only)
static walk_fn pick_walk(const axx_plan* p) {
    const int ax = p->ax, gm = p->gmask;
    const bool gen = p->gen, rej = p->run.boundary_mode == AXX_BOUNDARY_REJECTION;
    const bool alt = p->run.step_mode != AXX_STEP_ISOTROPIC;
    switch (p->Pt) {
    case 1: return pick_walk_p1(ax, gen, rej, alt, gm);
    case 2: return pick_walk_p2(ax, gen, rej, alt, gm);
    case 4: return pick_walk_p4(ax, gen, rej, alt, gm);
    case 8: return pick_walk_p8(ax, gen, rej, alt, gm);
    case 16: return pick_walk_p16(ax, gen, rej, alt, gm);
    case 32: return pick_walk_p32(ax, gen, rej, alt, gm);
    default: return pick_walk_p64(ax, gen, rej, alt, gm);
    }
}

// Default threads per block: of 256, 512 and 1024 (within the kernel's
// bound) the one that keeps the most waves resident per CU for this plan's
// kernel, register use and LDS tables, the smaller on a tie (more, smaller
// workgroups also spread small walker counts over every CU).  The 1000-
// cylinder hexagonal workload in its unit cell: 256 (6 waves per SIMD),
// 2.18e11 walker-steps/s against 1.97e11 at 1024 (4 waves).
static int default_block(const axx_plan* p) {
    const walk_fn fn = pick_walk(p);
    int best = 256, best_waves = -1;
    for (int blk : {256, 512, 1024}) {
        if (blk > walk_max_block(p->Pt)) continue;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, blk, p->lds_bytes) != hipSuccess)
            continue;
        if (per_cu * blk / 64 > best_waves) {
            best = blk;
            best_waves = per_cu * blk / 64;
        }
    }
    return best;
}

static int axes_mask(const axx_acq* q) {
    int mask = 0;
    for (int j = 0; j < q->n_waveforms; ++j)
        for (int a = 0; a < 3; ++a)
            if (q->wf_dir[3 * j + a] != 0.0 && q->wf_amplitude[j] != 0.0) mask |= 1 << a;
    static const int widen[8] = {1, 1, 3, 3, 5, 5, 7, 7};
    return widen[mask];
}

// Device memory of a plan: hipMalloc, or (allocation-free runs) the next
// 256-byte aligned piece of the scratch array.
static int dev_alloc(axx_ctx* ctx, DevArena* ar, void** out, size_t bytes) {
    if (!ar) {
        HIP_TRY(ctx, hipMalloc(out, bytes));
        return AXX_OK;
    }
    const size_t off = (ar->used + 255) & ~(size_t)255;
    if (off > ar->size || bytes > ar->size - off)
        return fail(ctx, AXX_ERR_LIMIT, "run needs more than the %zu bytes of scratch of an allocation-free run",
                    ar->size);
    *out = ar->base + off;
    ar->used = off + bytes;
    return AXX_OK;
}

template <typename T>
static int dupload(axx_ctx* ctx, DevArena* ar, T** dst, const T* src, size_t n) {
    if (n == 0) n = 1;
    int rc = dev_alloc(ctx, ar, (void**)dst, n * sizeof(T));
    if (rc) return rc;
    if (src) HIP_TRY(ctx, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return AXX_OK;
}

static __global__ void zero_words(uint32_t* p, size_t n) {
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0u;
}

// Zero n words: hipMemset for a plan's own buffers, a fill kernel on the
// context stream for scratch (allocation-free runs).
static int zero_dev(axx_ctx* ctx, DevArena* ar, void* d, size_t n_words) {
    if (!ar) {
        HIP_TRY(ctx, hipMemset(d, 0, n_words * sizeof(uint32_t)));
        return AXX_OK;
    }
    const unsigned blocks = (unsigned)std::min<size_t>((n_words + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_words, dim3(std::max(blocks, 1u)), dim3(256), 0, ctx->stream, (uint32_t*)d, n_words);
    HIP_TRY(ctx, hipGetLastError());
    return AXX_OK;
}

void plan_free(axx_plan* p) {
    if (!p) return;
    if (p->arena) {   // buffers are scratch, events were never made
        delete p;
        return;
    }
    void* bufs[] = {p->d_geom, p->d_meta, p->d_cyl, p->d_inv, p->d_thr, p->d_bins, p->d_cand, p->d_L, p->d_rate,
                    p->d_prof, p->d_state, p->d_err, p->d_wfp, p->d_wford,
                    p->d_wfdir, p->d_wfscale, p->d_S, p->d_partial, p->d_out, p->d_cnt, p->d_queue};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (p->ev0) (void)hipEventDestroy(p->ev0);
    if (p->ev1) (void)hipEventDestroy(p->ev1);
    delete p;
}

extern "C" int axx_plan_create(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q,
                               const axx_run* r, axx_plan** out) {
    try {
        return axx_plan_create_ex(ctx, s, q, r, nullptr, out);
    } catch (...) {
        return api_exception();
    }
}

extern "C" void axx_plan_destroy(axx_plan* plan) {
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    (void)hipStreamSynchronize(plan->ctx->stream);
    plan_free(plan);
}

int plan_create_impl(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q, const axx_run* r,
                     const axx_tuning* tune, axx_plan* p, int min_pt, DevArena* ar) {
    int rc = validate(ctx, s, q);
    if (rc) return rc;
    if (r->n_walkers < 1) return fail(ctx, AXX_ERR_INVALID, "n_walkers must be >= 1");
    if (r->n_walkers > (1ll << 32)) return fail(ctx, AXX_ERR_LIMIT, "n_walkers > 2^32");
    if (r->boundary_mode < 0 || r->boundary_mode > 2) return fail(ctx, AXX_ERR_INVALID, "boundary_mode");
    if (r->init_mode < 0 || r->init_mode > 2) return fail(ctx, AXX_ERR_INVALID, "init_mode");
    if (r->init_mode == AXX_INIT_INSIDE_PRIM && (r->init_arg < 0 || r->init_arg >= s->n_prims))
        return fail(ctx, AXX_ERR_INVALID, "init_arg: no such primitive");
    if (r->init_mode == AXX_INIT_COMPARTMENT && (r->init_arg < 0 || r->init_arg > s->n_prims))
        return fail(ctx, AXX_ERR_INVALID, "init_arg: no such compartment");
    if ((r->init_mode == AXX_INIT_INSIDE_PRIM && s->type[r->init_arg] == AXX_PRIM_PLANE) ||
        (r->init_mode == AXX_INIT_COMPARTMENT && r->init_arg > 0 && s->type[r->init_arg - 1] == AXX_PRIM_PLANE))
        return fail(ctx, AXX_ERR_INVALID, "init_arg: planes bound no region");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HostPrep H;
    size_t budget = 150 * 1024;
    if (tune && tune->lds_budget > 0) budget = (size_t)tune->lds_budget;
    budget = std::min<size_t>(budget, ctx->lds_max - kWalkStaticLds);   // the walk kernel's own LDS words
    if (tune && tune->block && (tune->block < 64 || tune->block > 1024 || tune->block % 64))
        return fail(ctx, AXX_ERR_INVALID, "block must be a multiple of 64 in [64, 1024]");
    if (tune && tune->bin_width < 0) return fail(ctx, AXX_ERR_INVALID, "bin_width < 0");
    if (r->step_mode < 0 || r->step_mode > 2) return fail(ctx, AXX_ERR_INVALID, "step_mode");
    if (r->step_mode == AXX_STEP_GAUSSIAN) {
        double Lmax = std::sqrt(6.0 * s->D_lat * q->dt);
        for (int k = 0; k < s->n_prims; ++k)
            if (s->type[k] != AXX_PRIM_PLANE) Lmax = std::max(Lmax, std::sqrt(6.0 * s->D[k] * q->dt));
        const double need = 2.0 * step_reach_factor(AXX_STEP_GAUSSIAN) * Lmax;
        if (need >= s->a || need >= s->b || need >= s->c)
            return fail(ctx, AXX_ERR_GEOMETRY, "Gaussian steps need a cell wider than %.3g", need);
    }
    // the unit cell of a repeating substrate unless the tuning asks for the
    // whole cell or the placement names one particular primitive
    const bool allow_unit = !(tune && tune->whole_cell) &&
                            (r->init_mode == AXX_INIT_UNIFORM ||
                             (r->init_mode == AXX_INIT_COMPARTMENT && r->init_arg == 0));
    rc = prepare(ctx, s, q, budget, tune ? tune->bin_width : 0.0, r->step_mode, allow_unit, H);
    if (rc) return rc;
    p->ctx = ctx;
    p->run = *r;
    p->N = r->n_walkers;
    p->n_prims = s->n_prims;
    p->n_steps = q->n_steps;
    p->dt = q->dt;
    p->P = q->n_profiles;
    p->Pt = std::max(1, min_pt);
    while (p->Pt < p->P) p->Pt *= 2;
    p->M = q->n_waveforms;
    p->ax = axes_mask(q);
    p->gen = H.gen;
    p->plane_eps = H.plane_eps;
    p->gmask = H.group - 1;
    p->cell[0] = (float)H.cell[0]; p->cell[1] = (float)H.cell[1]; p->cell[2] = (float)H.cell[2];
    p->whole[0] = (float)s->a; p->whole[1] = (float)s->b;
    p->units = H.units;
    {
        const double hc = 0.5 * std::min(H.cell[0], H.cell[1]);
        const double hs = 0.5 * std::min(std::min(H.cell[0], H.cell[1]), H.cell[2]);
        p->H2c = (float)(hc * hc);
        p->H2s = (float)(hs * hs);
    }
    p->n_img = (int)H.geom.size();
    p->n_bins = (int)H.bins.size();
    p->n_cand = (int)H.cand.size();
    for (int a = 0; a < 3; ++a) { p->nb[a] = H.nb[a]; p->ib[a] = H.ib[a]; }
    p->ib4[0] = H.ib4[0]; p->ib4[1] = H.ib4[1];
    p->bin_width = H.bin_width;
    p->lds_bytes = lds_size(p->n_img, p->n_bins, p->n_cand, p->gen);
    p->lds_bytes = (p->lds_bytes + 15) & ~(size_t)15;
    p->init_lds = ((size_t)p->n_img * 20 + 15) & ~(size_t)15;
    if (p->init_lds > ctx->lds_max) return fail(ctx, AXX_ERR_LIMIT, "too many primitive images for LDS");
    p->init_img = 0;
    if (r->init_mode == AXX_INIT_INSIDE_PRIM) {
        p->init_img = -1;
        for (int i = 0; i < p->n_img; ++i)
            if (H.orig[i] == r->init_arg) { p->init_img = i; break; }
        if (p->init_img < 0) return fail(ctx, AXX_ERR_INVALID, "init primitive has no image in the cell");
    }
    // profiles padded to the template width
    std::vector<float> prof((size_t)q->n_steps * p->Pt, 0.0f);
    for (int i = 0; i < q->n_steps; ++i)
        for (int k = 0; k < p->P; ++k) prof[(size_t)i * p->Pt + k] = q->profiles[(size_t)i * p->P + k];
    p->S.assign(p->Pt, 0.0);
    for (int i = 0; i < q->n_steps; ++i)
        for (int k = 0; k < p->P; ++k) p->S[k] += (double)q->profiles[(size_t)i * p->P + k];
    p->wf_p.assign(q->wf_profile, q->wf_profile + p->M);
    p->wf_ord.resize(p->M);
    for (int j = 0; j < p->M; ++j) p->wf_ord[j] = j;
    std::stable_sort(p->wf_ord.begin(), p->wf_ord.end(),
                     [&](int a, int b) { return p->wf_p[a] < p->wf_p[b]; });
    p->wf_dir.assign(q->wf_dir, q->wf_dir + 3 * p->M);
    p->wf_scale.resize(p->M);
    for (int j = 0; j < p->M; ++j) p->wf_scale[j] = q->gamma * q->dt * q->wf_amplitude[j];

    if ((rc = dupload(ctx, ar, &p->d_geom, H.geom.data(), H.geom.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_meta, H.meta.data(), H.meta.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_cyl, H.cyl.data(), H.cyl.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_inv, H.inv.data(), H.inv.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_thr, H.thr.data(), H.thr.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_bins, H.bins.data(), H.bins.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_cand, H.cand.data(), H.cand.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_L, H.L.data(), H.L.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_rate, H.rate.data(), H.rate.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_prof, prof.data(), prof.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_wfp, p->wf_p.data(), p->wf_p.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_wford, p->wf_ord.data(), p->wf_ord.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_wfdir, p->wf_dir.data(), p->wf_dir.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_wfscale, p->wf_scale.data(), p->wf_scale.size()))) return rc;
    if ((rc = dupload(ctx, ar, &p->d_S, p->S.data(), p->S.size()))) return rc;
    const size_t nst = (size_t)AXX_STATE_FIELDS(p->Pt) * (size_t)p->N;
    if (ar) {
        if ((rc = dev_alloc(ctx, ar, (void**)&p->d_state, nst * sizeof(uint32_t)))) return rc;
    } else if (hipMalloc((void**)&p->d_state, nst * sizeof(uint32_t)) != hipSuccess) {
        return fail(ctx, AXX_ERR_NOMEM, "walker state (%zu bytes)", nst * 4);
    }
    if ((rc = zero_dev(ctx, ar, p->d_state, nst))) return rc;
    if ((rc = dupload<int>(ctx, ar, &p->d_err, nullptr, 1))) return rc;
    if ((rc = zero_dev(ctx, ar, p->d_err, 1))) return rc;
    p->chunk = kRedThreads * 16;
    p->n_chunks = (int)((p->N + p->chunk - 1) / p->chunk);
    if ((rc = dupload<double>(ctx, ar, &p->d_partial, nullptr, (size_t)p->M * p->n_chunks * kRedVals))) return rc;
    if ((rc = dupload<double>(ctx, ar, &p->d_out, nullptr, (size_t)p->M * kRedVals))) return rc;
    if ((rc = dupload<unsigned long long>(ctx, ar, &p->d_cnt, nullptr, 4))) return rc;
    if ((rc = dupload<unsigned int>(ctx, ar, &p->d_queue, nullptr, 1 + (size_t)((p->N + 63) / 64)))) return rc;
    if (!ar) {
        HIP_TRY(ctx, hipEventCreate(&p->ev0));
        HIP_TRY(ctx, hipEventCreate(&p->ev1));
    }
    p->arena = ar != nullptr;
    if (tune && tune->block) {
        p->block = tune->block;
    } else {
        p->block = default_block(p);
    }
    p->wpt = 1;
    HIP_TRY(ctx, hipDeviceSynchronize());
    return AXX_OK;
}

extern "C" int axx_plan_create_ex(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q,
                                  const axx_run* r, const axx_tuning* tune, axx_plan** out) {
    try {
        if (!ctx || !s || !q || !r || !out) return fail(ctx, AXX_ERR_INVALID, "null argument");
        *out = nullptr;
        axx_plan* p = new (std::nothrow) axx_plan();
        if (!p) return AXX_ERR_NOMEM;
        p->ctx = ctx;
        int rc;
        try {
            rc = plan_create_impl(ctx, s, q, r, tune, p, 1);
        } catch (...) {
            plan_free(p);
            throw;
        }
        if (rc) {
            plan_free(p);
            return rc;
        }
        *out = p;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

WalkArgs plan_walk_args(const axx_plan* p, int i0, int i1) {
    WalkArgs A{};
    A.a = p->cell[0]; A.b = p->cell[1]; A.c = p->cell[2];
    A.ca = p->whole[0]; A.cb = p->whole[1];
    A.H2c = p->H2c; A.H2s = p->H2s; A.plane_eps = p->plane_eps;
    A.gmask = p->gmask;
    A.nbx = p->nb[0]; A.nby = p->nb[1]; A.nbz = p->nb[2];
    A.ibz = p->ib[2];
    A.ibx4 = p->ib4[0]; A.iby4 = p->ib4[1];
    A.n_img = p->n_img; A.n_bins = p->n_bins; A.n_cand = p->n_cand;
    A.g_geom = p->d_geom; A.g_meta = p->d_meta; A.g_cyl = p->d_cyl; A.g_bins = p->d_bins; A.g_cand = p->d_cand;
    A.g_inv = p->d_inv;
    A.g_thr = p->d_thr;
    A.compL = p->d_L; A.rate = p->d_rate;
    A.prof = p->d_prof;
    A.state = p->d_state;
    A.N = p->N;
    A.k0 = (uint32_t)p->run.seed; A.k1 = (uint32_t)(p->run.seed >> 32);
    A.inst = p->run.instance; A.woff = p->run.walker_offset;
    A.mode = p->run.boundary_mode; A.init_mode = p->run.init_mode; A.init_arg = p->run.init_arg;
    A.step_mode = p->run.step_mode;
    A.init_img = p->init_img;
    A.i0 = i0; A.i1 = i1;
    A.cap = p->run.bounce_cap > 0 ? p->run.bounce_cap : AXX_DEFAULT_BOUNCE_CAP;
    A.err = p->d_err;
    return A;
}

// Automatic task length: a sixteenth of the walk within [512, 4096] steps.
// The hexagonal workload (1e6 walkers, 250,000 steps) runs 2.662e11,
// 2.667e11, 2.668e11 and 2.660e11 walker-steps/s at 1024, 2048, 4096 and
// 6144 steps per task, 2.642e11 at 16384, 2.375e11 with whole walks; the
// launch ends within one task of the last group's last chunk, so a sixteenth
// bounds that tail for short walks.
constexpr int kAutoChunks = 16;
constexpr int kMinTaskSteps = 512;
constexpr int kMaxTaskSteps = 4096;
constexpr uint32_t kWaitTicks = 2000000000u; // 20 s of 100 MHz real time: longest wait of one task

// Steps per walk task.  A task is one group of 64 walkers over one time
// chunk; chunks pay off once the groups fill the resident waves at least
// twice: then every wave stays busy until the last chunk of the last group,
// instead of the launch ending with a partly filled round of whole walks.
static int task_steps(const axx_plan* p, long long groups, long long resident_waves, int steps) {
    if (p->sched > 0) return std::min(steps, (p->sched + 1) & ~1);
    if (p->sched < 0 || groups < 2 * resident_waves) return steps;
    if (p->auto_task > 0) return std::min(steps, (p->auto_task + 1) & ~1);
    const int target = (steps + kAutoChunks - 1) / kAutoChunks;
    return std::min(steps, std::min(kMaxTaskSteps, std::max(kMinTaskSteps, (target + 1) & ~1)));
}

// Grid and task queue of a persistent walk launch of kernel fn (threads
// per block `block`, the plan's LDS) over steps [i0, i1): at most
// the resident workgroups, no more waves than walker groups; zeroes the
// queue words the launch uses on the context stream.
int plan_tasks(axx_plan* p, const void* fn, int block, int i0, int i1, WalkQueue* Q, unsigned* blocks) {
    axx_ctx* ctx = p->ctx;
    int per_cu = 0;
    HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, p->lds_bytes));
    if (per_cu < 1) return fail(ctx, AXX_ERR_LIMIT, "walk kernel does not fit a CU at %d threads", block);
    const long long wpb = block / 64;
    const long long resident = (long long)per_cu * ctx->n_cu;
    const long long groups = (p->N + 63) / 64;
    const int steps = std::max(i1 - std::max(i0, 1), 1);   // chunks count from step 1 (step 0: chunk 0)
    int chunk = task_steps(p, groups, resident * wpb, steps);
    long long chunks = (steps + chunk - 1) / chunk;
    if (groups * chunks >= 0x7fffffffll) {   // one chunk per walk keeps the task count below 2^31
        chunk = steps;
        chunks = 1;
    }
    *Q = WalkQueue{};
    Q->q = p->d_queue;
    Q->n_groups = (uint32_t)groups;
    Q->n_tasks = (uint32_t)(groups * chunks);
    Q->chunk = chunk;
    Q->wait_ticks = kWaitTicks;
    *blocks = (unsigned)std::min(resident, (groups + wpb - 1) / wpb);
    const size_t qwords = chunks > 1 ? (size_t)groups + 1 : 1;
    hipLaunchKernelGGL(zero_words, dim3((unsigned)std::min<size_t>((qwords + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, (uint32_t*)p->d_queue, qwords);
    HIP_TRY(ctx, hipGetLastError());
    p->last_task_steps = chunk;
    p->last_grid = (int)*blocks;
    return AXX_OK;
}

static int launch_walk(axx_plan* p, int i0, int i1) {
    axx_ctx* ctx = p->ctx;
    WalkArgs A = plan_walk_args(p, i0, i1);
    walk_fn fn = pick_walk(p);
    const int block = std::min(p->block, walk_max_block(p->Pt));
    const long long grid = (p->N + block - 1) / block;
    if (grid > 0x7fffffff) return fail(ctx, AXX_ERR_LIMIT, "grid too large");
    if (i0 == 0) {
        hipLaunchKernelGGL(init_kernel, dim3((unsigned)grid), dim3(block), p->init_lds, ctx->stream, A,
                           3 * p->Pt);
        HIP_TRY(ctx, hipGetLastError());
        A.i0 = 1;
    }
    if (A.i0 < A.i1) {
        WalkQueue Q;
        unsigned blocks = 0;
        int rc = plan_tasks(p, (const void*)fn, block, A.i0, A.i1, &Q, &blocks);
        if (rc) return rc;
        hipLaunchKernelGGL(fn, dim3(blocks), dim3(block), p->lds_bytes, ctx->stream, A, Q, (const float*)p->d_prof);
        HIP_TRY(ctx, hipGetLastError());
    }
    return AXX_OK;
}

extern "C" int axx_plan_walk(axx_plan* p, int step_begin, int step_end) {
    try {
        if (!p) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        if (step_begin < 0 || step_end > p->n_steps || step_begin > step_end)
            return fail(ctx, AXX_ERR_INVALID, "step range [%d, %d) outside [0, %d)", step_begin, step_end,
                        p->n_steps);
        if (step_begin != 0 && step_begin != p->next_step)
            return fail(ctx, AXX_ERR_STATE, "walk must restart at 0 or continue at step %d", p->next_step);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        int rc = launch_walk(p, step_begin, step_end);
        if (rc) return rc;
        p->next_step = step_end;
        p->reduced = false;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_time_walk(axx_plan* p, int step_begin, int step_end, float* ms) {
    try {
        if (!p || !ms) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        int rc;
        if (step_begin == 0) {
            // placement first, outside the timed region: the time is the walk
            // kernel's alone
            if ((rc = axx_plan_walk(p, 0, 1))) return rc;
            step_begin = 1;
        }
        HIP_TRY(ctx, hipEventRecord(p->ev0, ctx->stream));
        rc = axx_plan_walk(p, step_begin, step_end);
        if (rc) return rc;
        HIP_TRY(ctx, hipEventRecord(p->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(p->ev1));
        HIP_TRY(ctx, hipEventElapsedTime(ms, p->ev0, p->ev1));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

RedArgs plan_red_args(const axx_plan* p) {
    RedArgs R{};
    R.state = p->d_state;
    R.N = p->N;
    R.P = p->Pt;
    R.M = p->M;
    R.n_steps = p->n_steps;
    R.relaxation = p->run.relaxation;
    R.chunk = p->chunk;
    R.n_chunks = p->n_chunks;
    R.rate = p->d_rate;
    R.wf_p = p->d_wfp;
    R.wf_ord = p->d_wford;
    R.ax = p->ax;
    R.wf_dir = p->d_wfdir;
    R.wf_scale = p->d_wfscale;
    R.S = p->d_S;
    R.partial = p->d_partial;
    R.counters = p->d_cnt;
    return R;
}

extern "C" int axx_plan_reduce(axx_plan* p) {
    try {
        if (!p) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        if (p->next_step != p->n_steps)
            return fail(ctx, AXX_ERR_STATE, "walk not finished (at step %d of %d)", p->next_step, p->n_steps);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const RedArgs R = plan_red_args(p);
        hipLaunchKernelGGL(zero_u64, dim3(1), dim3(64), 0, ctx->stream, p->d_cnt, 4);
        const int groups = (p->M + kRedJB - 1) / kRedJB;
        hipLaunchKernelGGL(reduce_kernel, dim3(p->n_chunks, groups), dim3(kRedThreads), 0, ctx->stream, R);
        HIP_TRY(ctx, hipGetLastError());
        const int tot = p->M * kRedVals;
        hipLaunchKernelGGL(finalize_kernel, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream,
                           (const double*)p->d_partial, p->M, p->n_chunks, p->d_out);
        HIP_TRY(ctx, hipGetLastError());
        p->reduced = true;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_sync(axx_plan* p) {
    try {
        if (!p) return AXX_ERR_INVALID;
        HIP_TRY(p->ctx, hipStreamSynchronize(p->ctx->stream));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// Error bits the kernels leave in a plan's error word: 1 = a placement by
// rejection did not finish, kErrWaitLimit = a walk task gave up waiting for
// its group's previous chunk.
static int walk_error(axx_ctx* ctx, int err) {
    if (err & kErrWaitLimit)
        return fail(ctx, AXX_ERR_DEVICE, "a walk task waited more than %.0f s for its walker group's previous chunk",
                    kWaitTicks * 1e-8);
    return fail(ctx, AXX_ERR_INIT, "rejection initialisation did not finish for some walkers");
}

extern "C" int axx_plan_results(axx_plan* p, axx_result* out) {
    try {
        if (!p || !out) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        if (!p->reduced) return fail(ctx, AXX_ERR_STATE, "results requested before reduce");
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        std::vector<double> h((size_t)p->M * kRedVals);
        unsigned long long cnt[4];
        int err = 0;
        HIP_TRY(ctx, hipMemcpyAsync(h.data(), p->d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(cnt, p->d_cnt, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (err) return walk_error(ctx, err);
        for (int j = 0; j < p->M; ++j) {
            const double* v = &h[(size_t)j * kRedVals];
            if (out->sum_cos) out->sum_cos[j] = v[0];
            if (out->sum_sin) out->sum_sin[j] = v[1];
            if (out->sum_cos2) out->sum_cos2[j] = v[2];
            if (out->sum_sin2) out->sum_sin2[j] = v[3];
            if (out->intra_cos) out->intra_cos[j] = v[4];
            if (out->intra_sin) out->intra_sin[j] = v[5];
        }
        out->n_walkers = p->N;
        out->collisions = cnt[0];
        out->transmissions = cnt[1];
        out->cap_hits = cnt[2];
        out->n_intra = (int64_t)cnt[3];
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// State is exchanged in the public layout (AXX_STATE_FIELDS(n_profiles));
// internally moments are kept for the padded profile width.
extern "C" int axx_plan_get_state(axx_plan* p, uint32_t* host) {
    try {
        if (!p || !host) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const size_t nf = AXX_STATE_FIELDS(p->P);
        HIP_TRY(ctx, hipMemcpy(host, p->d_state, nf * (size_t)p->N * sizeof(uint32_t), hipMemcpyDeviceToHost));
        int err = 0;
        HIP_TRY(ctx, hipMemcpy(&err, p->d_err, sizeof(int), hipMemcpyDeviceToHost));
        if (err) return walk_error(ctx, err);
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_set_state(axx_plan* p, const uint32_t* host) {
    try {
        if (!p || !host) return AXX_ERR_INVALID;
        axx_ctx* ctx = p->ctx;
        // the kernels index per-compartment tables with the compartment and
        // assume positions within a cell length of the cell and one common step
        {
            const size_t N = (size_t)p->N;
            const int* is = reinterpret_cast<const int*>(host);
            const float* fs = reinterpret_cast<const float*>(host);
            const int step = is[(size_t)F_STEP * N];
            if (step < 1 || step > p->n_steps) return fail(ctx, AXX_ERR_INVALID, "state: step %d", step);
            const float lo[3] = {-p->cell[0], -p->cell[1], -p->cell[2]};
            const float hi[3] = {2 * p->cell[0], 2 * p->cell[1], 2 * p->cell[2]};
            for (size_t w = 0; w < N; ++w) {
                const int c = is[(size_t)F_COMP * N + w], c0 = is[(size_t)F_COMP0 * N + w];
                if (c < 0 || c > p->n_prims || c0 < 0 || c0 > p->n_prims)
                    return fail(ctx, AXX_ERR_INVALID, "state: walker %zu compartment out of range", w);
                if (is[(size_t)F_STEP * N + w] != step)
                    return fail(ctx, AXX_ERR_INVALID, "state: walkers at different steps");
                for (int a = 0; a < 3; ++a) {
                    const float v = fs[(size_t)(F_X + a) * N + w];
                    if (!(v >= lo[a] && v <= hi[a])) return fail(ctx, AXX_ERR_INVALID, "state: walker %zu position", w);
                }
            }
        }
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        const size_t nf = AXX_STATE_FIELDS(p->P);
        HIP_TRY(ctx, hipMemcpy(p->d_state, host, nf * (size_t)p->N * sizeof(uint32_t), hipMemcpyHostToDevice));
        const int* is = reinterpret_cast<const int*>(host);
        p->next_step = is[(size_t)F_STEP * p->N];
        p->reduced = false;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_get_config(const axx_plan* p, int* block, int* wpt, double* bin_width, int* lds) {
    try {
        if (!p) return AXX_ERR_INVALID;
        if (block) *block = std::min(p->block, walk_max_block(p->Pt));   // as launched
        if (wpt) *wpt = p->wpt;
        if (bin_width) *bin_width = p->bin_width;
        if (lds) *lds = (int)p->lds_bytes;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_set_config(axx_plan* p, int block, int wpt) {
    try {
        if (!p) return AXX_ERR_INVALID;
        if (block < 64 || block > 1024 || (block % 64) != 0) return fail(p->ctx, AXX_ERR_INVALID, "block must be a multiple of 64 in [64, 1024]");
        if (wpt != 1) return fail(p->ctx, AXX_ERR_INVALID, "walkers per thread must be 1");
        p->block = block;
        p->wpt = wpt;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_set_schedule(axx_plan* p, int task_steps) {
    try {
        if (!p) return AXX_ERR_INVALID;
        if (task_steps < -1) return fail(p->ctx, AXX_ERR_INVALID, "task_steps must be -1, 0 or positive");
        p->sched = task_steps;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_get_schedule(const axx_plan* p, int* task_steps, int* last_task_steps, int* last_grid) {
    try {
        if (!p) return AXX_ERR_INVALID;
        if (task_steps) *task_steps = p->sched;
        if (last_task_steps) *last_task_steps = p->last_task_steps;
        if (last_grid) *last_grid = p->last_grid;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_plan_profile_count(const axx_plan* p) { return p ? p->P : AXX_ERR_INVALID; }

extern "C" int axx_plan_unit_cells(const axx_plan* p) { return p ? p->units : AXX_ERR_INVALID; }

// Unwrapped positions of the traced walkers into out[(j * n_steps + i) * 3].
static __global__ void trace_gather_kernel(const uint32_t* __restrict__ state, long long N,
                                           const uint32_t* __restrict__ ids, int k, int n_steps, int i,
                                           float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const float* fs = reinterpret_cast<const float*>(state);
    const size_t w = ids[j];
    for (int a = 0; a < 3; ++a)
        out[((size_t)j * n_steps + i) * 3 + a] = fs[(size_t)(F_X0 + a) * N + w] + fs[(size_t)(F_DX + a) * N + w];
}

extern "C" int axx_trace(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q, const axx_run* r,
                         const uint32_t* walker_ids, int k, float* xyz, uint32_t* counts) {
    try {
        if (!ctx || !s || !q || !r || !walker_ids || !xyz || k < 1) return fail(ctx, AXX_ERR_INVALID, "trace: bad arguments");
        for (int j = 0; j < k; ++j)
            if ((long long)walker_ids[j] >= r->n_walkers) return fail(ctx, AXX_ERR_INVALID, "trace: walker id %u", walker_ids[j]);
        axx_plan* p = nullptr;
        int rc = axx_plan_create(ctx, s, q, r, &p);
        if (rc) return rc;
        uint32_t* d_ids = nullptr;
        float* d_out = nullptr;
        const size_t n_out = (size_t)k * p->n_steps * 3;
        if (hipMalloc((void**)&d_ids, (size_t)k * sizeof(uint32_t)) != hipSuccess ||
            hipMalloc((void**)&d_out, n_out * sizeof(float)) != hipSuccess ||
            hipMemcpy(d_ids, walker_ids, (size_t)k * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
            if (d_ids) (void)hipFree(d_ids);
            if (d_out) (void)hipFree(d_out);
            axx_plan_destroy(p);
            return fail(ctx, AXX_ERR_NOMEM, "trace: device buffers");
        }
        const int blocks = (k + 255) / 256;
        for (int i = 0; !rc && i < p->n_steps; ++i) {
            rc = axx_plan_walk(p, i == 0 ? 0 : i, i + 1);
            if (!rc) {
                hipLaunchKernelGGL(trace_gather_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                                   (const uint32_t*)p->d_state, p->N, (const uint32_t*)d_ids, k, p->n_steps, i, d_out);
                if (hipGetLastError() != hipSuccess) rc = fail(ctx, AXX_ERR_DEVICE, "trace: gather launch");
            }
        }
        if (!rc && hipMemcpy(xyz, d_out, n_out * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ctx, AXX_ERR_DEVICE, "trace: copy back");
        if (!rc && counts) {
            std::vector<uint32_t> st((size_t)AXX_STATE_FIELDS(p->P) * (size_t)p->N);
            rc = axx_plan_get_state(p, st.data());
            for (int j = 0; !rc && j < k; ++j) {
                counts[j * 3 + 0] = st[(size_t)F_NCOLL * p->N + walker_ids[j]];
                counts[j * 3 + 1] = st[(size_t)F_NTRAN * p->N + walker_ids[j]];
                counts[j * 3 + 2] = st[(size_t)F_NCAP * p->N + walker_ids[j]];
            }
        }
        (void)hipFree(d_ids);
        (void)hipFree(d_out);
        axx_plan_destroy(p);
        return rc;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_run_phase_noalloc(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q,
                                     const axx_run* r, axx_result* out) {
    try {
        if (!ctx || !s || !q || !r || !out) return fail(ctx, AXX_ERR_INVALID, "null argument");
        if (!ctx->scratch) return fail(ctx, AXX_ERR_DEVICE, "no scratch address on device %d", ctx->device);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        DevArena ar{ctx->scratch, (size_t)AXX_NOALLOC_SCRATCH_BYTES, 0};
        axx_plan pl;
        pl.ctx = ctx;
        int rc = plan_create_impl(ctx, s, q, r, nullptr, &pl, 1, &ar);
        if (!rc) rc = launch_walk(&pl, 0, q->n_steps);
        if (!rc) {
            pl.next_step = q->n_steps;
            rc = axx_plan_reduce(&pl);
        }
        if (!rc) rc = axx_plan_results(&pl, out);
        if (rc) (void)hipStreamSynchronize(ctx->stream);   // nothing of this run left in flight
        return rc;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_run_phase(axx_ctx* ctx, const axx_substrate* s, const axx_acq* q, const axx_run* r,
                             axx_result* out) {
    try {
        if (!ctx || !out) return AXX_ERR_INVALID;
        axx_plan* p = nullptr;
        int rc = axx_plan_create(ctx, s, q, r, &p);
        if (rc) return rc;
        rc = axx_plan_walk(p, 0, q->n_steps);
        if (!rc) rc = axx_plan_reduce(p);
        if (!rc) rc = axx_plan_results(p, out);
        axx_plan_destroy(p);
        return rc;
    } catch (...) {
        return api_exception();
    }
}
