/*
 * axonexa.h — C ABI of the MI355X random-walk diffusion-MRI engine.
 *
 * The reference exposes this path as C++ templates that a driver instantiates
 * and runs in one call:
 *   phaseAcquisitionStream<F>::runAcquisitionLattice(num, basis*, Lattice&, dt,
 *                                                    blocks, threads, numOfSM)
 *   (src/acquisition/phaseAcquisitionStream.cuh:126-190)
 * which allocates device buffers, seeds the RNG (src/kernels/kernelSetup.cuh:1-5),
 * runs updateWalkersLatticePhase (src/kernels/kernelLattice.cuh:1-104), reduces
 * cos/sin of the phases (src/util/reduction.cuh:76-195) and copies Mx/My back
 * into the acquisition (phaseAcquisitionStream.cuh:185-189).
 *
 * This header is that boundary restated as a C ABI: plain structs, plain
 * pointers, negative error codes, no C++ or HIP types.  Every entry point
 * below names the reference routine it replaces.  Host buffers are owned by
 * the caller; device memory is owned by the context/plan.
 *
 * Units are the caller's (the reference uses mm, ms, T with GAMMA in
 * rad/(ms*T)); the engine only needs them to be consistent.
 */
#ifndef AXONEXA_H
#define AXONEXA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference aborts via safe_cuda, src/util/misc.cuh:5-11;
 *      this ABI returns instead) ---- */
#define AXX_OK               0
#define AXX_ERR_INVALID     -1  /* bad argument / descriptor                 */
#define AXX_ERR_DEVICE      -2  /* HIP runtime error                         */
#define AXX_ERR_NOMEM       -3  /* device or host allocation failed          */
#define AXX_ERR_LIMIT       -4  /* size beyond a compiled limit              */
#define AXX_ERR_GEOMETRY    -5  /* overlapping primitives / step >= half cell */
#define AXX_ERR_INIT        -6  /* rejection initialisation did not finish   */
#define AXX_ERR_NODEVICE    -7  /* no HIP device                              */
#define AXX_ERR_STATE       -8  /* call out of order (e.g. results before run) */

/* ---- primitive types (src/primitives/cylinderXY.cuh, Sphere.cuh, plane.cuh) ---- */
#define AXX_PRIM_CYLINDER_XY 0  /* infinite cylinder along z */
#define AXX_PRIM_SPHERE      1
#define AXX_PRIM_PLANE       2  /* infinite reflecting plane (center = a point on it) */

/* ---- boundary policies (kernelLattice.cuh:64-79, boundaryCheck.cuh:34-113) ---- */
#define AXX_BOUNDARY_SPECULAR  0  /* multi-reflection, impermeable (SPECULAR_REFLECTION)   */
#define AXX_BOUNDARY_PERMEABLE 1  /* reflect or transmit with p = 4*kappa/v (USE_PERMEABLE) */
#define AXX_BOUNDARY_REJECTION 2  /* stay put if the region changes (simple rejection)    */

/* ---- initial placement (kernelLattice.cuh:29-37) ---- */
#define AXX_INIT_UNIFORM      0  /* uniform in the cell (Lattice::initializeUniformly)          */
#define AXX_INIT_COMPARTMENT  1  /* rejection until compartment == init_arg (initializeInRegion) */
#define AXX_INIT_INSIDE_PRIM  2  /* rejection inside the bounding cube of primitive init_arg
                                    (basis[k].randUnif, USE_INITIALIZE_IN_REGION = k+1)           */

/* Limits compiled into the kernels. */
#define AXX_MAX_PROFILES   64   /* distinct waveform time profiles per run      */
/* Boundary events allowed in one step before it is cut short (counted in
 * axx_result.cap_hits).  The reference loops without bound
 * (boundaryCheck.cuh:45,79); a walker sliding along the inside of a surface
 * at a grazing angle a needs about L / (2 R a) reflections for a step of
 * length L, so the cap only ends such grazing runs. */
#define AXX_DEFAULT_BOUNCE_CAP 4096
#define AXX_MAX_PRIMS      32768

/*
 * Substrate: a periodic cell a x b x c (src/substrate/lattice.cuh:3-23) holding
 * n_prims primitives (the reference's basis[] array).  Primitive k is the
 * reference's Cylinder_XY(cx, cy, R, T2, T1, D, region, permeability) or
 * Sphere(cx, cy, cz, R, ...).  Compartment ids: 0 = the lattice (extra-
 * cellular) space, k+1 = inside primitive k.  Primitives must not overlap
 * (under the minimum-image convention).  T2 <= 0 or +inf means no decay.
 * Planes (reference Plane, plane.cuh:1-40) take a point (center) and a
 * normal; they bound no compartment, reflect always and ignore radius, D,
 * T2 and kappa.
 */
typedef struct axx_substrate {
    double a, b, c;          /* cell lengths                                    */
    double D_lat;            /* diffusivity of the lattice space (length^2/time)*/
    double T2_lat;           /* T2 of the lattice space (time)                   */
    int n_prims;
    const int*    type;      /* [n_prims] AXX_PRIM_*                             */
    const double* center;    /* [n_prims][3] (z ignored for cylinders)           */
    const double* radius;    /* [n_prims]                                        */
    const double* D;         /* [n_prims] diffusivity inside                     */
    const double* T2;        /* [n_prims] T2 inside                              */
    const double* kappa;     /* [n_prims] membrane permeability (length/time); may be NULL */
    const double* normal;    /* [n_prims][3] plane normals; may be NULL without planes */
} axx_substrate;

/*
 * Acquisition: the reference's phaseAcquisition<F> (src/acquisition/
 * phaseAcquisition.cuh:1-52) holds M gradient functors G_j(t) evaluated at
 * t_i = i*dt.  Here each waveform j is amplitude_j * profile_{p(j)}(t_i) *
 * direction_j, with the unit-amplitude profiles sampled once on the host
 * (profiles[i*n_profiles + p], i in [0, n_steps)).  Phase of walker w:
 *   phi_j = gamma*dt * sum_{i=0}^{n_steps-1} G_j(t_i) . r_unbounded_w(t_i)
 * exactly as kernelLattice.cuh:38-44,81-85.
 */
typedef struct axx_acq {
    int n_steps;             /* steps per walk (SimuParamsPhase::steps)          */
    double dt;               /* time step                                        */
    double gamma;            /* gyromagnetic ratio (GAMMA)                       */
    int n_profiles;          /* 1 .. AXX_MAX_PROFILES                            */
    const float* profiles;   /* [n_steps][n_profiles]                            */
    int n_waveforms;         /* M                                                */
    const int*    wf_profile;   /* [M] profile index                             */
    const double* wf_amplitude; /* [M] gradient amplitude                        */
    const double* wf_dir;       /* [M][3] gradient direction                     */
} axx_acq;

typedef struct axx_run {
    int64_t  n_walkers;      /* N (any N >= 1; no multiple-of-block rule)        */
    uint64_t seed;           /* Philox key                                       */
    uint32_t instance;       /* independent stream selector (replicas, batch)    */
    uint32_t walker_offset;  /* global id of walker 0 (sharding)                 */
    int init_mode;           /* AXX_INIT_*                                       */
    int init_arg;
    int boundary_mode;       /* AXX_BOUNDARY_*                                   */
    int relaxation;          /* 1: weight by exp(-sum dt/T2) (USE_RELAXATION)    */
    int bounce_cap;          /* max boundary events per step (0 -> AXX_DEFAULT_BOUNCE_CAP) */
    int step_mode;           /* AXX_STEP_*                                       */
} axx_run;

/* ---- step generators (kernelLattice.cuh:52-57; kernelPhase.cuh:31-35;
 *      kernelWC.cuh:56-67) ---- */
#define AXX_STEP_ISOTROPIC 0  /* length sqrt(6 D dt), direction uniform on the sphere   */
#define AXX_STEP_AXES      1  /* +-sqrt(2 D dt) on each axis, random signs              */
#define AXX_STEP_GAUSSIAN  2  /* N(0, 2 D dt) on each axis (Box-Muller); the cell must
                                 exceed 9.5 sqrt(6 D dt), the longest pair of steps     */

/*
 * Results: the reference returns raw Mx_j = sum_w cos(phi_j,w)*e^{-T2f_w} and
 * My_j likewise (phaseAcquisitionStream.cuh:171-189).  Sums of squares give
 * the Monte-Carlo standard error.  "intra" sums cover walkers that started
 * inside a primitive (compartment != 0 at t = 0).  Any pointer may be NULL.
 */
typedef struct axx_result {
    double* sum_cos;         /* [M] */
    double* sum_sin;         /* [M] */
    double* sum_cos2;        /* [M] */
    double* sum_sin2;        /* [M] */
    double* intra_cos;       /* [M] */
    double* intra_sin;       /* [M] */
    int64_t n_walkers;
    int64_t n_intra;
    uint64_t collisions;     /* boundary events (reflections + transmissions)  */
    uint64_t transmissions;
    uint64_t cap_hits;       /* steps truncated by bounce_cap                   */
} axx_result;

/*
 * Walker state, structure-of-arrays, n_walkers entries per field (4 bytes
 * each).  Exposed for parity checks and checkpoint/resume.  Field f of walker
 * w lives at state[f*n_walkers + w].  Moments: AXX_F_MOM + 3*p + axis.
 */
#define AXX_F_X      0   /* wrapped position                                */
#define AXX_F_Y      1
#define AXX_F_Z      2
#define AXX_F_DX     3   /* displacement since t=0 (r_unbounded - r0)       */
#define AXX_F_DY     4
#define AXX_F_DZ     5
#define AXX_F_X0     6   /* initial position                                */
#define AXX_F_Y0     7
#define AXX_F_Z0     8
#define AXX_F_E      9   /* accumulated sum dt/T2 up to the last compartment change */
#define AXX_F_LF    10   /* fraction of step AXX_F_LI at the last compartment change */
#define AXX_F_COMP  11   /* int: current compartment                        */
#define AXX_F_COMP0 12   /* int: compartment at t = 0                       */
#define AXX_F_LI    13   /* int: step index of the last compartment change  */
#define AXX_F_NCOLL 14   /* uint: boundary events                           */
#define AXX_F_NTRAN 15   /* uint: transmissions                             */
#define AXX_F_NCAP  16   /* uint: bounce-cap truncations                    */
#define AXX_F_STEP  17   /* int: next step index to run                     */
#define AXX_F_BUD   18   /* float: transmission budget left: a 2 Exp(1) draw
                          * less the hazards of the permeable boundary events
                          * since the last transmission                     */
#define AXX_F_MOM   19
#define AXX_STATE_FIELDS(P) (AXX_F_MOM + 3 * (P))

/* ---- context ---- */
typedef struct axx_ctx axx_ctx;
typedef struct axx_plan axx_plan;

/* Select the device and create the context's stream.  One context per
 * device; a context is used by one host thread at a time. */
int  axx_create(axx_ctx** out, int device);
void axx_destroy(axx_ctx* ctx);
const char* axx_strerror(int code);
/* Human-readable detail of the last error on this context (or "" ). */
const char* axx_last_error(const axx_ctx* ctx);

/*
 * One-shot run: the whole of runAcquisitionLattice
 * (phaseAcquisitionStream.cuh:126-190) — upload, walk, reduce, download.
 * Synchronous.  Replaces runAcquisitionLattice / runAcquisition (single pore:
 * a cell larger than the pore with AXX_INIT_INSIDE_PRIM).
 */
int axx_run_phase(axx_ctx* ctx, const axx_substrate* s, const axx_acq* acq,
                  const axx_run* run, axx_result* out);

/*
 * axx_run_phase without device memory allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Every
 * device buffer of the run (substrate tables, profile table, walker state,
 * reduction partials) is carved from one static scratch array of
 * AXX_NOALLOC_SCRATCH_BYTES per GPU, held by the process from the library's
 * load until it ends; host memory for the preparation of the substrate
 * tables is still allocated.  A run whose buffers exceed the scratch returns
 * AXX_ERR_LIMIT (axx_last_error says so).  The scratch keeps the last run's
 * contents; runs on one GPU that overlap (two threads, or two contexts of
 * one device) race.  Results equal axx_run_phase's bit for bit.
 */
#define AXX_NOALLOC_SCRATCH_BYTES 268435456ull   /* 256 MiB */
int axx_run_phase_noalloc(axx_ctx* ctx, const axx_substrate* s, const axx_acq* acq,
                          const axx_run* run, axx_result* out);

/*
 * Trace (parity hook): the walk of axx_run_phase, step by step, with the
 * unwrapped positions (start + displacement, fp32) of k chosen walkers after
 * every step: xyz[(j * n_steps + i) * 3 + a] for walker walker_ids[j]
 * (index within the run) after step i (i = 0: the start), and counts[j * 3 +
 * 0..2] = its collisions, transmissions and bounce-cap hits at the end.  The
 * same kernel walks all run.n_walkers walkers (one launch per step), so the
 * traced walkers are those of the full run.  No reference counterpart: the
 * reference offers no way to observe a walker.
 */
int axx_trace(axx_ctx* ctx, const axx_substrate* s, const axx_acq* acq, const axx_run* run,
              const uint32_t* walker_ids, int k, float* xyz, uint32_t* counts);

/*
 * Plan API: the same run split into its parts so that the device-resident
 * walk can be timed, chunked (checkpoint/resume) and inspected.
 *   create  : validate, build periodic images + cell bins, upload tables,
 *             allocate walker state.             (phaseAcquisitionStream.cuh:134-158)
 *   walk    : advance every walker over steps [step_begin, step_end);
 *             step_begin is 0 (initialise the walkers; also restarts a
 *             finished run) or the step where the previous call stopped.
 *                                                  (kernelLattice.cuh:1-104)
 *   reduce  : fp64 signal sums into device buffers  (reduction.cuh:76-195)
 *   results : copy sums to the host (synchronises)  (phaseAcquisitionStream.cuh:185-189)
 * walk/reduce are asynchronous on the context stream.
 */
int  axx_plan_create(axx_ctx* ctx, const axx_substrate* s, const axx_acq* acq,
                     const axx_run* run, axx_plan** out);

/* Launch tuning (no reference counterpart: the reference takes blocks,
 * threads and numOfSM from the driver, phaseAcquisitionStream.cuh:127-129).
 * Zero fields select the defaults. */
typedef struct axx_tuning {
    int    block;       /* threads per block: multiple of 64 in [64, 1024]; 0 = the
                         * size (256, 512 or 1024) of highest occupancy */
    double bin_width;   /* cell-bin edge length; 0 = automatic               */
    int    lds_budget;  /* LDS bytes the substrate tables may use; 0 = 150 KiB */
    int    whole_cell;  /* 1: walk in the substrate's cell as given; 0: a substrate
                         * whose primitives repeat along x and y is walked in its
                         * unit cell (placement AXX_INIT_UNIFORM or compartment 0)
                         * -- the same physics; positions wrap in the unit cell,
                         * a walker inside a primitive carries the compartment of
                         * the lowest-numbered primitive of its translation class */
} axx_tuning;
int  axx_plan_create_ex(axx_ctx* ctx, const axx_substrate* s, const axx_acq* acq,
                        const axx_run* run, const axx_tuning* tune, axx_plan** out);
int  axx_plan_walk(axx_plan* plan, int step_begin, int step_end);
int  axx_plan_reduce(axx_plan* plan);
int  axx_plan_results(axx_plan* plan, axx_result* out);
int  axx_plan_sync(axx_plan* plan);
/* Copy walker state to / from the host (n_walkers * AXX_STATE_FIELDS(P) words). */
int  axx_plan_get_state(axx_plan* plan, uint32_t* host_state);
int  axx_plan_set_state(axx_plan* plan, const uint32_t* host_state);
/* Walk over [step_begin, step_end) and time the walk kernel with HIP events on
 * the plan's stream (from step 0 the placement runs first, untimed). */
int  axx_plan_time_walk(axx_plan* plan, int step_begin, int step_end, float* ms);
/* Launch configuration the plan uses (threads per block, walkers per thread,
 * bin width, LDS bytes) and a setter used by the autotuner. */
int  axx_plan_get_config(const axx_plan* plan, int* block, int* wpt,
                         double* bin_width, int* lds_bytes);
int  axx_plan_set_config(axx_plan* plan, int block, int wpt);
/* Walk schedule (no reference counterpart: the reference launches one thread
 * per walker for the whole walk, phaseAcquisitionStream.cuh:163-167,
 * kernelLattice.cuh:12,46).  A walk
 * launch is a persistent grid whose waves take tasks -- one group of 64
 * walkers over one time chunk of task_steps steps -- every group's chunk k
 * before any group's chunk k + 1, so every wave stays busy to the end of the
 * walk.  task_steps: 0 = automatic (chunks when the groups fill the resident
 * waves twice or more, one chunk per walk otherwise), -1 = the whole walk as
 * one chunk, > 0 = that many steps (rounded up to even).  Results are the same
 * bits for every schedule.  get: the setting, and the steps per task and
 * workgroups of the last walk launch. */
int  axx_plan_set_schedule(axx_plan* plan, int task_steps);
int  axx_plan_get_schedule(const axx_plan* plan, int* task_steps, int* last_task_steps,
                           int* last_grid);
int  axx_plan_profile_count(const axx_plan* plan);
/* Unit cells of the substrate's cell the plan walks in one of (1: the whole
 * cell; see axx_tuning.whole_cell).  No reference counterpart: the reference
 * scans the whole basis (lattice.cuh:184-247). */
int  axx_plan_unit_cells(const axx_plan* plan);
void axx_plan_destroy(axx_plan* plan);

/*
 * Batched runs: many substrates and acquisitions of uneven walker counts in
 * one persistent launch, captured together with placement and reduction into
 * one hipGraph.  No reference counterpart beyond the loop over acquisitions
 * of phaseAcquisitionStream<F>::runAcquisitionStreamLattice
 * (phaseAcquisitionStream.cuh:425-613); every instance gives bitwise the
 * walkers and sums of its own axx_run_phase.  Limits: at most
 * AXX_BATCH_MAX_PROFILES profiles per instance, n_walkers < 2^31; every
 * boundary policy and step generator.  Instances keep their own seed / instance id / walker
 * offset, so replicas stay independent.
 */
#define AXX_BATCH_MAX_PROFILES 32
typedef struct axx_instance {
    const axx_substrate* substrate;
    const axx_acq* acq;
    const axx_run* run;
} axx_instance;
typedef struct axx_batch axx_batch;
/* Build every instance's tables and state, the tile list and the graph.
 * tune: block (threads, <= 512) and lds_budget (default 72 KiB per instance)
 * are honoured; block 0 takes the cached autotuned choice or 256. */
int  axx_batch_create(axx_ctx* ctx, int n_inst, const axx_instance* inst,
                      const axx_tuning* tune, axx_batch** out);
int  axx_batch_run(axx_batch* b);                    /* launch the graph (async)   */
int  axx_batch_time(axx_batch* b, float* ms);        /* run once, HIP-event timed  */
int  axx_batch_sync(axx_batch* b);
int  axx_batch_results(axx_batch* b, int i, axx_result* out);   /* synchronises */
int  axx_batch_get_state(axx_batch* b, int i, uint32_t* host_state);
/* Time a probe walk of probe_steps steps for each (threads per block,
 * walkers per thread) candidate, keep the fastest, re-capture the graph and
 * cache the choice (per process; also appended to the file named by the
 * AXX_TUNE_CACHE environment variable, read at the first batch). */
int  axx_batch_autotune(axx_batch* b, int probe_steps, int* best_block, int* best_wpt,
                        float* best_ms);
int  axx_batch_set_config(axx_batch* b, int block, int wpt);
int  axx_batch_get_config(const axx_batch* b, int* block, int* wpt, int* grid, int* n_tiles);
void axx_batch_destroy(axx_batch* b);
/* One-shot: create, run, copy every instance's results (out[n_inst]), destroy. */
int  axx_run_batch(axx_ctx* ctx, int n_inst, const axx_instance* inst, axx_result* out);

/*
 * Learned surrogate: a BF16 multilayer perceptron on MFMA (FP32 accumulate)
 * with hand-written training kernels.  No reference counterpart.  dims[0]
 * inputs, dims[n_layers] outputs; hidden layers use GELU (tanh form), the
 * last is linear.  Parameters are packed per layer as W[out][in] (row
 * major) then b[out], fp32.  Loss: mean squared error over all outputs.
 */
typedef struct axx_mlp axx_mlp;
int  axx_mlp_create(axx_ctx* ctx, int n_layers, const int* dims, int max_batch, axx_mlp** out);
int  axx_mlp_param_count(const axx_mlp* m, int64_t* n);
int  axx_mlp_set_params(axx_mlp* m, const float* params);    /* also resets Adam */
int  axx_mlp_get_params(axx_mlp* m, float* params);
/* y[n][out] = f(x[n][in]) (host buffers, n <= max_batch). */
int  axx_mlp_forward(axx_mlp* m, const float* x, int n, float* y);
/* Re-run the forward pass over the inputs already on the device (the last
 * axx_mlp_forward's) and time it with HIP events. */
int  axx_mlp_time_forward(axx_mlp* m, int n, float* ms);
/* Mean loss of (x, t) and its gradient w.r.t. the packed parameters. */
int  axx_mlp_loss_grad(axx_mlp* m, const float* x, const float* t, int n, float* loss, float* grad);
/* Training set kept on the device; Adam (b1 0.9, b2 0.999, eps 1e-8) over
 * `steps` minibatches of `batch` rows (a multiple of 512) drawn with Philox. */
int  axx_mlp_set_data(axx_mlp* m, const float* x, const float* t, int n);
int  axx_mlp_train(axx_mlp* m, int steps, int batch, float lr, uint64_t seed, float* loss_trace);
void axx_mlp_destroy(axx_mlp* m);

/*
 * Magnetization (Bloch) path: magAcquisitionStream<F>::runAcquisition
 * (magAcquisitionStream.cuh:37-103) running updateWalkersMag
 * (kernelMag.cuh:1-141).  The walkers move exactly as in axx_run_phase with
 * the same substrate and run (same random numbers, so the same trajectories);
 * each carries one magnetization vector per field, integrated in double
 * precision at t_i = i dt, i = 0 .. n_steps-1, with the field at its
 * position after step i (its start for i = 0):
 *   AXX_MAG_ROTATE  exact rotation about B by -gamma |B| h per substep
 *                   (updateMag_noT2_rotate / _multistep, blochdiff.cuh:150-210)
 *   AXX_MAG_RK4     classical Runge-Kutta on dM/dt = gamma M x B, with
 *                   -(Mx/T2, My/T2, (Mz-1)/T1) when run.relaxation
 *                   (updateMagRK4, blochdiff.cuh:24-67)
 *   AXX_MAG_DOPRI   Dormand-Prince fifth-order steps, same equation
 *                   (updateMagDORP, blochdiff.cuh:68-148; fixed substeps)
 * Step i is split into substeps[i] substeps of h = dt / substeps[i].  Fields
 * are affine in the position, B(r, t) = b0(t) + J(t) (r - origin), r the
 * unwrapped position; samples[] holds, step after step and for each sample
 * time of the step, one record of 12 doubles (b0[3], J[3][3] row major) per
 * field.  Sample times of step i, t_i = i dt:
 *   ROTATE  t_i + k h (k < s_i);  RK4  t_i + k h/2 (k <= 2 s_i);
 *   DOPRI   t_i + (k + c) h (k < s_i; c = 0, 1/5, 3/10, 4/5, 8/9, 1),
 * i.e. s_i, 2 s_i + 1 and 6 s_i samples (axx_mag_sample_count).
 * Fields are walked in groups of AXX_MAG_GROUP; every group walks the same
 * walkers.  T1: per compartment (lattice, then primitives), NULL = none;
 * T2 comes from the substrate.  Results are the sums over walkers of Mx, My,
 * Mz per field (magAcquisition::getMx/getMy/getMz).
 */
#define AXX_MAG_ROTATE 0
#define AXX_MAG_RK4    1
#define AXX_MAG_DOPRI  2
#define AXX_MAG_GROUP  8
typedef struct axx_field {
    int           n_steps;
    double        dt;
    double        gamma;
    int           n_fields;
    int           integrator;   /* AXX_MAG_*                                  */
    const int*    substeps;     /* [n_steps], each in [1, 4096]; NULL = all 1 */
    const double* samples;      /* see above                                  */
    double        origin[3];
    double        m0[3];        /* magnetization at t = 0                     */
    const double* T1;           /* [n_prims + 1] or NULL                      */
} axx_field;

typedef struct axx_mag_result {
    double*  mx;                /* [n_fields] sums over walkers (or NULL)     */
    double*  my;
    double*  mz;
    int64_t  n_walkers;
    uint64_t collisions, transmissions, cap_hits;
} axx_mag_result;

typedef struct axx_mag axx_mag;
/* Number of field samples the table must hold (per field). */
int  axx_mag_sample_count(int integrator, int n_steps, const int* substeps, int64_t* n_samples);
int  axx_mag_create(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                    axx_mag** out);
/* Steps [step_begin, step_end) of every group; 0 places the walkers and sets
 * M = m0 first (same chunking rules as axx_plan_walk). */
int  axx_mag_walk(axx_mag* m, int step_begin, int step_end);
int  axx_mag_time_walk(axx_mag* m, int step_begin, int step_end, float* ms);
/* Walk schedule of every group's launches: as axx_plan_set_schedule (no
 * reference counterpart). */
int  axx_mag_set_schedule(axx_mag* m, int task_steps);
int  axx_mag_results(axx_mag* m, axx_mag_result* out);   /* synchronises */
/* Walker state of one group: AXX_F_MOM + 6 * F words per walker (F fields
 * in the group); M of field j, component c is the double whose low word is
 * AXX_F_MOM + 6 j + 2 c and high word the next one. */
int  axx_mag_get_state(axx_mag* m, int group, uint32_t* host_state);
void axx_mag_destroy(axx_mag* m);
int  axx_run_mag(axx_ctx* ctx, const axx_substrate* s, const axx_field* f, const axx_run* r,
                 axx_mag_result* out);

/*
 * Many magnetization runs in one persistent launch: the acquisitions of
 * magAcquisitionStream<F>::runAcquisitionStream (magAcquisitionStream.cuh:
 * 105-342).  Every run keeps its own substrate, fields, seed, instance id and
 * walkers, and gives bitwise the walkers and sums of its own axx_run_mag.
 * Limits: all runs of a batch carry the same number of fields (1 to
 * AXX_MAG_GROUP) and the same integrator; n_walkers < 2^31 per run.
 */
typedef struct axx_mag_instance {
    const axx_substrate* substrate;
    const axx_field*     field;
    const axx_run*       run;
} axx_mag_instance;
typedef struct axx_mag_batch axx_mag_batch;
int  axx_mag_batch_create(axx_ctx* ctx, int n_inst, const axx_mag_instance* inst, axx_mag_batch** out);
int  axx_mag_batch_run(axx_mag_batch* b);                /* async: placement, walk, sums */
int  axx_mag_batch_time(axx_mag_batch* b, float* ms);    /* run once, HIP-event timed   */
int  axx_mag_batch_results(axx_mag_batch* b, int i, axx_mag_result* out);   /* synchronises */
int  axx_mag_batch_get_state(axx_mag_batch* b, int i, uint32_t* host_state);
void axx_mag_batch_destroy(axx_mag_batch* b);
int  axx_run_mag_batch(axx_ctx* ctx, int n_inst, const axx_mag_instance* inst, axx_mag_result* out);

/* Library version string. */
const char* axx_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AXONEXA_H */
