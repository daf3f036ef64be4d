/* atum_hip.h — C ABI of libatumhip.so, the gfx950 (MI355X) DGSEM tendency and
 * time-step library that stands in for Atum.jl's DGSEM / LSRK / ARK hot path.
 *
 * Plain C: opaque handles, POD descriptors, double / int64 / int32 pointers,
 * int status (0 = ok, < 0 = error; details from atm_last_error).  No C++
 * exceptions and no torch types cross this boundary, so Julia can `ccall` it
 * directly.
 *
 * Data layout (the reference's data model, Atum/Bennu field arrays):
 *   - a field with S components is S contiguous blocks of Np*K doubles; block s
 *     holds component s as a column-major (Np, K) array, node ijk = i + Nq*(j + Nq*k)
 *     (0-based here; src/dgsem.jl:247 uses the 1-based form);
 *   - metric g: dim*dim blocks, block (a + dim*b) = d xi_a / d x_b (Bennu metrics .g);
 *   - face arrays: (Nfp, K) with Nfp = 2*dim*Nq^(dim-1), grouped by face 2r+side
 *     (side 0: xi_r = -1), src/dgsem.jl:706-707; node indices are 0-based linear
 *     indices into the (Np, K) node array;
 *   - boundary tags: (2*dim, K), 0 = interior or periodic.
 * Field pointers passed to the compute entry points are DEVICE pointers (allocated
 * with atm_alloc or by any HIP allocator on the context's device); geometry
 * pointers in atm_dg_desc are HOST pointers that atm_dg_create copies.
 * All compute entry points are asynchronous on the context's stream; call
 * atm_sync before reading results on the host.
 */
#ifndef ATUM_HIP_H
#define ATUM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ATM_OK 0
#define ATM_ERR_ARG (-1)
#define ATM_ERR_UNSUPPORTED (-2)
#define ATM_ERR_HIP (-3)
#define ATM_ERR_NOMEM (-4)

/* balance laws (files under src/balancelaws/) */
#define ATM_LAW_ADVECTION 0
#define ATM_LAW_EULER 1
#define ATM_LAW_EULER_GRAVITY 2
#define ATM_LAW_LINEAR_EULER_GRAVITY 3
#define ATM_LAW_SHALLOW_WATER 4

/* numerical fluxes (src/numericalfluxes.jl:31-47) */
#define ATM_FLUX_CENTRAL 0
#define ATM_FLUX_RUSANOV 1
#define ATM_FLUX_EC 2
#define ATM_FLUX_KG 3
#define ATM_FLUX_ROE 4
#define ATM_FLUX_MATRIX 5

/* volume forms (src/dgsem.jl:5-14) */
#define ATM_FORM_WEAK 0
#define ATM_FORM_FLUX_DIFFERENCING 1

/* boundarystate plug-ins (enumerated; Julia closures cannot cross a C ABI) */
#define ATM_BC_COPY 0       /* src/balancelaw.jl:21 default */
#define ATM_BC_REFLECT 1    /* momentum-reflecting wall, experiment/euler_gravity/risingbubble.jl:10-15 */
#define ATM_BC_ZERO_INFLOW 2 /* advection inflow q+ = 0, test/advection/imex_wave_2d.jl:19-26 */
#define ATM_BC_PRESCRIBED 3 /* q+ = bc_states[tag-1], experiment/euler/sod.jl:20-23 */

/* problem source! plug-ins */
#define ATM_SOURCE_NONE 0
#define ATM_SOURCE_CORIOLIS 1 /* -2Ω x ρu (3-D), experiment/euler_gravity/baroclinicwave.jl:27-34 */

/* arithmetic order of the volume term:
 *  REFERENCE: the reference's summation order, every product and sum rounded
 *             (no FMA): agrees with the reference to round-off of log/sqrt;
 *  FAST:      flux differencing evaluates each symmetric pair once, log means
 *             by series, FMA contraction: same operator, different last bits
 *             (EC flux differencing of the Euler laws; other configurations
 *             run the REFERENCE kernels). */
#define ATM_PRECISION_REFERENCE 0
#define ATM_PRECISION_FAST 1

/* summation order of the flux-differencing volume term in REFERENCE precision:
 * the `kernel_type` of FluxDifferencingForm (src/dgsem.jl:8-14,135-212).
 * The three orders compute the same sum; their last bits differ.  NAIVE needs
 * all directions (as the reference asserts, src/dgsem.jl:139); FAST precision
 * has its own order and takes PER_DIR here. */
#define ATM_FD_PER_DIR 0           /* :per_dir,  src/dgsem.jl:470-561 */
#define ATM_FD_PER_DIR_SYMMETRIC 1 /* :per_dir_symmetric, src/dgsem.jl:563-681 */
#define ATM_FD_NAIVE 2             /* :naive,    src/dgsem.jl:308-468 */

/* time integrators */
#define ATM_LSRK54 0  /* src/odesolvers.jl:148-174 */
#define ATM_LSRK144 1 /* src/odesolvers.jl:176-229 */

typedef struct atm_ctx atm_ctx;
typedef struct atm_dg atm_dg;
typedef struct atm_banded atm_banded;

/* Everything a DGSEM operator needs: replaces the DGSEM constructor
 * `DGSEM(; law, grid, surface_numericalflux, volume_form, directions, auxstate)`
 * (src/dgsem.jl:38-56).  MJ, MJI and faceMJ are passed as computed there. */
typedef struct atm_dg_desc {
  int32_t law;            /* ATM_LAW_* */
  int32_t dim;            /* 1, 2 or 3 */
  int32_t N;              /* polynomial degree, Nq = N+1 LGL points per direction (1 <= N <= 7) */
  int32_t K;              /* number of elements */
  int32_t volume_form;    /* ATM_FORM_* */
  int32_t volume_flux;    /* two-point flux for flux differencing (ATM_FLUX_CENTRAL/EC/KG) */
  int32_t surface_flux;   /* ATM_FLUX_* */
  int32_t dirmask;        /* bit r set <=> direction r+1 in `directions` */
  int32_t bc;             /* ATM_BC_* */
  int32_t source;         /* ATM_SOURCE_* */
  int32_t pde_level_balance; /* EulerGravityLaw pde_level_balance flag */
  int32_t naux;           /* aux components per node (0 for laws without aux) */
  int32_t n_bc_states;    /* rows of bc_states (ATM_BC_PRESCRIBED) */
  int32_t precision;      /* ATM_PRECISION_* */
  int32_t fd_kernel;      /* ATM_FD_* (flux differencing in REFERENCE precision; else 0) */
  int32_t reserved;       /* 0 */
  double gamma, grav, omega;
  double u[3];            /* advection velocity */
  const double* bc_states;  /* n_bc_states x Ns, host */
  const double* D;          /* Nq x Nq column major: D[r + Nq*c] = l_c'(xi_r) */
  const double* g;          /* dim*dim x (Np*K) */
  const double* MJ;         /* Np*K */
  const double* MJI;        /* Np*K */
  const double* faceMJ;     /* Nfp*K */
  const double* normal;     /* dim x (Nfp*K) */
  const double* aux;        /* naux x (Np*K) */
  const int64_t* faceix_m;  /* Nfp*K */
  const int64_t* faceix_p;  /* Nfp*K */
  const int32_t* bctag;     /* 2*dim*K */
} atm_dg_desc;

/* ---- context ---------------------------------------------------------------
 * A context binds one device and one HIP stream (default priority).  Not
 * thread safe: use one context per host thread. */
int atm_ctx_create(int device, atm_ctx** out);
int atm_ctx_destroy(atm_ctx* ctx);
/* Launch on an existing HIP stream (e.g. the caller's current stream) instead of
 * the context's own; NULL restores the context's stream. */
int atm_ctx_set_stream(atm_ctx* ctx, void* hip_stream);
/* the hipStream_t the context launches on (for the caller's event timing) */
void* atm_ctx_get_stream(atm_ctx* ctx);
/* Additional streams on the context's device (default priority, non-blocking),
 * e.g. to run independent operators concurrently by switching the context's
 * stream with atm_ctx_set_stream; destroy waits for the stream's work. */
int atm_stream_create(atm_ctx* ctx, void** stream);
int atm_stream_destroy(atm_ctx* ctx, void* stream);
int atm_stream_sync(atm_ctx* ctx, void* stream);
/* Work enqueued on `waiter` after this call starts after the work enqueued on
 * `waitee` before it (an event, no host wait); NULL names the context's stream. */
int atm_stream_wait(atm_ctx* ctx, void* waiter, void* waitee);
/* the context's last error message; the string stays valid until the calling
 * thread's next atm_last_error */
const char* atm_last_error(atm_ctx* ctx);
int atm_sync(atm_ctx* ctx);
const char* atm_version(void);

/* ---- device memory ------------------------------------------------------- */
int atm_alloc(atm_ctx* ctx, size_t bytes, void** dptr);
int atm_free(atm_ctx* ctx, void* dptr);
int atm_upload(atm_ctx* ctx, void* dst_device, const void* src_host, size_t bytes);
int atm_download(atm_ctx* ctx, void* dst_host, const void* src_device, size_t bytes);
int atm_copy(atm_ctx* ctx, void* dst_device, const void* src_device, size_t bytes);

/* ---- DGSEM operator ------------------------------------------------------ */
/* Replaces DGSEM(...) (src/dgsem.jl:38-56).  Copies the geometry to the device. */
int atm_dg_create(atm_ctx* ctx, const atm_dg_desc* desc, atm_dg** out);
int atm_dg_destroy(atm_dg* dg);
/* number of state components Ns of the law (numberofstates, src/balancelaw.jl:14) */
int atm_dg_nstates(const atm_dg* dg);
/* ---- workgroup shape and autotuning ----------------------------------------
 * Elements per workgroup (epw) of the operator's element kernel: the default
 * fills 256 lanes with small elements and gives large ones a workgroup each.
 * atm_dg_set_epw changes it (results do not depend on it: every element is
 * computed by its own threads in the same order).  atm_dg_epw_candidates lists
 * the shapes that fit (whole waves, at most 512 threads -- 768 for the FAST
 * kernel's 3-D N = 6 elements with the Roe or Rusanov surface flux, whose
 * default is two elements per workgroup -- and the LDS); it returns their count.  atm_dg_autotune times `reps` tendency launches dq = L(q) per
 * candidate with HIP events on the context's stream, keeps the fastest (the
 * operator's current shape unless another is more than 2 % faster) and
 * returns it (per-candidate epw and microseconds per launch go to the optional
 * arrays of capacity max); dq is overwritten.  atm_dg_set_epw waits for the
 * context's stream; it must not be called while other streams (a serving
 * scheduler holding requests of this operator) may launch the operator.  No
 * counterpart in the reference (KernelAbstractions workgroup sizes are fixed in
 * src/dgsem.jl:74-76). */
int atm_dg_get_epw(const atm_dg* dg);
int atm_dg_set_epw(atm_dg* dg, int epw);
int atm_dg_epw_candidates(const atm_dg* dg, int32_t* out, int max);
int atm_dg_autotune(atm_dg* dg, const double* q, double* dq, int reps, int32_t* epw_out, float* us_out, int max);
/* Order in which the element kernels take the elements (order[i] = the element
 * processed i-th, a permutation of 0..K-1; NULL: element-id order).  Results do
 * not depend on it; a compact order (e.g. Morton) keeps a workgroup's face
 * neighbours in the same L2 (measured slower than element-id order on the
 * brick meshes here).  Same threading rule as atm_dg_set_epw.  No counterpart in
 * the reference. */
int atm_dg_set_schedule(atm_dg* dg, const int32_t* order, int64_t n);
/* total nodes Np*K */
int64_t atm_dg_nnodes(const atm_dg* dg);
/* device bytes the operator holds (geometry, aux, boundary states, reduction and
 * schedule buffers), as the serving scheduler counts it against its budget.  No
 * counterpart in the reference. */
int64_t atm_dg_device_bytes(const atm_dg* dg);

/* The rhs! protocol: replaces `(dg::DGSEM)(dq, q, time; increment)`
 * (src/dgsem.jl:59-98).  increment != 0: dq += L(q); else dq = L(q).
 * `time` is unused (autonomous laws), as in the reference. */
int atm_dg_tendency(atm_dg* dg, double* dq, const double* q, double time, int increment);

/* One fused low-storage RK stage (src/odesolvers.jl:50-53):
 *   dq = a*dq + L(q_in);  q_out = q_in + bdt*dq.
 * q_out must not alias q_in (neighbour elements read q_in); q_out may alias
 * nothing else.  a == 0 ignores the previous dq. */
int atm_lsrk_stage(atm_dg* dg, const double* q_in, double* q_out, double* dq, double a,
                   double bdt, double time);

/* Replaces dostep!(q, ::LSRK) for `nsteps` steps of size dt (src/odesolvers.jl:47-57).
 * q (in/out) and dq (in/out, solver storage) are device fields of Ns*Np*K doubles;
 * scratch is a device field of the same size used as the ping-pong state.
 * The whole run is launched as one stream of fused stage kernels; with
 * use_graph != 0 each step is captured once in a hipGraph and replayed. */
int atm_lsrk_run(atm_dg* dg, int scheme, double* q, double* dq, double* scratch, double dt,
                 double t0, int64_t nsteps, int use_graph);

/* atm_lsrk_run without the scratch argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The ping-pong state
 * lives in a fixed device array of the library (64 MiB, 8,388,608 doubles, held
 * on every device the library runs on until the process ends); the call
 * allocates nothing, launches the same stage kernels eagerly (no graph
 * instantiation) and gives bit-for-bit the result of atm_lsrk_run.  States of
 * more than 8,388,608 doubles (Ns*Np*K) are refused with ATM_ERR_ARG.  Every
 * call on one device shares that array: calls that overlap on one GPU (other
 * streams, other host threads, other contexts) race and must be ordered by the
 * caller. */
int atm_lsrk_run_noalloc(atm_dg* dg, int scheme, double* q, double* dq, double dt, double t0,
                         int64_t nsteps);

/* ---- banded vertical solve (src/bandedsystem.jl + Bennu banded LU / ldiv!) ---
 * Assembles A = I + scale*L for the linear vertical operator `lin` (its
 * dirmask must be the last direction only) by probing it with coloured unit
 * vectors, as Bennu.batchedbandedmatrix does, then factors A (LU, no pivoting).
 * One matrix per vertical node line: a line is a fixed horizontal node position
 * inside one column of KV elements; colelem[c*KV + ez] lists the elements of
 * column c bottom to top.  Unknowns are ordered m = s + Ns*(k + Nq*ez).
 * For ARK23, scale = -dt * (1 - 1/sqrt(2)) (src/odesolvers.jl:264-268). */
int atm_banded_create(atm_dg* lin, int32_t KV, const int32_t* colelem_host, int32_t ncols,
                      double scale, atm_banded** out);
int atm_banded_destroy(atm_banded* b);
/* device bytes the factorisation holds (factors, block form, index tables).  No
 * counterpart in the reference. */
int64_t atm_banded_device_bytes(const atm_banded* b);

/* Solve kernels of atm_banded_solve (all give the solution of the same factors;
 * the block forms agree with substitution to round-off):
 *   MFMA   block form on the matrix cores in 4x4 FP64 tiles, 16 lines per wave, 64 lines
 *          of one matrix per 4-wave workgroup sharing each block's operators in LDS (one
 *          wave per workgroup when that LDS does not fit) (Ns in {1,4,5}, Nq <= 6);
 *   BLOCK  the same block form on the vector units (Ns in {1,4,5}, Nq <= 6);
 *   REG    register-blocked substitution, 64 lines per wave (Ns in {1,4,5}, Nq <= 6);
 *   LDS    substitution with the line matrix staged in LDS (when it fits);
 *   PLAIN  substitution from global memory, one lane per line (always).
 * AUTO takes the first available in that order.  atm_banded_set_kernel fails with
 * ATM_ERR_UNSUPPORTED for a kernel that cannot solve this factorisation;
 * atm_banded_get_kernel returns the kernel a solve would use.  atm_banded_autotune
 * times `reps` solves x = A \ rhs with each available kernel (HIP events on the
 * context's stream), keeps the fastest and returns it (per-kernel ids and
 * microseconds per solve go to the optional arrays of capacity max); x is
 * overwritten.  No counterpart in the reference (Bennu's ldiv! has one form). */
#define ATM_BAND_AUTO (-1)
#define ATM_BAND_MFMA 0
#define ATM_BAND_BLOCK 1
#define ATM_BAND_REG 2
#define ATM_BAND_LDS 3
#define ATM_BAND_PLAIN 4
int atm_banded_set_kernel(atm_banded* b, int kernel);
int atm_banded_get_kernel(const atm_banded* b);
/* Waves per workgroup of the MFMA solve: 4 (64 lines of one matrix share each block's
 * operators in LDS) or 1 (16 lines); 0 = the largest whose LDS fits.  atm_banded_set_waves
 * fails with ATM_ERR_UNSUPPORTED for a shape that does not fit this factorisation;
 * atm_banded_get_waves returns the shape a solve would use (0: no MFMA solve here).  The
 * results are the same bits for either shape.  atm_banded_autotune times both shapes of the
 * MFMA solve, keeps the faster one and reports its time.  No counterpart in the reference. */
int atm_banded_set_waves(atm_banded* b, int waves);
int atm_banded_get_waves(const atm_banded* b);
int atm_banded_autotune(atm_banded* b, double* x, const double* rhs, int reps, int32_t* kernel_out, float* us_out,
                        int max);
/* x = A^{-1} b for every line (ldiv!, src/odesolvers.jl:355); x may alias b. */
int atm_banded_solve(atm_banded* b, double* x, const double* rhs);
/* Copy the factored band to the host: lines x n x (2*bw+1) in [line][row][bw+col-row]. */
int atm_banded_shape(const atm_banded* b, int64_t* nlines, int32_t* n, int32_t* bw);
int atm_banded_download(atm_banded* b, double* host);
/* Number of distinct line matrices kept (lines with bit-identical bands share one
 * factor, e.g. every column of a horizontally homogeneous mesh). */
int64_t atm_banded_nunique(const atm_banded* b);

/* ---- ARK23 (src/odesolvers.jl:231-388) -------------------------------------
 * Replaces ARK23(rhs!, linrhs!, q, dt; split_rhs, paperversion) + solve!.
 * `ex` may be NULL (no explicit operator); `band` must be created from `lin`
 * with scale = -dt_fac*(1 - 1/sqrt(2)).  work is a device buffer of
 * 9*Ns*Np*K doubles.  Runs nsteps steps of size dt. */
int atm_ark23_run(atm_dg* ex, atm_dg* lin, atm_banded* band, double* q, double* work, double dt,
                  double t0, int64_t nsteps, int split_rhs, int paperversion);

/* ---- batched instances ------------------------------------------------------
 * Many operators (instances of uneven size and N) advance together: members that
 * share a kernel (law, dim, N, forms, fluxes, precision) form a group whose
 * elements are laid out back to back in a flattened element table (per-member
 * first workgroup), and each group advances in ONE launch per stage; with
 * use_graph the two-step stage sequence of all groups is one hipGraph.  Kernel
 * families with a multi-degree kernel (weak-form shallow water with Rusanov
 * surface flux; the 2-D FAST Euler and Euler-gravity kernels with Roe or
 * Rusanov surface flux; N = 3..7) put all their degrees in one group, whose
 * workgroups have the size of its largest member's (smaller members take more
 * elements per workgroup).  Each member's elements per workgroup are taken when
 * the batch is created.  No counterpart in the reference (it runs one DGSEM per
 * solver). */
typedef struct atm_batch atm_batch;
int atm_batch_create(atm_ctx* ctx, int32_t n, atm_dg* const* members, atm_batch** out);
int atm_batch_destroy(atm_batch* b);
/* launches per stage (number of kernel groups) */
int atm_batch_ngroups(const atm_batch* b);
/* nsteps LSRK steps for every member i with its own dt[i]: q[i], dq[i], scratch[i] are
 * device fields of member i (host arrays of n device pointers). */
int atm_batch_lsrk_run(atm_batch* b, int scheme, double* const* q, double* const* dq,
                       double* const* scratch, const double* dt, int64_t nsteps, int use_graph);

/* ---- serving scheduler (continuous batching on one device) ------------------
 * Simulation requests are submitted from any host thread.  A worker thread of
 * the server admits them in submission order (while fewer than max_active are
 * resident and their device footprint fits hbm_budget: a request's own buffers
 * -- state, solver storage -- plus the device memory of the operators it runs
 * (atm_dg_device_bytes, atm_banded_device_bytes), an operator counted once while
 * any resident request runs it), advances every resident
 * request by a common chunk of steps per tick (the LSRK requests in one batched
 * launch per kernel group and stage as atm_batch_lsrk_run, the ARK23 requests on
 * side streams beside it), and retires the finished ones: the final state goes
 * to the request's host result buffer and its buffers return to a pool of idle
 * buffers that later admissions reuse (trimmed so that pool and resident requests
 * stay within hbm_budget).  A
 * request's result is bitwise that of its steps run alone with atm_lsrk_run /
 * atm_ark23_run.  The server launches on streams of its own; the context's
 * stream stays with the caller.  Work that a request's operators still need
 * (creation, factorisation) must be enqueued on the context's stream of the
 * submitting thread, or finished, before atm_server_submit.  The operators must
 * outlive their requests.  The worker captures no graphs (HIP would refuse other
 * threads' copies meanwhile), and no graph may be captured on the context while
 * the server holds requests.  No counterpart in the reference (it runs one solver
 * per process): these entry points serve many independent runs at once. */
typedef struct atm_server atm_server;
typedef struct {
  int32_t max_active;   /* requests resident at once, >= 1 */
  int32_t chunk;        /* steps per tick, >= 1 */
  int32_t imex_streams; /* side streams for ARK23 requests, 0..8 (0: the main stream) */
  int32_t reserved;     /* 0 */
  double hbm_budget;    /* bytes of device memory the resident requests may hold */
} atm_server_desc;

#define ATM_REQ_LSRK 0  /* atm_lsrk_run(dg, scheme, ...) */
#define ATM_REQ_ARK23 1 /* atm_ark23_run(dg, lin, band, ...) */

typedef struct {
  int32_t kind;         /* ATM_REQ_LSRK or ATM_REQ_ARK23 */
  int32_t scheme;       /* LSRK: ATM_LSRK54 or ATM_LSRK144 */
  int32_t split_rhs;    /* ARK23, as atm_ark23_run */
  int32_t paperversion; /* ARK23, as atm_ark23_run */
  atm_dg* dg;           /* LSRK: the operator; ARK23: the explicit operator or NULL */
  atm_dg* lin;          /* ARK23: the linear operator */
  atm_banded* band;     /* ARK23: the factors of lin for this dt */
  const double* q0;     /* host initial state, Ns*nnodes doubles (copied by submit) */
  double* result;       /* host buffer for the final state (NULL: discarded); written
                           before the request reports ATM_REQ_DONE */
  double dt;
  int64_t steps;        /* >= 0 */
} atm_request;

/* request states (negative: the request failed with that status) */
#define ATM_REQ_QUEUED 0
#define ATM_REQ_ACTIVE 1
#define ATM_REQ_DONE 2

typedef struct {
  int64_t submitted, completed, failed, ticks, max_active;
  double dof_updates; /* sum over steps taken of Ns*nnodes*stages (LSRK54 5, LSRK144 14, ARK23 3) */
  double busy_s;      /* wall time the worker spent in ticks */
  double gpu_span_s;  /* device time from the first to the last launch of each tick */
  double admit_s;     /* worker: allocating and uploading admitted requests */
  double enqueue_s;   /* worker: launching the ticks */
  double wait_s;      /* worker: waiting for the device */
  double retire_s;    /* worker: downloading results, freeing */
  int64_t pool_hits;   /* request buffers taken from the server's pool of idle buffers */
  int64_t pool_misses; /* request buffers allocated (hipMalloc) */
  double peak_bytes;   /* peak device bytes counted against hbm_budget (buffers + operators) */
  double peak_operator_bytes; /* the operators' share of it (each operator counted once) */
  double pool_bytes;   /* idle pooled buffers now (kept within hbm_budget minus what is resident) */
} atm_server_stats;

int atm_server_create(atm_ctx* ctx, const atm_server_desc* desc, atm_server** out);
/* finishes every submitted request, then stops the worker; no other thread may
 * be inside a call on this server (e.g. blocked in atm_server_wait) */
int atm_server_destroy(atm_server* s);
/* copies q0; validates the request (operators of this context, footprint within
 * the budget); *id identifies it from then on */
int atm_server_submit(atm_server* s, const atm_request* req, int64_t* id);
/* the request's state (ATM_REQ_*) or its negative error status */
int atm_server_poll(atm_server* s, int64_t id);
/* blocks until the request is done or failed; returns as atm_server_poll
 * (ATM_REQ_DONE if another thread released it meanwhile) */
int atm_server_wait(atm_server* s, int64_t id);
/* the next request that finished or failed and was not yet handed out here:
 * returns 1 and sets *id, or 0 when timeout_s (< 0: no limit) passes first */
int atm_server_next_done(atm_server* s, double timeout_s, int64_t* id);
/* blocks until every submitted request is done or failed */
int atm_server_drain(atm_server* s);
/* seconds on the server's clock (0 at creation) */
double atm_server_clock(atm_server* s);
/* submission, admission and completion times of a request (server clock; -1: not yet) */
int atm_server_times(atm_server* s, int64_t id, double* times3);
/* message of a failed request ("" otherwise); a copy owned by the calling
 * thread, valid until its next call of this function */
const char* atm_server_error(atm_server* s, int64_t id);
/* forgets a finished or failed request */
int atm_server_release(atm_server* s, int64_t id);
int atm_server_get_stats(atm_server* s, atm_server_stats* out);

/* ---- timing: HIP events recorded on the context's stream ------------------
 * slot in [0, 16); atm_event_elapsed_ms waits for the later event. */
int atm_event_record(atm_ctx* ctx, int slot);
int atm_event_elapsed_ms(atm_ctx* ctx, int slot_start, int slot_end, float* ms);

/* ---- elementwise helpers (broadcasts of the reference solvers) ------------ */
/* y = a*y + b*x over n doubles */
int atm_axpby(atm_ctx* ctx, double* y, const double* x, double a, double b, int64_t n);

/* ---- brick meshes built on the device ---------------------------------------
 * Replaces brickgrid(LobattoCell, coordinates; periodic, ordering) (Bennu,
 * called at test/euler/isentropicvortex.jl:43 and the experiment/ scripts) followed by
 * DGSEM(...) (src/dgsem.jl:38-56): the library builds the node coordinates,
 * metric terms, MJ, MJI, faceMJ, normals, face tables and boundary tags on the
 * device, bit-identical to the host builder.  Vertex coordinates, LGL nodes and
 * weights come from the host; in the operator description only D, aux and
 * bc_states are read (the other geometry pointers are ignored). */
typedef struct {
  int32_t Kd[3];            /* elements per dimension (dim entries used) */
  int32_t periodic[3];
  int32_t stacked;          /* element order: 0 first dimension fastest, 1 last (vertical) dimension fastest */
  const double* coords[3];  /* host arrays of Kd[d]+1 vertex coordinates */
  const double* xi;         /* host: N+1 LGL nodes */
  const double* w;          /* host: N+1 LGL weights */
} atm_brick_desc;

int atm_dg_create_brick(atm_ctx* ctx, const atm_dg_desc* d, const atm_brick_desc* b, atm_dg** out);

/* Copies one geometry array of an operator to host memory, in the layout the
 * operator description uses for it; POINTS (brick operators only) is dim
 * blocks of Np*K node coordinates. */
#define ATM_GEOM_POINTS 0
#define ATM_GEOM_G 1
#define ATM_GEOM_MJ 2
#define ATM_GEOM_MJI 3
#define ATM_GEOM_FACEMJ 4
#define ATM_GEOM_NORMAL 5
#define ATM_GEOM_FACEIX_P 6
#define ATM_GEOM_BCTAG 7
int atm_dg_geometry(atm_dg* dg, int which, void* host);

/* ---- snapshot stream ---------------------------------------------------------
 * The output hook of a time loop (solve!'s after_step callback,
 * src/odesolvers.jl:9-26; VTK/state output in experiment/*) without stalling
 * the compute stream.  atm_snapshot copies `bytes` from the device buffer src
 * to pinned host memory host_dst (from atm_host_alloc): through the device
 * staging buffer `staging` (same size) if given, so that the compute stream
 * continues while a second stream of the context does the device-to-host
 * copy; without staging the compute stream waits for the copy.  slot in
 * [0, 16) names the copy: atm_snapshot_query returns 1 once it has landed (0
 * before), atm_snapshot_wait blocks until then.  A slot's staging buffer may
 * be reused by the next atm_snapshot on that slot (ordered on the device);
 * host_dst must not be reused before the slot's copy has landed.  The query
 * and wait calls may come from another host thread. */
int atm_host_alloc(atm_ctx* ctx, size_t bytes, void** host);
int atm_host_free(atm_ctx* ctx, void* host);
int atm_snapshot(atm_ctx* ctx, int slot, void* host_dst, void* staging, const void* src, size_t bytes);
int atm_snapshot_query(atm_ctx* ctx, int slot);
int atm_snapshot_wait(atm_ctx* ctx, int slot);

/* ---- measures (src/measures.jl:6-33) ---------------------------------------
 * Reductions over all nodes on the device, in double-double arithmetic with a
 * fixed order (reproducible bit for bit); the result is copied to the host and
 * the call returns after the stream has reached it.
 * atm_entropy_integral: sum MJ*entropy(law, q, aux)   (entropyintegral(dg, q))
 * atm_entropy_product:  sum MJ*entropyvariables(law, p, aux).q
 *                                                     (entropyproduct(dg, p, q))
 * ATM_ERR_UNSUPPORTED for a law without entropy (LinearEulerGravity). */
int atm_entropy_integral(atm_dg* dg, const double* q, double* out);
int atm_entropy_product(atm_dg* dg, const double* p, const double* q, double* out);
/* weightednorm(dg, q, p; componentwise) (src/measures.jl:6-21): p > 0 or
 * INFINITY; out holds Ns values if componentwise, else one. */
int atm_weightednorm(atm_dg* dg, const double* q, double p, int componentwise, double* out);

/* ---- relaxation LSRK (src/odesolvers.jl:69-146) ----------------------------
 * Replaces RLSRK54/RLSRK144(rhs!, q, dt) + nsteps calls of dostep!.  Each step
 * runs the LSRK stages with the stage entropy production accumulated on the
 * device, then solves eta(q0 + gamma*k) - eta0 - gamma*deta = 0 for gamma by
 * Newton's method on the device, starting from *gamma (the previous step's;
 * 1 for a new solver), and sets q = q0 + gamma*k.  work is a device buffer of
 * 3*Ns*Np*K doubles (dq, q0, k); dq must be finite (zero for a new solver).
 * *gamma returns the last step's relaxation factor. */
int atm_rlsrk_run(atm_dg* dg, int scheme, double* q, double* work, double dt, double t0, int64_t nsteps,
                  double* gamma);

#ifdef __cplusplus
}
#endif
#endif /* ATUM_HIP_H */
