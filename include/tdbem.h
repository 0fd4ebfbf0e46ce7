/*
 * tdbem.h — C ABI of the MI355X-native TD-BEM hot paths (libtdbem.so, built from td-bem_amd/csrc).
 *
 * The reference (XiaoyongBai/TD-BEM) has no FFI; its hot paths sit behind three C++ classes that
 * main.cpp drives (src/main.cpp:43-128).  Each entry point below names the member function(s) it
 * replaces.  Conventions:
 *   - every function returns TDBEM_OK (0) or a negative status; tdbem_last_error(ctx) gives the
 *     message (for a failed tdbem_create, tdbem_last_error(NULL) gives it), which the C++ host
 *     rethrows as `const char*` the way the reference reports errors (src/main.cpp:132-135);
 *   - ids are 0-based (the deck's 1-based ids are converted on reading, src/InputT.cpp:529-534);
 *     dof = 3*node + direction;
 *   - dense matrices are row-major doubles; "host" pointers are ordinary host memory;
 *   - one host thread per context; each context owns one HIP stream of default priority, so
 *     independent contexts can run concurrently on one GPU.
 */
#ifndef TDBEM_H
#define TDBEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdbem_context tdbem_context;

enum {
    TDBEM_OK = 0,
    TDBEM_ERR_ARG = -1,     /* bad argument (the reference's thrown const char* cases) */
    TDBEM_ERR_HIP = -2,     /* HIP runtime failure */
    TDBEM_ERR_STATE = -3,   /* call out of order */
    TDBEM_ERR_NUMERIC = -4  /* singular step-0 matrix */
};

/* Temporal kernels, GreenFunctionT::Compute(type) (src/GreenFunctionT.cpp:119-145). */
enum {
    TDBEM_KERNEL_HEAVISIDE = 2, /* TDElementHeavisideLinearT (assemble_step only) */
    TDBEM_KERNEL_STEP = 4,      /* TDElementStepLinearT, the one main.cpp runs
                                   (src/ElastoDynamicsT.cpp:27) */
    TDBEM_KERNEL_BSPLINE = 5,   /* TDElementRampHeaviLinearT: cubic B-spline pulse
                                   (src/TDElementRampHeaviLinearT.cpp:77-108; assemble_step only) */
    TDBEM_KERNEL_RAMP = 6       /* ramped-Heaviside pulse of slope 2/dt: Int_Stokes_RampedHeaviside_Linear
                                   and Int_Kelvin_RampedHeaviside_Linear (src/GreenFunctionT.cpp:
                                   589-622, 661-680), which no Compute case reaches, in the element
                                   loop of TDElementRampHeaviLinearT (assemble_step only) */
};

/* Problem set-up.  Replaces ElastoDynamicsT::ElastoDynamicsT + TDElementT::SetNQ/SetMaterial
 * (src/ElastoDynamicsT.cpp:21-43, src/TDElementT.cpp:100-139) and the SolverT constructor with
 * SetTimeControl and ComputeMaxStep (src/SolverT.cpp:23-51, :101-155, :211-250).
 * coords [nnd][3]; ien [nel][4 or 8]; etype 1 (4-node) or 2 (8-node); nq in {6,10,20,40}. */
int tdbem_create(int nnd, const double* coords, int nel, const int* ien, int etype, int nq,
                 double E, double nu, double rho, int if_ext, double dt, int numstep, int kernel,
                 tdbem_context** out);
/* Frees the context (TDBEM_OK for NULL).  A context that is a member of a batch is not freed:
 * TDBEM_ERR_STATE until it has left the batch (tdbem_batch_remove / tdbem_batch_destroy). */
int tdbem_destroy(tdbem_context* ctx);
const char* tdbem_last_error(const tdbem_context* ctx);

/* SolverT::GetMaxStep (src/SolverT.cpp:252-255). */
int tdbem_max_step(const tdbem_context* ctx);

/* ElastoDynamicsT::G_H_Driver(t = step*dt, t_1 = dt) followed by GetGHLinear
 * (src/ElastoDynamicsT.cpp:108-231, :283-289): dense row blocks of collocation nodes
 * [lower, upper], each [3(upper-lower+1)][3 nnd].  G1/H1 are fG_1/fH_1 (the phi_1 part, called
 * G_0/H_0 in main.cpp:100-105), G2/H2 are fG_2/fH_2 (phi_2, main.cpp's G_1/H_1).
 * The _device variant takes device pointers and does not synchronise. */
int tdbem_assemble_step(tdbem_context* ctx, int step, int lower, int upper, double* G1,
                        double* G2, double* H1, double* H2);
int tdbem_assemble_step_device(tdbem_context* ctx, int step, int lower, int upper, double* G1,
                               double* G2, double* H1, double* H2);
/* Allocation-free variant of tdbem_assemble_step_device (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments and results; it calls no allocating
 * routine and does not wait: the kernels are queued on the context's stream (tdbem_synchronize
 * waits for them).  The element slots of each row batch live in a 64 MiB device array of the
 * library, one per GPU, held until the process ends: a mesh whose one collocation row needs more
 * (more than 69 905 4-node or 34 952 8-node elements) is refused with TDBEM_ERR_ARG.  All calls
 * on one GPU share that array: calls that overlap on one GPU (contexts on different streams)
 * race, and it keeps the last call's contents. */
int tdbem_assemble_step_device_noalloc(tdbem_context* ctx, int step, int lower, int upper,
                                       double* G1, double* G2, double* H1, double* H2);

/* The lag matrices that SolverStepLinearT::UpdateGHLinear accumulates over steps
 * 1..min(numstep, MaxStep) (src/SolverStepLinearT.cpp:58-97), built on the device in one pass and
 * kept only inside each node pair's wavefront shell. */
int tdbem_build_lag_store(tdbem_context* ctx);
/* SolverStepLinearT::UpdateGHLinear(G_0, G_1, H_0, H_1) at fCurrStep = step
 * (src/SolverStepLinearT.cpp:58-97) with caller-supplied blocks: the rows of collocation nodes
 * lower..upper ([3 (upper - lower + 1)][3 nnd] each, host) of the step's phi_1 part (G0, H0:
 * main.cpp's G_0, H_0 = fG_1, fH_1) and phi_2 part (G1, H1).  As the reference does, the phi_2
 * blocks are inserted into lag 0 at step 1 and added to lag step - 1 after, the phi_1 blocks
 * inserted into lag step; steps past the kept lags (MaxStep, or numstep) are ignored.  The first
 * call replaces the store the device assembles itself (and drops a march on it) by a dense lag
 * database on the device (the reference's (MaxStep + 1) x 2 N x N footprint); the march builds
 * its shell store from it before each step that needs newer lags, so steps and updates may
 * alternate as in main.cpp:89-128.  March step s needs the blocks of step s + 1 (lag s complete),
 * march_init those of steps 1 and 2.  FP64 precision only; G's 3 x 3 node blocks must be
 * symmetric (to 1e-12): the store keeps their upper triangle.  tdbem_build_lag_store returns to
 * the device's own assembly. */
int tdbem_update_gh_linear(tdbem_context* ctx, int step, int lower, int upper, const double* G0,
                           const double* G1, const double* H0, const double* H1);
/* Bytes of the shell store (predicted when not yet built). */
int64_t tdbem_lag_store_bytes(tdbem_context* ctx);
/* Device bytes a solve of this context needs at nrhs right-hand sides in the current precision,
 * from the store planner's exact counts (node-pair lags, stored sub-tiles, slot batches) and the
 * march's allocations, without building anything: *resident while it marches (tables, store,
 * march state), *peak at any point of its set-up (with the build's node-pair tables and slot
 * buffers, or the dense step-0 matrices).  Admission control for serving (no counterpart in the
 * reference).  The device's memory pool may keep up to 16 GiB of freed set-up memory besides. */
int tdbem_predict_bytes(tdbem_context* ctx, int nrhs, int64_t* resident, int64_t* peak);
/* Free and total device memory after a device synchronisation with the memory pool trimmed
 * (hipMemGetInfo): what the predictions above are checked against. */
int tdbem_device_memory(int64_t* free_bytes, int64_t* total_bytes);
/* Dense N x N (N = 3 nnd) lag k of G and H on the host: fG_DB[k], fH_DB[k]. */
int tdbem_get_lag(tdbem_context* ctx, int k, double* G, double* H);
/* The rows of collocation nodes lower..upper of lag k, [3 (upper - lower + 1)][N] each: one
 * rank's row block of fG_DB[k], fH_DB[k] (rows fDofLow..fDofHigh, src/SolverT.cpp:128-140), as
 * WriteGH writes them per rank (src/SolverT.cpp:670-732). */
int tdbem_get_lag_rows(tdbem_context* ctx, int k, int lower, int upper, double* G, double* H);

/* SolverStepLinearT::Initialize/SetLoad/Solve_Averaging set-up (src/SolverStepLinearT.cpp:30-56,
 * :314-393; src/SolverT.cpp:344-377): alpha (a1, a2), displacement (UBC) and traction (FBC)
 * boundary data, and nrhs in {1,2,4,8} load histories f_r(t) = amp_r * min(1, t / t_rise_r)
 * (t_rise_r <= 0: Heaviside; amp = t_rise = NULL: the reference's f = 1).  Builds the lag store if
 * needed and factors the step-0 matrices. */
int tdbem_march_init(tdbem_context* ctx, double a1, double a2, int n_ubc, const int* ubc_dof,
                     const double* ubc_val, int n_fbc, const int* fbc_dof, const double* fbc_val,
                     int nrhs, const double* amp, const double* t_rise);
/* Advances up to nsteps of main.cpp's step loop (src/main.cpp:89-128: SetLoad + Solve +
 * RecordResult), stopping after the last step. */
int tdbem_march_run(tdbem_context* ctx, int nsteps);
/* Queues up to nsteps on the context's stream and returns without waiting (for running many
 * contexts concurrently); tdbem_query returns 1 when the queued work has finished, 0 while it runs.
 * No counterpart in the reference, whose steps are synchronous. */
int tdbem_march_enqueue(tdbem_context* ctx, int nsteps);
int tdbem_query(tdbem_context* ctx);
/* Index s of the last solved history vector x_s (0 before the first solve). */
int tdbem_march_step_index(const tdbem_context* ctx);
/* Histories of displacement and traction, x_0 .. x_{numstep-1} as RecordResult sees them
 * (src/SolverStepLinearT.cpp:549-603): U, T are host [numstep][3 nnd]. */
int tdbem_get_history(tdbem_context* ctx, int rhs, double* U, double* T);
/* One history vector x_s (u and t, host [3 nnd] each): fDis_DB[s], fTrac_DB[s] as RecordResult
 * gathers them (src/SolverStepLinearT.cpp:558-564). */
int tdbem_get_state(tdbem_context* ctx, int rhs, int s, double* u, double* t);

/* Arithmetic of the lag store and the history convolution for the next tdbem_build_lag_store /
 * tdbem_march_init (no counterpart in the reference, which is FP64 throughout):
 *   TDBEM_PRECISION_FP64 (default): the plain path, nrhs in {1, 2, 4, 8};
 *   TDBEM_PRECISION_MIXED_BF16X3: the lags are assembled in FP64 as on the plain path and
 *   rounded once into FP32 16 x 32 sub-tiles (only those a node pair reaches), each value split
 *   exactly into three BF16 terms in registers before its MFMA products, which are accumulated
 *   in FP32 over one (column tile, lag) step and in FP64 across steps, nrhs in {16, 32, 48, 64}.
 *   The FP64 store of the lags is not kept (only lags 0 and 1, for the step-0 matrices); the
 *   step-0 LU/inverse and the step solve stay FP64.
 * Changing the precision discards a lag store built with the other one.  tdbem_assemble_step
 * is always FP64. */
enum { TDBEM_PRECISION_FP64 = 0, TDBEM_PRECISION_MIXED_BF16X3 = 1 };
int tdbem_set_precision(tdbem_context* ctx, int precision);

/* Diagnostic: one v_mfma_f32_16x16x32_bf16 product D[16][16] = A[16][32] B[32][16] (inputs
 * rounded to bf16) through the operand lane layouts the mixed convolution assumes. */
int tdbem_mfma_selftest(const float* A, const float* B, float* D);

/* Diagnostic: one v_mfma_f64_16x16x4 product D[16][16] = A[16][4] B[4][16] through the operand
 * lane layouts of the step solve. */
int tdbem_mfma_f64_selftest(const double* A, const double* B, double* D);

/* Diagnostic: a plain streaming copy (k_stream_copy) of `bytes` bytes (a multiple of 16) between
 * two device buffers, `launches` times; *avg_ms = mean launch time from HIP events.  It reads and
 * writes exactly `bytes` per launch, so hardware byte counters of other kernels can be scaled by
 * what they read for it.  No counterpart in the reference. */
int tdbem_stream_copy(size_t bytes, int launches, double* avg_ms);

/* Diagnostic: a read-only stream of `bytes` bytes (a multiple of 16) with non-temporal loads,
 * `launches` times; *avg_ms = mean launch time from HIP events.  bytes / avg_ms is the read
 * bandwidth a streaming kernel reaches on this GPU, the practical ceiling of the HBM-bound
 * convolutions.  No counterpart in the reference. */
int tdbem_stream_read(size_t bytes, int launches, double* avg_ms);

/* Diagnostic: the step-0 dense solver on its own.  A is n x n row-major (host), B is n x nrhs
 * row-major (host), overwritten with X = A^{-1} B.  mode 0: the LU of the march (no row
 * interchanges when partial pivoting would make none, else partial pivoting) and two triangular
 * solves, as in the n == 2 solve; TDBEM_SOLVE_INVERSE: X = (A^{-1}) B through the explicit
 * inverse the march applies every step; TDBEM_SOLVE_PIVOTED: factor with partial pivoting in
 * every case.  Stands in for KSPSolve (src/SolverStepLinearT.cpp:503-504); TDBEM_ERR_NUMERIC
 * when A is singular. */
enum { TDBEM_SOLVE_INVERSE = 1, TDBEM_SOLVE_PIVOTED = 2 };
int tdbem_dense_solve(int n, const double* A, int nrhs, double* B, int mode);

/* Waits for all work queued on the context's stream. */
int tdbem_synchronize(tdbem_context* ctx);

/* Batched march (no counterpart in the reference, which runs one problem per process): many
 * independent contexts, each after a mixed-precision tdbem_march_init with the same nrhs, are
 * stepped together, one step of all of them per six kernel launches, captured as a hipGraph of
 * steps_per_graph steps that is replayed until every context reaches its own last step.  The
 * batch's members may change between runs (continuous batching): tdbem_batch_add admits a
 * context (solving its first step on its own stream if it has not started), tdbem_batch_remove
 * lets one leave; the next run or tune re-plans the work table and captures a new graph.  While
 * a context is a member, the calls that would free or re-initialise what the batch reads
 * (tdbem_destroy, tdbem_march_init / run / enqueue, tdbem_build_lag_store, tdbem_set_precision,
 * tdbem_set_max_step, tdbem_mrt_create) return TDBEM_ERR_STATE; its histories
 * (tdbem_get_history) hold the results.  tdbem_batch_create takes count >= 0 initial members.
 * tdbem_batch_tune times the tunables (workgroup rounds of the convolution, wavefronts of the
 * step solve) on the batch and keeps the fastest; report (may be NULL) receives {rounds, solve
 * waves, convolution ms, solve ms}.  tdbem_batch_run returns the number of graph launches in
 * *launches (may be NULL); max_replays <= 0 runs to the end; members that reach their last step
 * stay members (idle) until removed. */
typedef struct tdbem_batch tdbem_batch;
int tdbem_batch_create(tdbem_context* const* ctxs, int count, tdbem_batch** out);
int tdbem_batch_add(tdbem_batch* b, tdbem_context* ctx);
int tdbem_batch_remove(tdbem_batch* b, tdbem_context* ctx);
/* number of members */
int tdbem_batch_size(const tdbem_batch* b);
int tdbem_batch_tune(tdbem_batch* b, double* report);
int tdbem_batch_set_params(tdbem_batch* b, int rounds, int solve_waves);
int tdbem_batch_run(tdbem_batch* b, int steps_per_graph, int max_replays, int* launches);
int tdbem_batch_destroy(tdbem_batch* b);

/* SolverMRT::SetMaxStep (src/SolverMRT.cpp:1318-1321), for any context: the lag store holds
 * M_0 .. M_K (K capped at numstep), M_K without the part the step after K would add, as the
 * reference's lag database of fMaxStep + 1 matrices does (UpdateGHLinear, :850-873).  A change
 * drops a built store.  tdbem_create sets K to SolverT::ComputeMaxStep's value. */
int tdbem_set_max_step(tdbem_context* ctx, int max_step);

/* Two-region solver: SolverMRT (src/SolverMRT.cpp) over two region contexts (FP64, Step kernel,
 * the same dt and numstep), which must outlive it and are not marched on their own meanwhile.
 *   tdbem_mrt_create  — SolverMRT::SolverMRT with LocalIndexing / InterfaceIndexing (:22-60,
 *                       :160-568) and SetMaxStep(max_step) (the reference's default is 10).
 *                       iface1/iface2: the n_iface interface node ids (0-based) of each region;
 *                       region 2's are matched to region 1's by coordinates (1e-10), and
 *                       TDBEM_ERR_ARG "Interface mismatch, interface nodes must match
 *                       one-to-one" is returned when they do not.
 *   tdbem_mrt_init    — builds the region lag stores if needed and the global step matrix
 *                       (Assemble_GH_R1/R2, :685-850) and its inverse.  scheme:
 *                       TDBEM_MRT_DIRECT (Solve_Direct, :1035-1099, what SolverMRT::Solve runs)
 *                       or TDBEM_MRT_AVERAGING (Solve_Averaging, :1101-1249).  load_kind:
 *                       TDBEM_MRT_LOAD_PULSE, f(t) = 5.46366e9 t exp(-2000 t) of SetLoad
 *                       (:935-939; amp, t_rise unused), or TDBEM_MRT_LOAD_RAMP, f(t) =
 *                       amp min(1, t / t_rise) (amp when t_rise <= 0).  The loads are the
 *                       regions' traction conditions (fbc_dof per region, 0-based dofs) times
 *                       f(fCurrStep dt); conditions on interface dofs are ignored, as are
 *                       displacement conditions (SetLoad reads only the FBC lists).
 *   tdbem_mrt_run     — the next nsteps steps (UpdateGHLinear + SetLoad + Solve per step).
 *                       Solve_Direct solves x_1 .. x_numstep; Solve_Averaging solves x_s at
 *                       fCurrStep = s + 1, so x_1 .. x_{numstep-1}, with x_numstep left 0.
 *   tdbem_mrt_get_state — step s (0 .. numstep): x [n] the global unknowns fDis_DB_Global
 *                       (n = tdbem_mrt_num_unknowns = 3 (nnd1 + nnd2)); u [n] and t [n] the
 *                       regions' displacements (ExtractField, :1251-1301) and tractions, region
 *                       1's 3 nnd1 dofs then region 2's.  Any of x, u, t may be NULL.
 * Errors are read with tdbem_last_error(NULL). */
enum { TDBEM_MRT_DIRECT = 0, TDBEM_MRT_AVERAGING = 1 };
enum { TDBEM_MRT_LOAD_PULSE = 0, TDBEM_MRT_LOAD_RAMP = 1 };
typedef struct tdbem_mrt tdbem_mrt;
int tdbem_mrt_create(tdbem_context* region1, tdbem_context* region2, int max_step, int n_iface,
                     const int* iface1, const int* iface2, tdbem_mrt** out);
int tdbem_mrt_init(tdbem_mrt* m, int scheme, int load_kind, double amp, double t_rise,
                   int n_fbc1, const int* fbc_dof1, const double* fbc_val1, int n_fbc2,
                   const int* fbc_dof2, const double* fbc_val2);
int tdbem_mrt_run(tdbem_mrt* m, int nsteps);
int tdbem_mrt_step_index(const tdbem_mrt* m);
int tdbem_mrt_num_unknowns(const tdbem_mrt* m);
int tdbem_mrt_get_state(tdbem_mrt* m, int s, double* x, double* u, double* t);
/* host wall clock (ms) of tdbem_mrt_init and tdbem_mrt_run, summed (no reference counterpart) */
int tdbem_mrt_get_timings(const tdbem_mrt* m, double* init_ms, double* run_ms);
int tdbem_mrt_destroy(tdbem_mrt* m);

/* Element matrices of (collocation point, element) pairs of no mesh: the TDElementT interface
 * (src/TDElementT.h:16-115) as a reusable evaluator.  tdbem_element_create is the constructor
 * TDElementT(type) of the element class the kernel names (TDBEM_KERNEL_STEP:
 * TDElementStepLinearT, TDBEM_KERNEL_HEAVISIDE: TDElementHeavisideLinearT, TDBEM_KERNEL_BSPLINE
 * and TDBEM_KERNEL_RAMP: TDElementRampHeaviLinearT) with SetNQ(nq) and
 * SetMaterial(E, nu, density) (TDElementT.h:19-30); it holds a stream and buffers on the current
 * device that later calls reuse.  Errors: TDBEM_ERR_ARG for an element type, rule, kernel or
 * material the reference refuses. */
typedef struct tdbem_element tdbem_element;
int tdbem_element_create(int etype, int kernel, int nq, double E, double nu, double density,
                         tdbem_element** out);
int tdbem_element_destroy(tdbem_element* el);
/* Message of the element's last failed call; with el NULL, of the last failed create. */
const char* tdbem_element_last_error(const tdbem_element* el);

/* SetCoords(coords[p]), SetCollocation(collocation[p]), SetTime(t, t_1), then FormGH_Regular()
 * for singular[p] == -1 or FormGH_Singular(singular[p]), then GetGHLinear (TDElementT.h:22-52;
 * the call sequence of src/ElastoDynamicsT.cpp:139-192), for npairs pairs at once.  Host arrays:
 * coords [npairs][ennd][3], collocation [npairs][3], singular [npairs]; out GE_1, GE_2, HE_1, HE_2
 * [npairs][3][3 ennd] and C_1, C_2 [npairs][3][3], row-major as the element's own arrays (_1: the
 * phi_1 hat, _2: the phi_2 hat).  The Step element's C block is nonzero only at t == t_1 (its
 * first step); the Heaviside and pulse elements fill C for regular pairs and move it into the
 * singular node's HE columns for singular ones, as the reference's elements do.  As in the
 * reference, the singular node is taken from the caller and is not checked against the geometry.
 * TDBEM_ERR_ARG: a singular id outside -1 .. ennd-1 (the reference throws), t_1 <= 0, null
 * arrays.  Waits for its copies. */
int tdbem_element_form(tdbem_element* el, int64_t npairs, const double* coords,
                       const double* collocation, const int* singular, double t, double t_1,
                       double* GE_1, double* GE_2, double* HE_1, double* HE_2, double* C_1,
                       double* C_2);
/* The same on device arrays, queued on the element's stream without waiting (no counterpart in
 * the reference: for callers whose pairs are already in device memory).  out = GE_1 | GE_2 |
 * HE_1 | HE_2 | C_1 | C_2 back to back (npairs * (36 ennd + 18) doubles).  A pair whose
 * singular id is out of range gets NaN.  tdbem_element_synchronize waits for the queued work. */
int tdbem_element_form_device(tdbem_element* el, int64_t npairs, const double* coords,
                              const double* collocation, const int* singular, double t,
                              double t_1, double* out);
int tdbem_element_synchronize(tdbem_element* el);

/* Instrumentation (no counterpart in the reference, which has no timers): level
 * TDBEM_PROFILE_ALL (any nonzero level other than TDBEM_PROFILE_CONV) brackets the lag kernel,
 * the store gather, the history convolution, the reduction and the step solve with HIP events on
 * the context stream and counts the lag kernel's active points; TDBEM_PROFILE_CONV brackets only
 * the history convolution (two events per step instead of six), 0 turns it off.  tdbem_get_timings fills out[TDBEM_NUM_TIMINGS] in this order:
 * store_ms, init_ms, march_ms (host wall clock of build_lag_store / march_init / march_run),
 * lag_kernel_ms, gather_ms, conv_ms, solve_ms (summed event times), lag_launches,
 * conv_launches, active_points (quadrature-point x lag evaluations done by the lag kernel),
 * store_nonzero (stored (i, j, k) blocks before padding), store_padded, tile_subtiles (16 x 32
 * split-BF16 sub-tiles of the mixed-precision store, 0 until built), reduce_ms (the mixed path's
 * reduction of partial sums into the step solve's right-hand side, summed event times), tile_blocks (48 x 96 blocks with a stored
 * sub-tile), tile_entries (steps of the merged per-group lists), active_levels (time levels
 * evaluated at those points, one to three each), then, of the mixed-precision store, the
 * in-shell node-pair lags of the near lags (k < 3), of the far lags of even and of odd k, the
 * stored sub-tiles of the same three classes, and the entries of the merged lists of lag parity
 * 0 and 1 (the time-blocked convolution streams the near lags and the far lags of one parity at
 * each step).  conv_ms covers k_conv or k_conv_mixed alone. */
#define TDBEM_NUM_TIMINGS 25
#define TDBEM_PROFILE_ALL 1
#define TDBEM_PROFILE_CONV 2
int tdbem_set_profiling(tdbem_context* ctx, int level);
int tdbem_reset_timings(tdbem_context* ctx);
int tdbem_get_timings(tdbem_context* ctx, double* out);

#ifdef __cplusplus
}
#endif

#endif
