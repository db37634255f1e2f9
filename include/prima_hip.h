/* libprima_hip: MI355X (gfx950) backend for Prima's restarted-PDHG LP solver.
 *
 * A plain C ABI: opaque handles, plain pointers and sizes, int status codes
 * (0 = ok, negative = error; the message is kept per context and read with
 * prima_last_error). Nothing is thrown across the ABI. Host pointers are
 * borrowed only for the duration of a call.
 *
 * Two layers, matching the two seams of the reference (paths relative to
 * modules/prima-core/src/main/scala/org/noaidi/prima/):
 *
 *  1. The per-operation kernel seam, `trait Kernels` (kernels/Kernels.scala:
 *     42-125): allocate / upload / download / uploadMatrix / length / close
 *     plus the eight numeric operations. A context of precision 64 behaves
 *     like ScalaKernels (kernels/ScalaKernels.scala), one of precision 32 like
 *     Float32Kernels (kernels/Float32Kernels.scala: values narrowed on upload,
 *     infinities preserved, capabilities.supportsFloat64 = false).
 *
 *  2. The coarse on-device solve, which replaces the host loop of
 *     Pdhg.solveWith (Pdhg.scala:48-55), MixedPrecision.solve
 *     (MixedPrecision.scala:104-199) and the fan-out of PrimaZio.solveAll
 *     (modules/prima-zio/.../PrimaZio.scala:61-66): the whole restarted-PDHG
 *     iteration, KKT checkpoints, restarts and certificates run on the GPU.
 *
 * Threading: a context is not thread-safe and is used from one thread
 * (Kernels.scala:39-40); prima_cancel and prima_progress are the calls that
 * may come from another thread.
 */
#ifndef PRIMA_HIP_H
#define PRIMA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define PRIMA_OK 0
#define PRIMA_EINVAL (-1)   /* IllegalArgumentException in the reference (require) */
#define PRIMA_ENOMEM (-2)
#define PRIMA_EDEVICE (-3)
#define PRIMA_ECLOSED (-4)

/* SolveStatus (SolveResult.scala:9-26), same ordinals. */
#define PRIMA_OPTIMAL 0
#define PRIMA_PRIMAL_INFEASIBLE 1
#define PRIMA_DUAL_INFEASIBLE 2
#define PRIMA_ITERATION_LIMIT 3
#define PRIMA_TIME_LIMIT 4
#define PRIMA_NUMERICAL_ERROR 5
#define PRIMA_INTERRUPTED 6

typedef struct prima_ctx prima_ctx;
typedef struct prima_vec prima_vec;
typedef struct prima_mat prima_mat;

/* KernelCapabilities (kernels/Kernels.scala:13-22). */
typedef struct {
  char name[32];
  char device[64];
  int32_t supports_float64;
} prima_capabilities;

/* ---- context ------------------------------------------------------------ */

/* Open a context on a HIP device with iteration precision 64 or 32.
 * Replaces constructing a Kernels backend (ScalaKernels() /
 * Float32Kernels(), kernels/ScalaKernels.scala:125, Float32Kernels.scala:157). */
int prima_open(int32_t device, int32_t precision, prima_ctx** out);
/* Kernels.close (kernels/Kernels.scala:65-66): releases device resources;
 * idempotent. The handle stays valid (calls then return PRIMA_ECLOSED)
 * until prima_destroy. */
int prima_close(prima_ctx* ctx);
/* Frees the handle itself (implies close). */
void prima_destroy(prima_ctx* ctx);
/* Kernels.capabilities (kernels/Kernels.scala:47). */
int prima_capabilities_of(const prima_ctx* ctx, prima_capabilities* out);
/* Message for the last failing call on this context ("" if none). */
const char* prima_last_error(const prima_ctx* ctx);

/* ---- memory (kernels/Kernels.scala:51-63) -------------------------------- */

/* Kernels.allocate: a zero-filled vector of length n. */
int prima_vec_alloc(prima_ctx* ctx, int64_t n, prima_vec** out);
/* Kernels.upload: copies (never aliases) host data; narrows to float in a
 * precision-32 context with infinities preserved. */
int prima_vec_upload(prima_ctx* ctx, const double* data, int64_t n, prima_vec** out);
/* Kernels.download: into a caller-owned array of length prima_vec_length. */
int prima_vec_download(prima_ctx* ctx, const prima_vec* v, double* into, int64_t n);
/* Kernels.length. */
int64_t prima_vec_length(const prima_vec* v);
int prima_vec_free(prima_ctx* ctx, prima_vec* v);
/* Kernels.uploadMatrix from canonical CSR (SparseMatrix.scala:15-21): rows
 * of row_ptr ascending, column indices in range, values finite. */
int prima_mat_upload_csr(prima_ctx* ctx, int32_t rows, int32_t cols, int64_t nnz,
                         const int32_t* row_ptr, const int32_t* col, const double* vals,
                         prima_mat** out);
int prima_mat_free(prima_ctx* ctx, prima_mat* m);

/* SpMV schedules for prima_spmv (no counterpart in the reference, whose
 * SpMV is a sequential loop): sub-waves of 1..64 lanes per row, or the
 * row-block (CSR-Adaptive) kernel the on-device solver uses. */
#define PRIMA_SPMV_AUTO 0
#define PRIMA_SPMV_VECTOR_1 1
#define PRIMA_SPMV_VECTOR_64 7
#define PRIMA_SPMV_ROW_BLOCKS 8
/* Column panels: for matrices whose x is larger than an XCD's L2 (at least
 * 2^20 columns in float32, 2^19 in float64), prima_mat_upload_csr also
 * stores the entries cut into (row range, column panel) tiles so that the
 * gathers of x stay in L2. Only available on such matrices. */
#define PRIMA_SPMV_PANELS 9
/* The row-block kernel over small blocks (2,048 nonzeros, 1,024 rows), so
 * that several workgroups share a CU: for matrices of short rows. */
#define PRIMA_SPMV_ROW_BLOCKS_SMALL 10
/* Force a schedule for a matrix (PRIMA_EINVAL for an unknown schedule or a
 * panel schedule on a matrix without panels). */
int prima_mat_set_schedule(prima_ctx* ctx, prima_mat* a, int32_t schedule);
/* Autotuner: time every schedule on this matrix (reps launches each, HIP
 * events on the context stream) and keep the fastest for later prima_spmv
 * calls. Returns the choice and its time per SpMV in milliseconds. */
int prima_mat_autotune(prima_ctx* ctx, prima_mat* a, int32_t reps, int32_t* chosen,
                       double* best_ms);

/* ---- the eight operations (kernels/Kernels.scala:73-123) ---------------- */

/* Kernels.spmv: out := a * x. */
int prima_spmv(prima_ctx* ctx, const prima_mat* a, const prima_vec* x, prima_vec* out);
/* Kernels.axpby: out := alpha x + beta y; out may alias x or y. */
int prima_axpby(prima_ctx* ctx, double alpha, const prima_vec* x, double beta,
                const prima_vec* y, prima_vec* out);
/* Kernels.scale: out := alpha x. */
int prima_scale(prima_ctx* ctx, double alpha, const prima_vec* x, prima_vec* out);
/* Kernels.copy. */
int prima_copy(prima_ctx* ctx, const prima_vec* src, prima_vec* dst);
/* Kernels.dot; synchronises the context's stream. */
int prima_dot(prima_ctx* ctx, const prima_vec* x, const prima_vec* y, double* out);
/* Kernels.squaredNorm; synchronises the context's stream. */
int prima_squared_norm(prima_ctx* ctx, const prima_vec* x, double* out);
/* Kernels.primalStep: out := clamp(x - tau (cost - ktY), lower, upper). */
int prima_primal_step(prima_ctx* ctx, const prima_vec* x, const prima_vec* kty,
                      const prima_vec* cost, const prima_vec* lower, const prima_vec* upper,
                      double tau, prima_vec* out);
/* Kernels.dualStep: out := proj(y + sigma (rhs - kxBar)), first n_eq free. */
int prima_dual_step(prima_ctx* ctx, const prima_vec* y, const prima_vec* kx_bar,
                    const prima_vec* rhs, double sigma, int32_t n_eq, prima_vec* out);

/* ---- coarse on-device solve --------------------------------------------- */

/* LpProblem in standard form (LpProblem.scala:22-104): min c'x + offset,
 * K x = q on the first n_eq rows, K x >= q on the rest, l <= x <= u.
 * K is canonical CSR, m x n. */
typedef struct {
  int32_t n, m, n_eq;
  int32_t pad;
  int64_t nnz;
  const int32_t* row_ptr;
  const int32_t* col;
  const double* val;
  const double* c;
  const double* q;
  const double* l;
  const double* u;
  double offset;
} prima_lp;

/* PdhgParams + ScalingParams (PdhgParams.scala:9-59, Scaling.scala:11-24),
 * plus launch tuning that has no counterpart in the reference. */
typedef struct {
  double eps_abs;
  double eps_rel;
  int32_t max_iterations;
  int32_t evaluation_frequency;
  int64_t time_limit_ms; /* < 0: none */
  int32_t ruiz_iterations;
  int32_t apply_pock_chambolle;
  double pock_chambolle_alpha;
  double infeasibility_tolerance;
  double restart_sufficient;
  double restart_necessary;
  double restart_artificial;
  double primal_weight_smoothing;
  int32_t workgroups_per_lp; /* 0: chosen by the tuner */
  int32_t chunk_iterations;  /* iterations per launch; 0: adaptive */
} prima_params;

/* Pdhg.WarmStart (Pdhg.scala:72-77) in the caller's coordinates; a step
 * size or primal weight that is not finite and positive means "not given". */
typedef struct {
  const double* primal; /* n */
  const double* dual;   /* m */
  double step_size;
  double primal_weight;
} prima_warm;

/* LpSolution scalars (SolveResult.scala:44-65) with its KktError. */
typedef struct {
  int32_t status;
  int32_t iterations;
  int32_t restarts;
  int32_t used_average;
  double objective;
  double dual_objective;
  double primal_residual;
  double dual_residual;
  double solve_time_ms;  /* iteration time, after scaling and upload (Pdhg.scala:204) */
  double total_time_ms;  /* including scaling, transpose and upload */
  double final_step_size;
  double final_primal_weight;
  int64_t trials;        /* step-size trials, accepted and rejected */
  int32_t evaluations;   /* checkpoints */
  int32_t workgroups;    /* workgroups that owned this LP */
} prima_result;

/* MixedPrecision.Params (MixedPrecision.scala:37-76), plus the choice of
 * finishing pass. refinement = 0 is the reference contract: a float64 PDHG
 * pass warm-started from the float32 result. refinement = 1 finishes with an
 * FP64 outer iterative refinement instead (no counterpart in the reference;
 * Gleixner, Steffy & Wolter, INFORMS J. Comput. 28(3), 2016): residuals
 * b - Ax and c - A'y and the KKT test in float64, correction LPs on the same
 * device matrix solved by float32 PDHG, at most max_refinements rounds, then
 * the reference's float64 pass if the test still fails.
 * stall_iterations > 0 (no counterpart in the reference; 0 keeps its
 * contract) caps the warm-started float64 pass at that many iterations; an LP
 * still unconverged then gets a cold float64 pass with the rest of
 * max_iterations. Without a time limit that cold pass races the warm one:
 * it starts once the warm pass has run a quarter of the cap while the device
 * has workgroups to spare, and whichever concludes first (checked at launch
 * boundaries, which fall on the same iterations every run) gives the result.
 * A warm start can put the adaptive primal weight on a path that stalls for
 * hundreds of thousands of iterations where a cold start converges.
 * fix_threshold > 0 (iterative refinement only; no counterpart in the
 * reference): a variable at a bound whose scaled reduced cost holds it there
 * by more than fix_threshold is fixed in that round's correction LP. */
typedef struct {
  double reduced_tolerance;
  int32_t reduced_max_iterations;
  int32_t always_refine;
  double reduced_time_share;
  int32_t refinement;
  int32_t max_refinements;
  double inner_tolerance;
  int32_t inner_max_iterations;
  int32_t stall_iterations;
  double fix_threshold;
} prima_mixed_params;

/* MixedPrecision.Result scalars (MixedPrecision.scala:82-91); rounds and
 * fallback_iterations describe the iterative refinement (0 otherwise). */
typedef struct {
  int32_t reduced_iterations;
  int32_t reduced_restarts;
  int32_t refinement_iterations;
  int32_t refined;
  double reduced_ms;
  double refinement_ms;
  int32_t rounds;
  int32_t fallback_iterations;
} prima_mixed_info;

void prima_default_params(prima_params* p);
void prima_default_mixed_params(prima_mixed_params* p);

/* Pdhg.solveWith(problem, params, kernels, cancelled, warmStart)
 * (Pdhg.scala:48-55) on this context's precision, entirely on the device.
 * warm may be NULL; x_out (n), y_out (m), rc_out (n) may be NULL. */
int prima_solve(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                const prima_warm* warm, prima_result* res, double* x_out, double* y_out,
                double* rc_out);

/* No counterpart in the reference (its only GPU SpMV, SpmvSpike.scala:43-75,
 * has a static maxRowNnz cap): an autotuner for the solver's launch
 * thresholds on one matrix. It runs `iterations` accepted iterations of the
 * LP (no tolerance met, no certificate tested) at each candidate: group sizes
 * of half, once and twice the rule of about 1k nonzeros per workgroup, each
 * with the coarse row-block cut (TILE nonzeros, ROWCAP rows per block) and
 * with about one and two row blocks per workgroup of the group. The device
 * time per accepted iteration of each candidate (the kernel's HIP events)
 * goes into the result; the fastest is kept on this context for the
 * matrix (by content) and used by later solves: the group size of the LP
 * solved alone (and the most it is regrouped to in a batch), and the block
 * cut of its groups. Iteration counts to a tolerance can change with these
 * settings (the order of the global sums changes); results stay
 * reproducible for fixed settings. */
#define PRIMA_TUNE_MAX_CANDIDATES 9
typedef struct {
  int32_t workgroups;            /* chosen group size */
  int32_t blocks_per_workgroup;  /* chosen cut: 0 coarse, k about k blocks per workgroup */
  double us_per_iteration;       /* its device time per accepted iteration */
  int32_t candidates;            /* candidates timed (entries filled below) */
  int32_t cand_workgroups[PRIMA_TUNE_MAX_CANDIDATES];
  int32_t cand_blocks_per_workgroup[PRIMA_TUNE_MAX_CANDIDATES];
  double cand_us_per_iteration[PRIMA_TUNE_MAX_CANDIDATES];
} prima_tune_result;
int prima_tune_solver(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                      int32_t iterations, prima_tune_result* out);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc). The same solve
 * as prima_solve, for one LP, whose device memory (the LP's arena, view,
 * launch plan and state) is carved from a 64 MiB workspace the library holds
 * on each device it runs on, taken once per device when a context is opened
 * there; nothing is allocated on the device. An LP that needs more than the
 * workspace is rejected with PRIMA_EINVAL and a message giving both sizes.
 * Every call on a device shares that device's workspace (its contents are
 * cleared at the start of each call), so calls that overlap on one device
 * race: one context, one thread, one call at a time. */
int prima_solve_noalloc(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                        const prima_warm* warm, prima_result* res, double* x_out,
                        double* y_out, double* rc_out);

/* MixedPrecision.solve (MixedPrecision.scala:104-199). A precision-64
 * context skips the reduced pass (MixedPrecision.scala:111-113); a
 * precision-32 context runs the float32 pass on the device, re-checks a
 * certificate in double precision and finishes with a float64 device pass
 * warm-started from (x, y, step size, primal weight). */
int prima_mixed_solve(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                      const prima_mixed_params* mixed, prima_result* res, prima_mixed_info* info,
                      double* x_out, double* y_out, double* rc_out);

/* Independent LPs solved together (PrimaZio.solveAll, PrimaZio.scala:61-66):
 * each LP gets its own workgroups, step size, restarts and termination inside
 * shared persistent launches. Results keep input order. warms may be NULL or
 * hold NULL entries; x_outs/y_outs/rc_outs may be NULL or hold NULL entries.
 * mixed may be NULL (plain solve in the context's precision); otherwise each
 * LP follows prima_mixed_solve's two-pass contract. */
int prima_solve_batch(prima_ctx* ctx, int32_t count, const prima_lp* lps,
                      const prima_params* params, const prima_mixed_params* mixed,
                      const prima_warm* warms, prima_result* results, prima_mixed_info* infos,
                      double** x_outs, double** y_outs, double** rc_outs);

/* ---- branch-and-bound over device relaxations ---------------------------- */

/* MilpStatus (BranchAndBound.scala:7-31), same ordinals. */
#define PRIMA_MILP_OPTIMAL 0
#define PRIMA_MILP_FEASIBLE 1
#define PRIMA_MILP_INFEASIBLE 2
#define PRIMA_MILP_UNBOUNDED 3
#define PRIMA_MILP_NO_SOLUTION 4

/* BnbParams (BranchAndBound.scala:57-108) plus `frontier`: up to this many
 * open nodes are popped together and their relaxations solved in one batched
 * device call (they share K and differ only in bounds). frontier = 1 is the
 * reference's sequential depth-first search. mixed may be NULL (plain solves
 * in the context's precision). */
typedef struct {
  prima_params lp;
  double integrality_tolerance;
  double gap_abs;
  double gap_rel;
  double pruning_safety_factor;
  int32_t max_nodes;
  int32_t warm_start;
  int64_t time_limit_ms; /* < 0: none */
  int32_t frontier;
  int32_t pad;
} prima_bnb_params;

/* MilpSolution scalars (BranchAndBound.scala:33-55). has_incumbent = 0 is the
 * reference's empty `primal`. relaxations counts LPs solved, which exceeds
 * nodes_explored when a node solved in a batch was skipped afterwards because
 * an earlier node of the same batch raised the incumbent. */
typedef struct {
  int32_t status;
  int32_t nodes_explored;
  int32_t unproven_nodes;
  int32_t has_incumbent;
  int32_t relaxations;
  int32_t batches;
  double objective;
  double best_bound;
  double solve_time_ms;
} prima_milp_result;

void prima_default_bnb_params(prima_bnb_params* p);
/* BranchAndBound.solveWith(problem, integers, params, kernels)
 * (BranchAndBound.scala:168-392) with the node relaxations on the device.
 * x_out (n) receives the incumbent (integer coordinates snapped) when one
 * exists. */
int prima_milp_solve(prima_ctx* ctx, const prima_lp* lp, int32_t n_int, const int32_t* integers,
                     const prima_bnb_params* params, const prima_mixed_params* mixed,
                     prima_milp_result* res, double* x_out);
/* BranchAndBound.skipsBeforeSolving (BranchAndBound.scala:407-413). */
int prima_bnb_skips_before_solving(double safe_bound, double incumbent, int32_t have_incumbent);
/* BranchAndBound.safeBoundOf (BranchAndBound.scala:422-427). */
double prima_bnb_safe_bound_of(int32_t conclusive, double bound, double slack);

/* Launch statistics of this context's solver kernel, accumulated since the
 * last reset: launches and their total duration, timed with HIP events on
 * the context's stream. No counterpart in the reference (its only timing is
 * solveTimeMillis); used by the benchmark for the roofline figure. */
typedef struct {
  int64_t launches;
  double kernel_ms;
  int64_t accepted_iterations;
  int64_t trials;
  int64_t evaluations;
  double algorithmic_bytes; /* compulsory HBM bytes of the work done: every array
                             * counted once, no gather amplification, and a
                             * matrix that LPs of one call share counted once
                             * per pass over it */
  /* launches, duration and algorithmic bytes split by the kernel's iteration
   * precision (index 0: float32, 1: float64) */
  int64_t launches_by_precision[2];
  double kernel_ms_by_precision[2];
  double algorithmic_bytes_by_precision[2];
} prima_stats;
int prima_stats_get(const prima_ctx* ctx, prima_stats* out);
int prima_stats_reset(prima_ctx* ctx);

/* Cooperative cancellation (the `cancelled` callback of Pdhg.scala:43-47):
 * a running solve on ctx stops at its next launch boundary with status
 * PRIMA_INTERRUPTED. Safe to call from another thread. */
int prima_cancel(prima_ctx* ctx);

/* Progress of the solves on ctx since it was opened, updated at every launch
 * boundary: solver launches, accepted PDHG iterations summed over LPs, and
 * passes finished (a mixed-precision LP counts its float32 and float64
 * passes, and a cold pass that raced the warm one). No counterpart in the
 * reference (its solves report nothing
 * until they return); for callers that watch a long batch. Safe to call from
 * another thread while a solve runs. */
int prima_progress(const prima_ctx* ctx, int64_t* launches, int64_t* accepted_iterations,
                   int64_t* finished_lps);

/* Test hook with no counterpart in the reference. While enabled, every solve
 * on ctx whose LPs run on more than one workgroup has the last workgroup of
 * each group wait, with a short limit, for a word nothing writes: its bounded
 * wait gives up and sets the group's abort word, the group's other
 * workgroups give up at their next barrier, and every workgroup exits. The
 * solve call then fails with PRIMA_EDEVICE, as it does when a group barrier
 * times out for real (a group barrier giving up is a device failure, never a
 * SolveStatus). For tests of that error path; 0 turns it off. */
int prima_test_stall_barriers(prima_ctx* ctx, int32_t enable);

#ifdef __cplusplus
}
#endif
#endif
