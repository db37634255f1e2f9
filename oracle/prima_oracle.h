/* CPU oracle for the Prima restarted-PDHG path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * double-precision CPU path (ScalaKernels + Pdhg + Scaling + Kkt +
 * Certificates), of its float32 emulation (Float32Kernels), of its two-pass
 * MixedPrecision driver, of LpBuilder/SparseMatrix canonicalisation and of the
 * LpFixtures instances (including a java.util.Random restatement so the
 * seeded ladder instances regenerate bit-for-bit). It is the checker for the
 * HIP backend and the CPU baseline timed beside it; the product never links or
 * calls it.
 *
 * Reference files restated (paths relative to the reference repository,
 * modules/prima-core/src/main/scala/org/noaidi/prima/ unless noted):
 *   kernels/ScalaKernels.scala, kernels/Float32Kernels.scala, Pdhg.scala,
 *   PdhgParams.scala, Scaling.scala, SparseMatrix.scala, Kkt.scala,
 *   Certificates.scala, MixedPrecision.scala, LpProblem.scala,
 *   SolveResult.scala, ../../../../test/scala/org/noaidi/prima/LpFixtures.scala
 *
 * Pinning (which checks hold each function; see tests/test_oracle.py):
 *   - java.util.Random + fdlibm log: the reference ladder's published
 *     objectives, iteration and restart counts (NOTES.md:43-58) only come out
 *     if the regenerated instances are bit-identical; fdlibm log is also
 *     compared against libm over random inputs (<= 1 ulp).
 *   - CSR build / transpose / spmv: SparseMatrixSuite KATs and dense brute
 *     force on random tiny matrices.
 *   - Scaling: ScalingSuite invariants (row/col inf-norms, round-trip).
 *   - Kkt / Certificates: hand-built directions with exact residuals
 *     (CertificatesSuite), closed-form KKT on constructed optima.
 *   - Pdhg: LpFixtures hand-derived optima and statuses, the ladder's
 *     iteration/restart counts, closed-form optima of constructed LPs and the
 *     HiGHS objectives committed under tests/golden/.
 *   - Float32 kernels: KernelsSuite float32 KATs (1+2^-30 -> 1, inf
 *     survives, 1e300 -> inf) and NOTES.md:228-233 (random-600x400 reaches
 *     1e-5 in 1,792 float32 iterations).
 *   - MixedPrecision: MixedPrecisionSuite behaviours (fp64 device skips the
 *     reduced pass; certificate re-check; refined answer is fp64-converged).
 */
#ifndef PRIMA_ORACLE_H
#define PRIMA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* SolveStatus (SolveResult.scala:9-26), same ordinal order. */
enum {
  ORC_OPTIMAL = 0,
  ORC_PRIMAL_INFEASIBLE = 1,
  ORC_DUAL_INFEASIBLE = 2,
  ORC_ITERATION_LIMIT = 3,
  ORC_TIME_LIMIT = 4,
  ORC_NUMERICAL_ERROR = 5,
  ORC_INTERRUPTED = 6
};

/* PdhgParams + ScalingParams (PdhgParams.scala:9-59, Scaling.scala:11-24). */
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
} orc_params;

/* LpSolution scalars (SolveResult.scala:44-65) and KktError (Kkt.scala:8-13). */
typedef struct {
  int32_t status;
  int32_t iterations;
  int32_t restarts;
  int32_t pad;
  double objective;
  double dual_objective;
  double primal_residual;
  double dual_residual;
  double solve_time_ms;
  double final_step_size;
  double final_primal_weight;
} orc_result;

/* MixedPrecision.Params (MixedPrecision.scala:37-76). */
typedef struct {
  double reduced_tolerance;
  int32_t reduced_max_iterations;
  int32_t always_refine;
  double reduced_time_share;
} orc_mixed_params;

typedef struct {
  int32_t reduced_iterations;
  int32_t reduced_restarts;
  int32_t refinement_iterations;
  int32_t refined;
  double reduced_ms;
  double refinement_ms;
} orc_mixed_info;

typedef struct orc_lp orc_lp;

void orc_default_params(orc_params* p);
void orc_default_mixed_params(orc_mixed_params* p);

/* LpProblem.apply (LpProblem.scala:62-104) from canonical CSR, copied. */
orc_lp* orc_lp_from_csr(int32_t n, int32_t m, int32_t n_eq, const int32_t* row_ptr,
                        const int32_t* col, const double* val, const double* c,
                        const double* q, const double* l, const double* u, double offset);
/* SparseMatrix.fromTriplets (SparseMatrix.scala:204-260) + LpProblem.apply. */
orc_lp* orc_lp_from_triplets(int32_t n, int32_t m, int32_t n_eq, int64_t count,
                             const int32_t* rows, const int32_t* cols, const double* vals,
                             const double* c, const double* q, const double* l, const double* u,
                             double offset);
void orc_lp_free(orc_lp* lp);
void orc_lp_dims(const orc_lp* lp, int32_t* n, int32_t* m, int32_t* n_eq, int64_t* nnz);
void orc_lp_export(const orc_lp* lp, int32_t* row_ptr, int32_t* col, double* val, double* c,
                   double* q, double* l, double* u, double* offset);

/* LpFixtures (LpFixtures.scala:25-165) by name; NULL if unknown. */
orc_lp* orc_fixture(const char* name);
/* LpFixtures.randomFeasible (LpFixtures.scala:174-213). */
orc_lp* orc_random_feasible(int64_t seed, int32_t n, int32_t n_eq, int32_t n_ineq, double density);

/* java.util.Random restatement, exposed for known-answer tests. */
typedef struct {
  int64_t seed;
  int32_t have_next_next_gaussian;
  double next_next_gaussian;
} orc_jrandom;
void orc_jrandom_init(orc_jrandom* r, int64_t seed);
int32_t orc_jrandom_next_int(orc_jrandom* r, int32_t bound);
double orc_jrandom_next_double(orc_jrandom* r);
double orc_jrandom_next_gaussian(orc_jrandom* r);
double orc_fdlibm_log(double x);

/* SparseMatrix helpers (SparseMatrix.scala:36-64, 76-86, 141-181). */
void orc_spmv(const orc_lp* lp, int32_t transpose, const double* x, double* out);
double orc_spectral_norm_bound(const orc_lp* lp);

/* Scaling.apply (Scaling.scala:77-155): row and column scale factors and the
 * scaled matrix values (same sparsity as the input). */
void orc_scaling(const orc_lp* lp, int32_t ruiz_iterations, int32_t apply_pc, double alpha,
                 double* row_scale, double* col_scale, double* scaled_vals);

/* Kkt.evaluate (Kkt.scala:68-125); kx/kty may be NULL. out[4] =
 * {primal residual, dual residual, primal objective, dual objective}. */
void orc_kkt_evaluate(const orc_lp* lp, const double* x, const double* y, const double* kx,
                      const double* kty, double* out);
/* Kkt.convergence(...).converged (Kkt.scala:130-143). */
int32_t orc_kkt_converged(const double* err4, double rhs_norm, double obj_norm, double eps_abs,
                          double eps_rel);
/* Certificates.primalInfeasibility / dualInfeasibility (Certificates.scala:60-195).
 * Return 1 and set *value when the direction is a candidate, else 0. */
int32_t orc_primal_infeasibility(const orc_lp* lp, const double* y, const double* kty,
                                 double* value);
int32_t orc_dual_infeasibility(const orc_lp* lp, const double* d, const double* kd,
                               double* value);
/* Certificates.classify (Certificates.scala:207-219); -1 when no certificate. */
int32_t orc_classify(const orc_lp* lp, const double* x, const double* y, double tol);

/* Pdhg.solveWith (Pdhg.scala:48-55) with ScalaKernels (precision 64) or
 * Float32Kernels (precision 32). warm_* may be NULL; a warm step size or
 * primal weight <= 0 or non-finite means "not given". x_out (n), y_out (m)
 * and rc_out (n) may be NULL. Returns 0, or -1 on a rejected argument. */
int32_t orc_pdhg_solve(const orc_lp* lp, const orc_params* params, int32_t precision,
                       const double* warm_x, const double* warm_y, double warm_step,
                       double warm_weight, orc_result* res, double* x_out, double* y_out,
                       double* rc_out);

/* MixedPrecision.solve (MixedPrecision.scala:104-199) with a Float32Kernels
 * device (device_precision 32) or an fp64 device (64, which skips the
 * reduced pass). */
int32_t orc_mixed_solve(const orc_lp* lp, const orc_params* params, int32_t device_precision,
                        const orc_mixed_params* mixed, orc_result* res, orc_mixed_info* info,
                        double* x_out, double* y_out, double* rc_out);

#ifdef __cplusplus
}
#endif
#endif
