// Structures shared by the host orchestration (host_api.cpp) and the
// persistent PDHG kernel (pdhg_kernel.hip). Plain data only.
#pragma once
#include <stdint.h>

namespace prima {

// Per-LP scalar state of the restarted-PDHG iteration (Pdhg.scala:194-202),
// resident in device memory across kernel launches.
struct SolveState {
  double step;            // eta, the adaptive step size
  double weight;          // omega, the primal weight
  double wsum;            // sum of accepted step sizes in the restart period
  double err_at_restart;  // weighted KKT error at the last restart
  double last_candidate;  // weighted KKT error at the last checkpoint
  double kkt[4];          // primal res, dual res, primal obj, dual obj (terminal)
  int32_t iteration;
  int32_t restarts;
  int32_t last_restart_it;
  int32_t status;         // -2: not initialised, -1: running, >= 0: SolveStatus
  int32_t trial_counter;  // trials taken in total (parity of partial slots)
  int32_t trials;         // trials of the current (unaccepted) iteration
  int32_t evals;          // checkpoints evaluated
  int32_t used_average;   // terminal iterate was the running average
  int64_t total_trials;
};

// Launch-uniform parameters (PdhgParams.scala:9-59).
struct KParams {
  double eps_abs, eps_rel, infeas_tol;
  double r_suff, r_nec, r_art, theta;
  int32_t eval_freq;     // iteration caps are per LP (LpView::max_iter)
  int32_t mode;          // 0: run, 1: finalize with final_status
  int32_t final_status;
  int32_t certificates;  // 1: test Farkas certificates at checkpoints
  int32_t test_stall;    // test hook: make every group's barrier wait give up
};

// Device view of one LP (one group of workgroups). Pointers are device
// pointers; T-typed arrays are stored as void* and reinterpreted.
struct LpView {
  int32_t n, m, n_eq;
  int32_t G;         // workgroups in this group
  int32_t max_iter;  // this LP's iteration cap
  int32_t iter_budget;  // end this LP's chunk after a checkpoint once this many ran
  const int32_t *a_ptr, *a_col;
  const void* a_val;
  const int32_t *t_ptr, *t_col;
  const void* t_val;
  const int32_t *a_rb, *t_rb;    // {first row, first nnz} per row block (blocks + 1 pairs)
  const int32_t *a_wrb, *t_wrb;  // first row block of each rank (G + 1)
  const void *c, *q, *l, *u;     // scaled data, iteration precision
  const double *oc, *oq, *ol, *ou, *rs, *cs;  // original data and scales
  // Optional lower bound on the inequality-row multipliers (default 0): the
  // shifted dual cone of an iterative-refinement correction problem. ylo is
  // scaled and in the iteration precision, oylo in original units.
  const void* ylo;
  const double* oylo;
  void *x, *y, *kx, *kty, *xn, *yn, *xbar, *kxbar, *sx, *sy, *skx, *skty, *xr, *yr;
  double *x_out, *y_out, *rc_out;
  double* part;       // 2 * G * 4 iteration partials, then G * kEvalParts
  uint32_t* bar;      // barrier words: count, generation, abort, then a flag line per rank
  SolveState* st;
  uint32_t* launch_abort;  // set (LP index + 1) when a group barrier gives up
  double offset, rhs_norm, obj_norm, spec_orig;
};

constexpr int kIterParts = 4;
constexpr int kEvalParts = 24;

}  // namespace prima
