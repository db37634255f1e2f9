// C-ABI of libprima_hip (include/prima_hip.h) and the host side of the
// device solve: validation, equilibration, transpose, SpMV row blocking,
// workgroup partitioning, device memory layout and the chunked persistent
// launches. Reference citations are relative to
// modules/prima-core/src/main/scala/org/noaidi/prima/.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/prima_hip.h"
#include "pdhg_shared.h"

namespace prima {
// pdhg_kernel.hip
int occupancy_blocks_per_cu(int precision);
hipError_t launch_pdhg(int precision, const LpView* views, const int32_t* wg_off,
                       const int32_t* wg_ent, int grid,
                       const KParams& P, hipStream_t stream);
int threads_per_block();
hipError_t scratch_init(int device);
void* scratch_of(int device, size_t* bytes);
int tile_nnz();
int block_row_cap();
// kernels_ops.hip
hipError_t op_spmv(int prec, int rows, int64_t nnz, const int32_t* ptr, const int32_t* col,
                   const void* val, const void* x, void* out, hipStream_t s);
hipError_t op_axpby(int prec, int64_t n, double a, const void* x, double b, const void* y,
                    void* out, hipStream_t s);
hipError_t op_scale(int prec, int64_t n, double a, const void* x, void* out, hipStream_t s);
hipError_t op_primal(int prec, int64_t n, const void* x, const void* kty, const void* c,
                     const void* l, const void* u, double tau, void* out, hipStream_t s);
hipError_t op_dual(int prec, int64_t m, const void* y, const void* kxb, const void* q,
                   double sigma, int64_t n_eq, void* out, hipStream_t s);
hipError_t op_dot(int prec, int64_t n, const void* x, const void* y, double* part,
                  hipStream_t s);
hipError_t op_spmv_vector(int prec, int tpr, int rows, const int32_t* ptr, const int32_t* col,
                          const void* val, const void* x, void* out, hipStream_t s);
hipError_t op_spmv_blocks(int prec, int nblocks, int bpw, const int32_t* ptr, const int32_t* col,
                          const void* val, const int32_t* rbd, const void* x, void* out,
                          hipStream_t s);
hipError_t op_spmv_panels(int prec, int P, int nranges, const int32_t* range_row0,
                          const int32_t* tile_off, const uint16_t* starts, const int32_t* pcol,
                          const void* pval, int nlong, const int32_t* long_rows,
                          const int32_t* ptr, const int32_t* col, const void* val, const void* x,
                          void* out, hipStream_t s);
int panel_rows();
int panel_tile();
int small_block_tile();
int small_block_rows();
hipError_t op_spmv_blocks_small(int prec, int nblocks, const int32_t* ptr, const int32_t* col,
                                const void* val, const int32_t* rbd, const void* x, void* out,
                                hipStream_t s);
// refine_kernels.hip
int ref_partial_blocks();
int ref_result_doubles();
hipError_t ref_evaluate(int n, int m, int n_eq, int64_t nnz, const int32_t* aptr,
                        const int32_t* acol, const double* aval, const int32_t* tptr,
                        const int32_t* tcol, const double* tval, const double* x,
                        const double* y, const double* c, const double* q, const double* l,
                        const double* u, double eps_abs, double eps_rel, double rhs_norm,
                        double obj_norm, double offset, double* rq, double* rc, double* part,
                        double* out, hipStream_t s);
hipError_t ref_build(int n, int m, int n_eq, double dp, double dd, double cap, double fix,
                     const double* rc, const double* rq, const double* x, const double* y,
                     const double* l, const double* u, double* cc, double* cl, double* cu,
                     double* cq, double* ylo, hipStream_t s);
hipError_t ref_apply(int n, int m, int n_eq, double dp, double dd, const double* cx,
                     const double* cy, const double* l, const double* u, double* x, double* y,
                     hipStream_t s);
}  // namespace prima

using namespace prima;

// Everything derived from the constraint matrix alone: equilibration,
// transpose, norm bounds and SpMV row blocks. LPs that share K (branch-and-
// bound nodes, scenario sweeps over one network) share one of these.
namespace prima {
struct MatrixPrep {
  uint64_t key = 0;
  std::vector<double> aval;  // scaled values, A order
  std::vector<int32_t> tptr, tcol;
  std::vector<double> tval;  // scaled transpose
  std::vector<double> rs, cs;
  double spec_scaled = 0, spec_orig = 0;
  std::vector<int32_t> a_rb, t_rb;    // block start rows (blocks + 1)
  std::vector<int32_t> a_rbd, t_rbd;  // device descriptors
  // choices of prima_tune_solver for this matrix (copied from the context's
  // table when the matrix is prepared): group size of an LP solved alone
  // (0: the nonzeros rule) and row blocks per workgroup of a group (0: the
  // coarse cut, k: about k blocks per workgroup; -1: the default, 1)
  mutable int tuned_G = 0;
  mutable int tuned_bpw = -1;
};
}  // namespace prima

struct prima_ctx {
  int device = 0;
  int precision = 64;
  bool closed = false;
  hipStream_t stream = nullptr;
  std::string err;
  int cu_count = 0;
  int xcds = 1;  // accelerator complex dies, each with its own L2
  int occ[2] = {0, 0};  // blocks per CU for float / double kernels
  double* scratch = nullptr;  // reduction partials for dot / squaredNorm
  std::atomic<int> cancel{0};
  int test_stall = 0;  // test hook: every group's barrier wait gives up
  bool scratch_ok = false;  // this device's allocation-free workspace is known
  // progress of the running solve, readable from another thread
  std::atomic<int64_t> prog_launches{0}, prog_iterations{0}, prog_finished{0};
  int64_t prog_base = 0;  // accepted iterations of finished waves
  std::string device_name;
  bool coop_ok = false;
  prima_stats stats{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // recently prepared matrices, by content (LPs that share K reuse them)
  std::vector<std::shared_ptr<const prima::MatrixPrep>> prep_cache;
  // solver settings chosen by prima_tune_solver, by matrix content key:
  // {key, group size, row blocks per workgroup}
  struct Tuned {
    uint64_t key;
    int G, bpw;
  };
  std::vector<Tuned> tuned;
};

struct prima_vec {
  void* d = nullptr;
  int64_t n = 0;
};

struct prima_mat {
  int32_t rows = 0, cols = 0;
  int64_t nnz = 0;
  int32_t* ptr = nullptr;
  int32_t* col = nullptr;
  void* val = nullptr;
  int32_t* rbd = nullptr;  // row-block descriptors for the row-block schedule
  int32_t nblocks = 0;
  int32_t* rbd_small = nullptr;  // the same for the small-block schedule
  int32_t nblocks_small = 0;
  int32_t schedule = 0;    // PRIMA_SPMV_* (0: by mean row length)
  // column-panel layout (PRIMA_SPMV_PANELS); panels == 0 when not built
  int32_t panels = 0, nranges = 0, nlong = 0;
  int32_t* range_row0 = nullptr;  // nranges + 1 row offsets
  int32_t* tile_off = nullptr;    // nranges * panels + 1 entry offsets
  uint16_t* starts = nullptr;     // per tile: row starts relative to the tile
  int32_t* pcol = nullptr;
  void* pval = nullptr;
  int32_t* long_rows = nullptr;   // rows summed outside the tiles
};

namespace {

using clk = std::chrono::steady_clock;
double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

int fail(prima_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int dev_fail(prima_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, PRIMA_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return dev_fail(ctx, _e, #expr); \
  } while (0)

size_t esize(int prec) { return prec == 32 ? sizeof(float) : sizeof(double); }

// Narrow to the iteration precision; infinities survive
// (Float32Kernels.scala:43-52).
void to_prec(int prec, const double* src, int64_t n, void* dst) {
  if (prec == 32) {
    float* f = (float*)dst;
    for (int64_t i = 0; i < n; ++i) f[i] = (float)src[i];
  } else {
    std::memcpy(dst, src, (size_t)n * sizeof(double));
  }
}

int check_open(prima_ctx* ctx) {
  if (!ctx) return PRIMA_EINVAL;
  if (ctx->closed) return fail(ctx, PRIMA_ECLOSED, "context is closed");
  ctx->err.clear();
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, PRIMA_EDEVICE, "hipSetDevice");
  return PRIMA_OK;
}

// ---------------------------------------------------------------------------
// Host preprocessing of one LP.

struct HostLp {
  int32_t n = 0, m = 0, n_eq = 0;
  int64_t nnz = 0;
  const prima_lp* src = nullptr;
  std::shared_ptr<const MatrixPrep> M;
  std::vector<double> sc, sq, sl, su;
  double rhs_norm = 0, obj_norm = 0;
  int G = 1;      // workgroups owning this LP in the current launch
  int G_cap = 1;  // most workgroups it may ever get (sizes its arrays)
};

// LpProblem.apply requirements (LpProblem.scala:62-104) and canonical CSR
// (SparseMatrix.scala:15-21, 204-260).
std::string validate_lp(const prima_lp* lp) {
  if (lp->n < 0 || lp->m < 0) return "negative dimensions";
  if (lp->n_eq < 0 || lp->n_eq > lp->m) return "numEqualities out of range";
  if (lp->m > 0 && !lp->row_ptr) return "row_ptr is null";
  if (lp->nnz > 0 && (!lp->col || !lp->val)) return "col or val is null";
  if (lp->n > 0 && (!lp->c || !lp->l || !lp->u)) return "c, l or u is null";
  if (lp->m > 0 && !lp->q) return "q is null";
  if (lp->nnz > INT32_MAX) return "nnz exceeds int32 indices";
  if (lp->m > 0) {
    if (lp->row_ptr[0] != 0 || lp->row_ptr[lp->m] != lp->nnz) return "row_ptr does not span nnz";
    for (int32_t r = 0; r < lp->m; ++r) {
      if (lp->row_ptr[r + 1] < lp->row_ptr[r]) return "row_ptr is not monotone";
      for (int32_t p = lp->row_ptr[r]; p < lp->row_ptr[r + 1]; ++p) {
        if (lp->col[p] < 0 || lp->col[p] >= lp->n) return "column index out of range";
        if (p > lp->row_ptr[r] && lp->col[p] <= lp->col[p - 1])
          return "column indices not strictly ascending within a row";
        if (!std::isfinite(lp->val[p])) return "matrix entry is not finite";
      }
    }
  } else if (lp->nnz != 0) {
    return "nnz without rows";
  }
  for (int32_t i = 0; i < lp->n; ++i) {
    if (!(lp->l[i] <= lp->u[i])) return "empty bound interval for variable " + std::to_string(i);
    if (!std::isfinite(lp->c[i])) return "objective coefficient is not finite";
  }
  for (int32_t r = 0; r < lp->m; ++r)
    if (!std::isfinite(lp->q[r])) return "right-hand side is not finite";
  if (!std::isfinite(lp->offset)) return "objective offset is not finite";
  return "";
}

// PdhgParams / ScalingParams requirements (PdhgParams.scala:61-86,
// Scaling.scala:16-20).
std::string validate_params(const prima_params* p) {
  if (!(p->eps_abs >= 0.0)) return "epsAbs must be non-negative";
  if (!(p->eps_rel >= 0.0)) return "epsRel must be non-negative";
  if (p->max_iterations <= 0) return "maxIterations must be positive";
  if (p->evaluation_frequency <= 0) return "evaluationFrequency must be positive";
  if (!(p->primal_weight_smoothing >= 0.0 && p->primal_weight_smoothing <= 1.0))
    return "primalWeightSmoothing must lie in [0, 1]";
  if (!(p->restart_sufficient > 0.0 && p->restart_sufficient <= p->restart_necessary))
    return "restartSufficientReduction must be positive and no greater than "
           "restartNecessaryReduction";
  if (!(p->restart_artificial > 0.0 && p->restart_artificial <= 1.0))
    return "restartArtificialFraction must lie in (0, 1]";
  if (!(p->infeasibility_tolerance >= 0.0)) return "infeasibilityTolerance must be non-negative";
  if (p->ruiz_iterations < 0) return "ruizIterations must be non-negative";
  if (!(p->pock_chambolle_alpha >= 0.0 && p->pock_chambolle_alpha <= 2.0))
    return "pockChambolleAlpha must lie in [0, 2]";
  if (p->workgroups_per_lp < 0 || p->chunk_iterations < 0) return "negative launch tuning";
  return "";
}

void transpose_csr(int32_t rows, int32_t cols, const int32_t* ptr, const int32_t* col,
                   const double* val, std::vector<int32_t>& tptr, std::vector<int32_t>& tcol,
                   std::vector<double>& tval) {
  const int64_t nnz = rows > 0 ? ptr[rows] : 0;
  tptr.assign((size_t)cols + 1, 0);
  tcol.resize((size_t)nnz);
  tval.resize((size_t)nnz);
  for (int64_t k = 0; k < nnz; ++k) tptr[col[k] + 1]++;
  for (int32_t c = 0; c < cols; ++c) tptr[c + 1] += tptr[c];
  std::vector<int32_t> cur(tptr.begin(), tptr.end() - 1);
  for (int32_t r = 0; r < rows; ++r)
    for (int32_t p = ptr[r]; p < ptr[r + 1]; ++p) {
      const int32_t d = cur[col[p]]++;
      tcol[d] = r;
      tval[d] = val[p];
    }
}

double norm_inf(int32_t rows, const int32_t* ptr, const double* val) {
  double mx = 0.0;
  for (int32_t r = 0; r < rows; ++r) {
    double s = 0.0;
    for (int32_t p = ptr[r]; p < ptr[r + 1]; ++p) s += std::fabs(val[p]);
    mx = std::max(mx, s);
  }
  return mx;
}

double norm2(const double* a, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * a[i];
  return std::sqrt(s);
}

// Scaling.apply (Scaling.scala:77-155), the matrix part: Ruiz passes then
// Pock-Chambolle, giving the scaled values and the row/column factors.
void equilibrate_matrix(const prima_lp* lp, const prima_params* p, MatrixPrep& M) {
  const int32_t m = lp->m, n = lp->n;
  const int64_t nnz = lp->nnz;
  M.aval.assign(lp->val, lp->val + nnz);
  M.rs.assign((size_t)m, 1.0);
  M.cs.assign((size_t)n, 1.0);
  std::vector<double> rf((size_t)m), cf((size_t)n);
  auto apply = [&]() {
    for (int32_t r = 0; r < m; ++r)
      for (int32_t q = lp->row_ptr[r]; q < lp->row_ptr[r + 1]; ++q)
        M.aval[q] = M.aval[q] * rf[r] * cf[lp->col[q]];
    for (int32_t r = 0; r < m; ++r) M.rs[r] *= rf[r];
    for (int32_t c = 0; c < n; ++c) M.cs[c] *= cf[c];
  };
  for (int it = 0; it < p->ruiz_iterations; ++it) {
    std::fill(cf.begin(), cf.end(), 0.0);
    for (int32_t r = 0; r < m; ++r) {
      double mx = 0.0;
      for (int32_t q = lp->row_ptr[r]; q < lp->row_ptr[r + 1]; ++q) {
        const double a = std::fabs(M.aval[q]);
        mx = std::max(mx, a);
        double& cm = cf[lp->col[q]];
        if (a > cm) cm = a;
      }
      rf[r] = mx > 0.0 ? 1.0 / std::sqrt(mx) : 1.0;
    }
    for (int32_t c = 0; c < n; ++c) cf[c] = cf[c] > 0.0 ? 1.0 / std::sqrt(cf[c]) : 1.0;
    apply();
  }
  if (p->apply_pock_chambolle) {
    const double alpha = p->pock_chambolle_alpha;
    std::fill(cf.begin(), cf.end(), 0.0);
    for (int32_t r = 0; r < m; ++r) {
      double s = 0.0;
      for (int32_t q = lp->row_ptr[r]; q < lp->row_ptr[r + 1]; ++q) {
        const double a = std::fabs(M.aval[q]);
        s += alpha == 1.0 ? a : std::pow(a, 2.0 - alpha);
        cf[lp->col[q]] += alpha == 1.0 ? a : std::pow(a, alpha);
      }
      rf[r] = s > 0.0 ? 1.0 / std::sqrt(s) : 1.0;
    }
    for (int32_t c = 0; c < n; ++c) cf[c] = cf[c] > 0.0 ? 1.0 / std::sqrt(cf[c]) : 1.0;
    apply();
  }
}

// Scaling.apply, the vector part: c~ = Dc c, q~ = Dr q, l~ = l / Dc, u~ = u / Dc.
void scale_vectors(const prima_lp* lp, const MatrixPrep& M, HostLp& h) {
  const int32_t m = lp->m, n = lp->n;
  h.sc.resize(n);
  h.sl.resize(n);
  h.su.resize(n);
  h.sq.resize(m);
  for (int32_t i = 0; i < n; ++i) {
    const double d = M.cs[i];
    h.sc[i] = lp->c[i] * d;
    h.sl[i] = lp->l[i] / d;
    h.su[i] = lp->u[i] / d;
  }
  for (int32_t r = 0; r < m; ++r) h.sq[r] = lp->q[r] * M.rs[r];
}

// Row blocks for the streamed SpMV: consecutive rows while their nnz fit the
// LDS tile (and at most a tile's worth of rows, so empty rows cannot make a
// block unbounded); a row longer than the tile is a block of its own.
std::vector<int32_t> row_blocks(int32_t rows, const int32_t* ptr, int tile = 0, int nt = 0) {
  if (tile <= 0) tile = tile_nnz();
  if (nt <= 0) nt = block_row_cap();
  std::vector<int32_t> rb;
  rb.push_back(0);
  int32_t r = 0;
  while (r < rows) {
    const int32_t start = r;
    if (ptr[r + 1] - ptr[r] > tile) {
      ++r;
    } else {
      int64_t acc = 0;
      while (r < rows && r - start < nt && acc + (ptr[r + 1] - ptr[r]) <= tile) {
        acc += ptr[r + 1] - ptr[r];
        ++r;
      }
    }
    rb.push_back(r);
  }
  return rb;
}

// Split row blocks into G contiguous ranges of near-equal cost (nnz plus a
// per-row charge for the fused elementwise work).
std::vector<int32_t> partition(const std::vector<int32_t>& rb, const int32_t* ptr, int G) {
  const size_t nb = rb.size() - 1;
  std::vector<double> pre(nb + 1, 0.0);
  for (size_t b = 0; b < nb; ++b) {
    const double nz = (double)(ptr[rb[b + 1]] - ptr[rb[b]]);
    pre[b + 1] = pre[b] + nz + 4.0 * (rb[b + 1] - rb[b]) + 16.0;
  }
  std::vector<int32_t> w((size_t)G + 1);
  w[0] = 0;
  for (int g = 1; g < G; ++g) {
    const double target = pre[nb] * g / G;
    w[g] = (int32_t)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
    if (w[g] < w[g - 1]) w[g] = w[g - 1];
    if (w[g] > (int32_t)nb) w[g] = (int32_t)nb;
  }
  w[G] = (int32_t)nb;
  return w;
}

// Row blocks for a group of G workgroups. A phase lasts as long as the
// slowest workgroup's chain of block passes, and one pass costs about the same
// few memory round trips whatever the block's size, so a group splits each
// matrix into about G blocks (never beyond the coarse caps, never under 64
// rows or 512 nonzeros); a single workgroup uses the shared coarse cut.
void group_blocks(const MatrixPrep& M, const int32_t* aptr, int32_t n, int32_t m, int64_t nnz,
                  int G, std::vector<int32_t>& a_rb, std::vector<int32_t>& t_rb) {
  const int bpw = M.tuned_bpw < 0 ? 1 : M.tuned_bpw;
  if (G <= 1 || bpw == 0) {
    a_rb = M.a_rb;
    t_rb = M.t_rb;
    return;
  }
  const int64_t parts = (int64_t)G * bpw;
  auto cap = [&](int64_t total, int64_t lo, int64_t hi) {
    return (int)std::max<int64_t>(lo, std::min<int64_t>(hi, (total + parts - 1) / parts));
  };
  const int tile = cap(nnz, 512, tile_nnz());
  a_rb = row_blocks(m, aptr, tile, cap(m, 64, block_row_cap()));
  t_rb = row_blocks(n, M.tptr.data(), tile, cap(n, 64, block_row_cap()));
}

// {first row, first nnz} pairs, one per block plus a sentinel.
std::vector<int32_t> block_descriptors(const std::vector<int32_t>& rb, const int32_t* ptr) {
  std::vector<int32_t> d(rb.size() * 2);
  for (size_t b = 0; b < rb.size(); ++b) {
    d[2 * b] = rb[b];
    d[2 * b + 1] = ptr[rb[b]];
  }
  return d;
}

// Content hash of a matrix and the scaling parameters (FNV-1a over the CSR
// bytes): the cache below reuses a preparation only for identical data.
uint64_t matrix_key(const prima_lp* lp, const prima_params* p) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* data, size_t bytes) {
    const unsigned char* b = (const unsigned char*)data;
    for (size_t i = 0; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
  };
  const int32_t dims[3] = {lp->n, lp->m, p->ruiz_iterations};
  mix(dims, sizeof(dims));
  mix(&p->apply_pock_chambolle, sizeof(p->apply_pock_chambolle));
  mix(&p->pock_chambolle_alpha, sizeof(p->pock_chambolle_alpha));
  if (lp->m > 0) mix(lp->row_ptr, ((size_t)lp->m + 1) * 4);
  mix(lp->col, (size_t)lp->nnz * 4);
  mix(lp->val, (size_t)lp->nnz * 8);
  return h;
}

std::shared_ptr<const MatrixPrep> prepare_matrix(const prima_lp* lp, const prima_params* p) {
  auto M = std::make_shared<MatrixPrep>();
  static const int32_t zero_ptr[1] = {0};
  const int32_t* ptr = lp->m > 0 ? lp->row_ptr : zero_ptr;
  equilibrate_matrix(lp, p, *M);
  transpose_csr(lp->m, lp->n, ptr, lp->col, M->aval.data(), M->tptr, M->tcol, M->tval);
  // spectralNormBound = sqrt(||K||_1 ||K||_inf) (SparseMatrix.scala:181)
  M->spec_scaled = std::sqrt(norm_inf(lp->n, M->tptr.data(), M->tval.data()) *
                             norm_inf(lp->m, ptr, M->aval.data()));
  {
    std::vector<int32_t> optr, ocol;
    std::vector<double> oval;
    transpose_csr(lp->m, lp->n, ptr, lp->col, lp->val, optr, ocol, oval);
    M->spec_orig = std::sqrt(norm_inf(lp->n, optr.data(), oval.data()) *
                             norm_inf(lp->m, ptr, lp->val));
  }
  M->a_rb = row_blocks(lp->m, ptr);
  M->t_rb = row_blocks(lp->n, M->tptr.data());
  M->a_rbd = block_descriptors(M->a_rb, ptr);
  M->t_rbd = block_descriptors(M->t_rb, M->tptr.data());
  return M;
}

// Certificates.classify (Certificates.scala:207-219) in double precision on
// the original data, used to re-check a certificate found by a float32 pass
// (MixedPrecision.scala:143-154). Returns -1 when neither holds.
int host_classify(const prima_lp* lp, const double* x, const double* y, double tol) {
  const int32_t n = lp->n, m = lp->m;
  static const int32_t zero_ptr[1] = {0};
  const int32_t* ptr = m > 0 ? lp->row_ptr : zero_ptr;
  std::vector<int32_t> tptr, tcol;
  std::vector<double> tval;
  transpose_csr(m, n, ptr, lp->col, lp->val, tptr, tcol, tval);
  const double spec =
      std::sqrt(norm_inf(n, tptr.data(), tval.data()) * norm_inf(m, ptr, lp->val));
  std::vector<double> aty(n, 0.0), ax(m, 0.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t q = tptr[i]; q < tptr[i + 1]; ++q) aty[i] += tval[q] * y[tcol[q]];
  for (int32_t r = 0; r < m; ++r)
    for (int32_t q = ptr[r]; q < ptr[r + 1]; ++q) ax[r] += lp->val[q] * x[lp->col[q]];
  auto terms = [](double rc, double lo, double hi, double& sq, double& lin) {
    if (rc > 0.0) {
      if (std::isinf(lo) && lo < 0) sq += rc * rc;
      else lin += lo * rc;
    } else if (rc < 0.0) {
      if (std::isinf(hi) && hi > 0) sq += rc * rc;
      else lin += hi * rc;
    }
  };
  {  // primal infeasibility from y
    double nsq = 0.0;
    for (int32_t r = 0; r < m; ++r) nsq += y[r] * y[r];
    if (nsq > 0.0 && std::isfinite(nsq)) {
      const double norm = std::sqrt(nsq);
      double value = 0.0, rsq = 0.0, csq = 0.0;
      for (int32_t r = 0; r < m; ++r) value += lp->q[r] * y[r];
      for (int32_t i = 0; i < n; ++i) terms(-aty[i], lp->l[i], lp->u[i], rsq, value);
      for (int32_t r = lp->n_eq; r < m; ++r)
        if (y[r] < 0.0) csq += y[r] * y[r];
      if (value > 0.0 && std::isfinite(value)) {
        const double rt = spec > 0.0 ? std::sqrt(rsq) / (spec * norm) : 0.0;
        const double h = std::hypot(rt, std::sqrt(csq) / norm);
        if (std::isfinite(h) && h <= tol) return PRIMA_PRIMAL_INFEASIBLE;
      }
    }
  }
  {  // unboundedness from x
    double nsq = 0.0, value = 0.0;
    for (int32_t i = 0; i < n; ++i) {
      nsq += x[i] * x[i];
      value += lp->c[i] * x[i];
    }
    if (nsq > 0.0 && std::isfinite(nsq) && value < 0.0 && std::isfinite(value)) {
      const double norm = std::sqrt(nsq);
      double rsq = 0.0, bsq = 0.0;
      for (int32_t r = 0; r < m; ++r)
        if (r < lp->n_eq || ax[r] < 0.0) rsq += ax[r] * ax[r];
      for (int32_t i = 0; i < n; ++i) {
        if (x[i] > 0.0 && !(std::isinf(lp->u[i]) && lp->u[i] > 0)) bsq += x[i] * x[i];
        else if (x[i] < 0.0 && !(std::isinf(lp->l[i]) && lp->l[i] < 0)) bsq += x[i] * x[i];
      }
      const double rt = spec > 0.0 ? std::sqrt(rsq) / (spec * norm) : 0.0;
      const double h = std::hypot(rt, std::sqrt(bsq) / norm);
      if (std::isfinite(h) && h <= tol) return PRIMA_DUAL_INFEASIBLE;
    }
  }
  return -1;
}

// ---------------------------------------------------------------------------
// Device memory for a wave of LPs: one arena, carved into 256-byte aligned
// arrays.

struct Carver {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  }
};

struct LpLayout {
  size_t a_ptr, a_col, t_ptr, t_col, a_rb, t_rb, a_wrb, t_wrb;
  size_t a_val, t_val, c, q, l, u;
  size_t x, y, kx, kty, xn, yn, xbar, kxbar, sx, sy, skx, skty, xr, yr;
  size_t oc, oq, ol, ou, rs, cs, x_out, y_out, rc_out, part, bar, st;
  size_t ylo, oylo;
};

// Caller-held device memory a wave is carved from instead of allocating.
struct Workspace {
  char* base = nullptr;
  size_t bytes = 0;
};

struct Wave {
  std::vector<int> idx;  // indices into the batch
  std::vector<LpLayout> lay;
  std::vector<LpView> views;
  // launch plan: workgroup w runs the chunks ent[off[w] .. off[w+1])
  std::vector<int32_t> off, ent;
  char* arena = nullptr;
  LpView* d_views = nullptr;
  int32_t* d_off = nullptr;
  int32_t* d_ent = nullptr;
  bool owned = true;  // false when carved from a Workspace
  int grid = 0;
};

// Sub-array offsets for one LP. With `owner`, the matrix arrays (CSR of A
// and of its transpose, block descriptors, scale factors) are those of an
// earlier LP of the wave with the same prepared matrix.
LpLayout carve(Carver& cv, const HostLp& h, int prec, bool with_ylo, const LpLayout* owner) {
  const size_t s = esize(prec), n = (size_t)h.n, m = (size_t)h.m, z = (size_t)h.nnz;
  const size_t G = (size_t)h.G_cap;
  LpLayout L;
  if (owner) {
    L = *owner;
  } else {
    L.a_ptr = cv.take((m + 1) * 4);
    L.a_col = cv.take(z * 4);
    L.t_ptr = cv.take((n + 1) * 4);
    L.t_col = cv.take(z * 4);
    L.a_val = cv.take(z * s);
    L.t_val = cv.take(z * s);
    L.rs = cv.take(m * 8);
    L.cs = cv.take(n * 8);
  }
  L.a_rb = cv.take((m + 2) * 8);  // block descriptors: at most one block per row
  L.t_rb = cv.take((n + 2) * 8);
  L.a_wrb = cv.take((G + 1) * 4);
  L.t_wrb = cv.take((G + 1) * 4);
  L.c = cv.take(n * s);
  L.l = cv.take(n * s);
  L.u = cv.take(n * s);
  L.q = cv.take(m * s);
  L.x = cv.take(n * s);
  L.kty = cv.take(n * s);
  L.xn = cv.take(n * s);
  L.xbar = cv.take(n * s);
  L.sx = cv.take(n * s);
  L.skty = cv.take(n * s);
  L.xr = cv.take(n * s);
  L.y = cv.take(m * s);
  L.kx = cv.take(m * s);
  L.yn = cv.take(m * s);
  L.kxbar = cv.take(m * s);
  L.sy = cv.take(m * s);
  L.skx = cv.take(m * s);
  L.yr = cv.take(m * s);
  L.oc = cv.take(n * 8);
  L.ol = cv.take(n * 8);
  L.ou = cv.take(n * 8);
  L.x_out = cv.take(n * 8);
  L.rc_out = cv.take(n * 8);
  L.oq = cv.take(m * 8);
  L.y_out = cv.take(m * 8);
  L.part = cv.take((2 * G * kIterParts + G * kEvalParts) * 8);
  L.bar = cv.take(128 + 128 * G);  // count, generation, abort; one flag line per rank
  L.st = 0;  // placed by the wave: all states of a wave are contiguous
  L.ylo = with_ylo ? cv.take(m * s) : 0;
  L.oylo = with_ylo ? cv.take(m * 8) : 0;
  return L;
}

struct SolveJob {
  const prima_lp* lp = nullptr;
  HostLp h;
  const prima_warm* warm = nullptr;
  const double* ylo = nullptr;  // shifted dual cone for rows >= n_eq (original units)
  prima_result* res = nullptr;
  double *x_out = nullptr, *y_out = nullptr, *rc_out = nullptr;
  double prep_ms = 0;
  // iteration cap of this LP (0: the params' max_iterations); with a cold
  // budget, a pass that hits the cap restarts cold in place with that budget
  int32_t max_iter = 0;
  int32_t cold_budget = 0;
  int32_t capped_iterations = -1;  // iterations of the capped pass, once restarted
  // a warm pass and its cold twin (indices into the same job list): the twin
  // waits dormant and races the warm pass once that runs long (see run_wave)
  int shadow = -1, shadow_of = -1;
};

// Compulsory bytes of one accepted iteration (each array counted once, no
// gather amplification): 2Z(s+4) + 4(m+n+2) + s(13n+11m).
double iter_bytes(const HostLp& h, int prec) {
  const double Z = (double)h.nnz, n = h.n, m = h.m, sb = (double)esize(prec);
  return 2 * Z * (sb + 4) + 4 * (m + n + 2) + sb * (13 * n + 11 * m);
}

// Cost model of one accepted iteration on G workgroups, in microseconds: a
// workgroup's share of HBM with the device full (about 16 GB/s: 4 TB/s over
// 256 workgroups, as measured on batches of dispatch LPs), two group barriers
// of about 4.5 us when G > 1, and about 8 us of dependent memory round trips
// per iteration at any size (one dispatch LP alone on 64 workgroups: 20.8 us
// measured, 21.4 modelled; 256 at one workgroup each: 415 us, 423 modelled).
// Deterministic on purpose: launch lengths, and with them the iterations at
// which LPs are regrouped, do not depend on timing.
double model_iter_us(const HostLp& h, int prec, int G) {
  constexpr double kWorkgroupBytesPerUs = 16000.0;
  return iter_bytes(h, prec) / (G * kWorkgroupBytesPerUs) + (G > 1 ? 9.0 : 0.0) + 8.0;
}

// Most workgroups one LP's group may have: ranks are packed into 9 bits of a
// launch-table entry.
constexpr int kMaxGroup = 512;

// Workgroups per LP for a set of LPs solved together. Alone an LP is latency
// bound and spreads at about 1k nonzeros per workgroup (G_cap, which sizes
// its per-rank arrays and bounds any later regrouping). Together, each LP
// gets a share of the device in proportion to its bytes per iteration, so
// that every LP iterates at about the same rate: an LP whose share rounds to
// fewer than two workgroups runs alone on one, packed with others onto the
// workgroups the groups leave free (enough for those LPs' bytes). From four
// LPs on a group is capped at 32 workgroups unless it moves more than 64 MB
// per iteration (at 65k nonzeros four LPs ran 30.8 us per iteration at 32
// workgroups each, 32.8 at 64). Twins (shadow_of >= 0) do not count towards
// the shares until they run.
void choose_groups(const prima_params* p, const std::vector<SolveJob*>& js, int capacity,
                   int prec) {
  const int gmax = std::max(1, std::min(capacity, kMaxGroup));
  for (SolveJob* J : js) {
    const int64_t nnz = std::max<int64_t>(J->h.nnz, 1);
    const int64_t rule = J->h.M && J->h.M->tuned_G > 0 ? J->h.M->tuned_G : (nnz + 1023) / 1024;
    const int alone = (int)std::max<int64_t>(1, std::min<int64_t>(rule, gmax));
    J->h.G_cap = std::max(J->h.G_cap, alone);
    if (p->workgroups_per_lp > 0) {
      J->h.G = std::min(p->workgroups_per_lp, gmax);
      J->h.G_cap = std::max(J->h.G_cap, J->h.G);
    }
  }
  if (p->workgroups_per_lp > 0) return;
  double total = 0;
  int64_t B = 0;
  for (SolveJob* J : js)
    if (J->shadow_of < 0) {
      total += iter_bytes(J->h, prec);
      ++B;
    }
  const double per_wg = total > 0 ? total / capacity : 1.0;
  for (SolveJob* J : js) {
    const double bytes = iter_bytes(J->h, prec);
    const double r = bytes / per_wg;
    int64_t g = r >= 1.5 ? (int64_t)std::llround(r) : 1;
    if (B >= 4) g = std::min<int64_t>(g, std::max<int64_t>(32, (int64_t)(bytes / 2.0e6)));
    J->h.G = (int)std::max<int64_t>(1, std::min<int64_t>(g, J->h.G_cap));
  }
  // leave the single-workgroup LPs enough workgroups for their bytes (at
  // least one): shrink the group with the lightest load per workgroup first
  for (;;) {
    int64_t grouped = 0;
    double single_bytes = 0;
    bool any_single = false;
    for (SolveJob* J : js) {
      if (J->shadow_of >= 0) continue;
      if (J->h.G > 1) grouped += J->h.G;
      else {
        single_bytes += iter_bytes(J->h, prec);
        any_single = true;
      }
    }
    const int64_t need =
        any_single ? std::max<int64_t>(1, (int64_t)std::ceil(single_bytes / per_wg - 1e-9)) : 0;
    if (grouped + need <= capacity) break;
    SolveJob* lightest = nullptr;
    double lw = INFINITY;
    for (SolveJob* J : js) {
      if (J->shadow_of >= 0 || J->h.G <= 1) continue;
      const double w = iter_bytes(J->h, prec) / J->h.G;
      if (w < lw) {
        lw = w;
        lightest = J;
      }
    }
    if (!lightest) break;
    lightest->h.G -= 1;
  }
}

// Launch table entries: LP index << 9 | rank in its group (G <= kMaxGroup).
inline int32_t wg_entry(size_t k, int rank) { return (int32_t)(((int64_t)k << 9) | rank); }

// Workgroup placement. The dispatcher hands a launch's workgroups to the
// XCDs round-robin (on MI355X's 8: workgroup w to XCD w mod 8), and each XCD
// has its own L2. The table is filled LP by LP; this reorders it so that
// consecutive entries take the slots of one XCD before the next (XCD 0 gets
// workgroups 0, 8, 16, ...), so a group sits on as few XCDs as its size and
// the launch's width allow (on a 256-workgroup launch, a group of up to 32 on
// one XCD). Placement changes no arithmetic (results are bit-identical); on the
// 65k-nnz dispatch LP, f64 microseconds per iteration went from 92.8 to 79.6
// for 32 LPs at G = 8, 138 to 126 for 64 at G = 4, 41.6 to 40.1 for 8 at
// G = 32 and 21.9 to 20.8 for one LP at G = 64. Returns the workgroup index
// of each of n consecutive slots.
std::vector<int32_t> xcd_order(size_t n, int xcds) {
  std::vector<int32_t> out;
  out.reserve(n);
  const size_t x_n = (size_t)std::max(1, xcds);
  for (size_t x = 0; x < x_n; ++x)
    for (size_t w = x; w < n; w += x_n) out.push_back((int32_t)w);
  return out;
}

// Upload one wave, run it to completion in chunks, download the results.
int run_wave(prima_ctx* ctx, int prec, std::vector<SolveJob>& jobs, const std::vector<int>& idx,
             const prima_params* p, bool certificates, int capacity,
             const Workspace* ws = nullptr) {
  if (idx.size() >= ((size_t)1 << 22))
    return fail(ctx, PRIMA_EINVAL, "more than 2^22 LPs in one wave");
  Wave W;
  W.idx = idx;
  Carver cv;
  std::vector<int> owner_of(idx.size(), -1);
  for (size_t k = 0; k < idx.size(); ++k) {
    const HostLp& h = jobs[idx[k]].h;
    for (size_t e = 0; e < k; ++e)
      if (jobs[idx[e]].h.M == h.M && owner_of[e] < 0) {
        owner_of[k] = (int)e;
        break;
      }
    W.lay.push_back(carve(cv, h, prec, jobs[idx[k]].ylo != nullptr,
                          owner_of[k] >= 0 ? &W.lay[owner_of[k]] : nullptr));
  }
  // one contiguous block of per-LP states, read back with a single copy at
  // every launch boundary
  const size_t st_base = cv.take(idx.size() * sizeof(SolveState) + 8);
  for (size_t k = 0; k < idx.size(); ++k) W.lay[k].st = st_base + k * sizeof(SolveState);
  // the launch-wide abort word follows the states (read back with them)
  const size_t abort_off = st_base + idx.size() * sizeof(SolveState);
  const size_t arena_bytes = cv.off;
  auto t_up = clk::now();
  const size_t views_bytes = (idx.size() * sizeof(LpView) + 255) & ~(size_t)255;
  const size_t off_bytes = (((size_t)capacity + 1) * sizeof(int32_t) + 255) & ~(size_t)255;
  const size_t ent_bytes = ((size_t)capacity + idx.size()) * sizeof(int32_t);
  void* arena = nullptr;
  if (ws) {
    // carve the wave from the caller's workspace: nothing is allocated
    const size_t need = arena_bytes + views_bytes + off_bytes + ent_bytes;
    if (need > ws->bytes)
      return fail(ctx, PRIMA_EINVAL,
                  "the LP needs " + std::to_string(need) + " bytes of device workspace; the "
                  "allocation-free solve holds at most " + std::to_string(ws->bytes));
    W.owned = false;
    arena = ws->base;
    W.d_views = (LpView*)(ws->base + arena_bytes);
    W.d_off = (int32_t*)(ws->base + arena_bytes + views_bytes);
    W.d_ent = (int32_t*)(ws->base + arena_bytes + views_bytes + off_bytes);
  } else {
    hipError_t e = hipMalloc(&arena, arena_bytes);
    if (e != hipSuccess) return fail(ctx, PRIMA_ENOMEM, "device arena of " +
                                               std::to_string(arena_bytes) + " bytes");
  }
  W.arena = (char*)arena;
  struct Guard {
    Wave& w;
    ~Guard() {
      if (!w.owned) return;
      if (w.arena) (void)hipFree(w.arena);
      if (w.d_views) (void)hipFree(w.d_views);
      if (w.d_off) (void)hipFree(w.d_off);
      if (w.d_ent) (void)hipFree(w.d_ent);
    }
  } guard{W};
  HIP_TRY(ctx, hipMemsetAsync(arena, 0, arena_bytes, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));

  // Uploads are synchronous copies: host buffers may be released as soon as
  // each call returns.
  const size_t s = esize(prec);
  auto put = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
    if (bytes == 0) return hipSuccess;
    return hipMemcpy(W.arena + off, src, bytes, hipMemcpyHostToDevice);
  };
  std::vector<char> conv;
  auto put_t = [&](size_t off, const double* src, size_t count) -> hipError_t {
    if (count == 0) return hipSuccess;
    if (prec == 64) return put(off, src, count * 8);
    conv.resize(count * s);
    to_prec(prec, src, (int64_t)count, conv.data());
    return put(off, conv.data(), count * s);
  };

  // Compulsory bytes of a pass (each array counted once, no gather
  // amplification): accepted iteration 2Z(s+4) + 4(m+n+2) + s(13n+11m);
  // rejected trial Z(s+4) + 4(m+1) + s(n+m) + s(7n+5m); checkpoint
  // Z(s+4) + 4(m+1) + s(n+m) + 8(4n+4m). The matrix terms (values, columns,
  // row pointers of A and of its transpose) are charged once per pass over a
  // matrix that LPs of the wave share: for the LP of the sharing set with the
  // most passes (charge_matrices, at the end of the wave).
  std::vector<double> a_passes(idx.size(), 0.0), t_passes(idx.size(), 0.0);
  auto account = [&](size_t k, const SolveState& st) {
    const SolveJob& J = jobs[idx[k]];
    const double n = J.h.n, m = J.h.m, sb = (double)s;
    const double it = st.iteration, rej = (double)(st.total_trials - st.iteration);
    const double v_iter = sb * (13 * n + 11 * m);
    const double v_rej = sb * (n + m) + sb * (7 * n + 5 * m);
    const double v_eval = sb * (n + m) + 8 * (4 * n + 4 * m);
    const double bytes = it * v_iter + rej * v_rej + st.evals * v_eval;
    ctx->stats.algorithmic_bytes += bytes;
    ctx->stats.algorithmic_bytes_by_precision[prec == 64] += bytes;
    ctx->stats.accepted_iterations += st.iteration;
    ctx->stats.trials += st.total_trials;
    ctx->stats.evaluations += st.evals;
    a_passes[k] += it + rej + st.evals;
    t_passes[k] += it;
  };
  auto charge_matrices = [&]() {
    for (size_t k = 0; k < idx.size(); ++k) {
      if (owner_of[k] >= 0) continue;  // charged with its owner
      double a = a_passes[k], t = t_passes[k];
      for (size_t e = k + 1; e < idx.size(); ++e)
        if (owner_of[e] == (int)k) {
          a = std::max(a, a_passes[e]);
          t = std::max(t, t_passes[e]);
        }
      const HostLp& h = jobs[idx[k]].h;
      const double Z = (double)h.nnz, sb = (double)s;
      const double bytes = a * (Z * (sb + 4) + 4.0 * (h.m + 1)) + t * (Z * (sb + 4) + 4.0 * (h.n + 1));
      ctx->stats.algorithmic_bytes += bytes;
      ctx->stats.algorithmic_bytes_by_precision[prec == 64] += bytes;
    }
  };

  // Starting point (Pdhg.scala:265-303): x0 = proj(warm / colScale or 0),
  // y0 = warm / rowScale with inequality multipliers clamped at 0; step from
  // the spectral bound, weight from data norms (Pdhg.scala:250-263), warm
  // values overriding when finite and positive.
  auto start_point = [&](size_t k, const prima_warm* w) -> hipError_t {
    const SolveJob& J = jobs[idx[k]];
    const HostLp& h = J.h;
    const LpLayout& L = W.lay[k];
    std::vector<double> x0((size_t)h.n, 0.0), y0((size_t)h.m, 0.0);
    if (w && w->primal)
      for (int32_t i = 0; i < h.n; ++i) x0[i] = w->primal[i] / h.M->cs[i];
    for (int32_t i = 0; i < h.n; ++i) {
      const double v = x0[i];
      x0[i] = v < h.sl[i] ? h.sl[i] : (v > h.su[i] ? h.su[i] : v);
    }
    if (w && w->dual)
      for (int32_t r = 0; r < h.m; ++r) {
        const double v = w->dual[r] / h.M->rs[r];
        y0[r] = (r >= h.n_eq && v < 0.0) ? 0.0 : v;
      }
    hipError_t e = put_t(L.x, x0.data(), h.n);
    if (e == hipSuccess) e = put_t(L.y, y0.data(), h.m);
    SolveState st;
    std::memset(&st, 0, sizeof(st));
    const bool ws = w && std::isfinite(w->step_size) && w->step_size > 0.0;
    const bool ww = w && std::isfinite(w->primal_weight) && w->primal_weight > 0.0;
    st.step = ws ? w->step_size : (h.M->spec_scaled > 0.0 ? 1.0 / h.M->spec_scaled : 1.0);
    const double cn = norm2(h.sc.data(), h.n), qn = norm2(h.sq.data(), h.m);
    st.weight = ww ? w->primal_weight : ((cn > 0.0 && qn > 0.0) ? cn / qn : 1.0);
    st.err_at_restart = INFINITY;
    st.last_candidate = INFINITY;
    st.status = -2;
    if (e == hipSuccess) e = put(L.st, &st, sizeof(st));
    return e;
  };

  // Every LP of the wave is admitted at once: an LP on more than one
  // workgroup holds them for a launch, and LPs on one workgroup are packed,
  // several to a workgroup if need be (see plan below).
  const size_t nlp = idx.size();
  std::vector<char> active(nlp, 1);
  // cold twins of warm passes wait dormant until they race (see below);
  // dropped slots ran but do not report
  std::vector<char> dormant(nlp, 0), dropped(nlp, 0);
  std::vector<int> slot_of(jobs.size(), -1);
  for (size_t k = 0; k < nlp; ++k) {
    slot_of[idx[k]] = (int)k;
    dormant[k] = jobs[idx[k]].shadow_of >= 0;
  }
  for (size_t k = 0; k < idx.size(); ++k) {
    SolveJob& J = jobs[idx[k]];
    HostLp& h = J.h;
    const prima_lp* lp = J.lp;
    const LpLayout& L = W.lay[k];
    static const int32_t zero_ptr[1] = {0};
    const int32_t* aptr = lp->m > 0 ? lp->row_ptr : zero_ptr;
    {
      std::vector<int32_t> arb, trb;
      group_blocks(*h.M, aptr, h.n, h.m, h.nnz, h.G, arb, trb);
      const std::vector<int32_t> ad = block_descriptors(arb, aptr);
      const std::vector<int32_t> td = block_descriptors(trb, h.M->tptr.data());
      HIP_TRY(ctx, put(L.a_rb, ad.data(), ad.size() * 4));
      HIP_TRY(ctx, put(L.t_rb, td.data(), td.size() * 4));
      std::vector<int32_t> awrb = partition(arb, aptr, h.G);
      std::vector<int32_t> twrb = partition(trb, h.M->tptr.data(), h.G);
      HIP_TRY(ctx, put(L.a_wrb, awrb.data(), awrb.size() * 4));
      HIP_TRY(ctx, put(L.t_wrb, twrb.data(), twrb.size() * 4));
    }
    if (owner_of[k] < 0) {  // the first LP of the wave with this matrix uploads it
      HIP_TRY(ctx, put(L.a_ptr, aptr, ((size_t)h.m + 1) * 4));
      HIP_TRY(ctx, put(L.a_col, lp->col, (size_t)h.nnz * 4));
      HIP_TRY(ctx, put(L.t_ptr, h.M->tptr.data(), ((size_t)h.n + 1) * 4));
      HIP_TRY(ctx, put(L.t_col, h.M->tcol.data(), (size_t)h.nnz * 4));
      HIP_TRY(ctx, put_t(L.a_val, h.M->aval.data(), (size_t)h.nnz));
      HIP_TRY(ctx, put_t(L.t_val, h.M->tval.data(), (size_t)h.nnz));
      HIP_TRY(ctx, put(L.cs, h.M->cs.data(), (size_t)h.n * 8));
      HIP_TRY(ctx, put(L.rs, h.M->rs.data(), (size_t)h.m * 8));
    }
    HIP_TRY(ctx, put_t(L.c, h.sc.data(), h.n));
    HIP_TRY(ctx, put_t(L.l, h.sl.data(), h.n));
    HIP_TRY(ctx, put_t(L.u, h.su.data(), h.n));
    HIP_TRY(ctx, put_t(L.q, h.sq.data(), h.m));
    HIP_TRY(ctx, put(L.oc, lp->c, (size_t)h.n * 8));
    HIP_TRY(ctx, put(L.ol, lp->l, (size_t)h.n * 8));
    HIP_TRY(ctx, put(L.ou, lp->u, (size_t)h.n * 8));
    HIP_TRY(ctx, put(L.oq, lp->q, (size_t)h.m * 8));
    if (J.ylo) {
      // y >= ylo in original units is y~ >= ylo / rowScale in scaled units
      std::vector<double> ys((size_t)h.m, 0.0);
      for (int32_t r = h.n_eq; r < h.m; ++r) ys[r] = J.ylo[r] / h.M->rs[r];
      HIP_TRY(ctx, put_t(L.ylo, ys.data(), h.m));
      std::vector<double> yo((size_t)h.m, 0.0);
      for (int32_t r = h.n_eq; r < h.m; ++r) yo[r] = J.ylo[r];
      HIP_TRY(ctx, put(L.oylo, yo.data(), (size_t)h.m * 8));
    }

    HIP_TRY(ctx, start_point(k, J.warm));

    LpView v;
    std::memset(&v, 0, sizeof(v));
    char* a = W.arena;
    v.n = h.n;
    v.m = h.m;
    v.n_eq = h.n_eq;
    v.G = h.G;
    v.max_iter = J.max_iter > 0 ? J.max_iter : p->max_iterations;
    v.a_ptr = (const int32_t*)(a + L.a_ptr);
    v.a_col = (const int32_t*)(a + L.a_col);
    v.a_val = a + L.a_val;
    v.t_ptr = (const int32_t*)(a + L.t_ptr);
    v.t_col = (const int32_t*)(a + L.t_col);
    v.t_val = a + L.t_val;
    v.a_rb = (const int32_t*)(a + L.a_rb);
    v.t_rb = (const int32_t*)(a + L.t_rb);
    v.a_wrb = (const int32_t*)(a + L.a_wrb);
    v.t_wrb = (const int32_t*)(a + L.t_wrb);
    v.c = a + L.c;
    v.q = a + L.q;
    v.l = a + L.l;
    v.u = a + L.u;
    v.oc = (const double*)(a + L.oc);
    v.oq = (const double*)(a + L.oq);
    v.ol = (const double*)(a + L.ol);
    v.ou = (const double*)(a + L.ou);
    v.rs = (const double*)(a + L.rs);
    v.cs = (const double*)(a + L.cs);
    v.x = a + L.x;
    v.y = a + L.y;
    v.kx = a + L.kx;
    v.kty = a + L.kty;
    v.xn = a + L.xn;
    v.yn = a + L.yn;
    v.xbar = a + L.xbar;
    v.kxbar = a + L.kxbar;
    v.sx = a + L.sx;
    v.sy = a + L.sy;
    v.skx = a + L.skx;
    v.skty = a + L.skty;
    v.xr = a + L.xr;
    v.yr = a + L.yr;
    v.x_out = (double*)(a + L.x_out);
    v.y_out = (double*)(a + L.y_out);
    v.rc_out = (double*)(a + L.rc_out);
    v.part = (double*)(a + L.part);
    v.bar = (uint32_t*)(a + L.bar);
    v.st = (SolveState*)(a + L.st);
    v.launch_abort = (uint32_t*)(a + abort_off);
    v.ylo = J.ylo ? (const void*)(a + L.ylo) : nullptr;
    v.oylo = J.ylo ? (const double*)(a + L.oylo) : nullptr;
    v.offset = lp->offset;
    v.rhs_norm = h.rhs_norm;
    v.obj_norm = h.obj_norm;
    v.spec_orig = h.M->spec_orig;
    W.views.push_back(v);
  }
  if (W.owned) {
    HIP_TRY(ctx, hipMalloc((void**)&W.d_views, W.views.size() * sizeof(LpView)));
    // a plan has at most `capacity` workgroups; each LP appears G times
    HIP_TRY(ctx, hipMalloc((void**)&W.d_off, ((size_t)capacity + 1) * sizeof(int32_t)));
    HIP_TRY(ctx, hipMalloc((void**)&W.d_ent, ((size_t)capacity + nlp) * sizeof(int32_t)));
  }
  const double upload_ms = ms_since(t_up);

  // chunked launches until every LP of the wave reached a status
  KParams P;
  std::memset(&P, 0, sizeof(P));
  P.eps_abs = p->eps_abs;
  P.eps_rel = p->eps_rel;
  P.infeas_tol = p->infeasibility_tolerance;
  P.r_suff = p->restart_sufficient;
  P.r_nec = p->restart_necessary;
  P.r_art = p->restart_artificial;
  P.theta = p->primal_weight_smoothing;
  P.eval_freq = p->evaluation_frequency;
  P.mode = 0;
  P.certificates = certificates ? 1 : 0;
  P.test_stall = ctx->test_stall;
  std::vector<SolveState> states(nlp);
  // A group barrier that gave up (the kernel's bounded wait) is a device
  // failure, not a solver outcome: the whole call fails with PRIMA_EDEVICE.
  auto check_abort = [&]() -> int {
    uint32_t word = 0;
    HIP_TRY(ctx, hipMemcpy(&word, W.arena + abort_off, sizeof(word), hipMemcpyDeviceToHost));
    if (word == 0) return PRIMA_OK;
    const size_t k = (size_t)word - 1;
    const int G = k < nlp ? W.views[k].G : 0;
    return fail(ctx, PRIMA_EDEVICE,
                "a group barrier of the solver kernel gave up (LP " +
                    std::to_string(k < nlp ? idx[k] : -1) + ", " + std::to_string(G) +
                    " workgroups): a workgroup of the group did not arrive in time");
  };
  for (auto& st : states) st.status = -1;
  // Launch plan. Each running LP on G > 1 workgroups holds G workgroups
  // for the launch; the LPs on one workgroup are packed, longest modelled
  // iteration first onto the least loaded, onto the workgroups the groups
  // leave free. Each LP's chunk (iterations before it stops at a checkpoint)
  // follows the cost model so that every workgroup's list takes about the
  // same time, about 50 ms, and at least one evaluation period. Without a
  // time limit the plan depends on the model alone, so launch boundaries --
  // and with them the points where LPs are regrouped -- fall on the same
  // iterations every run and batched results are reproducible. Under a time
  // limit the target shrinks with the limit and follows the measured launch
  // times. chunk_iterations > 0 fixes every chunk instead.
  const bool timed_chunks = p->chunk_iterations == 0 && p->time_limit_ms >= 0;
  const int64_t freq = std::max(1, p->evaluation_frequency);
  double model_scale = 1.0;    // measured / modelled launch time (time limits only)
  double planned_us = 0.0;     // modelled length of the launch just planned
  auto running_lp = [&](size_t k) {
    return active[k] && !dormant[k] && !dropped[k] && states[k].status < 0;
  };
  auto plan = [&](double target_us) -> int {
    std::vector<size_t> groups, singles;
    for (size_t k = 0; k < nlp; ++k) {
      if (!running_lp(k)) continue;
      (jobs[idx[k]].h.G > 1 ? groups : singles).push_back(k);
    }
    std::vector<std::vector<int32_t>> bins;
    std::vector<double> bin_us;
    for (size_t k : groups) {
      const HostLp& h = jobs[idx[k]].h;
      const double t = model_iter_us(h, prec, h.G);
      for (int g = 0; g < h.G; ++g) {
        bins.push_back({wg_entry(k, g)});
        bin_us.push_back(t);
      }
    }
    const size_t first_single = bins.size();
    if (bins.size() + (singles.empty() ? 0 : 1) > (size_t)capacity)
      return fail(ctx, PRIMA_EDEVICE, "launch plan exceeds the co-resident workgroups (" +
                                          std::to_string(bins.size()) + " grouped of " +
                                          std::to_string(capacity) + ")");
    if (!singles.empty()) {
      const size_t room = bins.size() < (size_t)capacity ? (size_t)capacity - bins.size() : 1;
      const size_t nb = std::min(singles.size(), std::max<size_t>(1, room));
      std::vector<double> t1(nlp, 0.0);
      for (size_t k : singles) t1[k] = model_iter_us(jobs[idx[k]].h, prec, 1);
      std::stable_sort(singles.begin(), singles.end(),
                       [&](size_t a, size_t b) { return t1[a] > t1[b]; });
      bins.resize(first_single + nb);
      bin_us.resize(first_single + nb, 0.0);
      for (size_t k : singles) {
        size_t best = first_single;
        for (size_t b = first_single + 1; b < bins.size(); ++b)
          if (bin_us[b] < bin_us[best]) best = b;
        bins[best].push_back(wg_entry(k, 0));
        bin_us[best] += t1[k];
      }
    }
    // one launch lasts as long as its longest list; every list gets at least
    // one evaluation period of iterations
    double T = target_us;
    for (double t : bin_us) T = std::max(T, (double)freq * t);
    planned_us = 0.0;
    for (size_t b = 0; b < bins.size(); ++b) {
      int64_t budget = p->chunk_iterations > 0
                           ? (int64_t)p->chunk_iterations
                           : std::max<int64_t>(freq, (int64_t)(T / bin_us[b]) / freq * freq);
      budget = std::min<int64_t>(budget, 100000000);
      planned_us = std::max(planned_us, budget * bin_us[b]);
      for (int32_t e : bins[b]) W.views[(size_t)(e >> 9)].iter_budget = (int32_t)budget;
    }
    // consecutive slots of one XCD first (see xcd_order)
    const std::vector<int32_t> where = xcd_order(bins.size(), ctx->xcds);
    std::vector<int32_t> slot_of_wg(bins.size());
    for (size_t b = 0; b < bins.size(); ++b) slot_of_wg[(size_t)where[b]] = (int32_t)b;
    W.off.assign(1, 0);
    W.ent.clear();
    for (size_t w = 0; w < bins.size(); ++w) {
      for (int32_t e : bins[(size_t)slot_of_wg[w]]) W.ent.push_back(e);
      W.off.push_back((int32_t)W.ent.size());
    }
    W.grid = (int)bins.size();
    HIP_TRY(ctx, hipMemcpy(W.d_views, W.views.data(), W.views.size() * sizeof(LpView),
                           hipMemcpyHostToDevice));
    HIP_TRY(ctx, hipMemcpy(W.d_off, W.off.data(), W.off.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice));
    if (!W.ent.empty())
      HIP_TRY(ctx, hipMemcpy(W.d_ent, W.ent.data(), W.ent.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice));
    return PRIMA_OK;
  };
  std::vector<double> done_ms(nlp, -1.0), admit_ms(nlp, 0.0);
  auto t_solve = clk::now();
  // Elastic batching: when the set of running LPs changes (LPs finish, a
  // twin starts), group sizes are chosen again for the survivors, and an LP
  // whose size moved by half or more is re-partitioned (row blocks, rank
  // ranges). Safe at a launch boundary, where every LP's state is committed.
  std::vector<char> was_running(nlp, 0);
  auto regroup = [&]() -> int {
    std::vector<size_t> which;
    bool changed = false;
    for (size_t k = 0; k < nlp; ++k) {
      const bool r = running_lp(k);
      changed |= r != (bool)was_running[k];
      was_running[k] = r;
      if (r) which.push_back(k);
    }
    if (!changed || which.empty()) return PRIMA_OK;
    std::vector<SolveJob> tmp(which.size());
    std::vector<SolveJob*> ptrs(which.size());
    for (size_t j = 0; j < which.size(); ++j) {
      const HostLp& h = jobs[idx[which[j]]].h;
      tmp[j].h.n = h.n;
      tmp[j].h.m = h.m;
      tmp[j].h.nnz = h.nnz;
      tmp[j].h.G_cap = h.G_cap;
      ptrs[j] = &tmp[j];
    }
    choose_groups(p, ptrs, capacity, prec);
    // keep a group size that is within a factor of 1.5 of the new choice,
    // unless the kept sizes would not leave the packed LPs a workgroup
    std::vector<int> next(which.size());
    int64_t grouped = 0;
    bool any_single = false;
    for (size_t j = 0; j < which.size(); ++j) {
      const int want = std::min(tmp[j].h.G, jobs[idx[which[j]]].h.G_cap);
      const int cur = jobs[idx[which[j]]].h.G;
      next[j] = (2 * want < 3 * cur && 2 * cur < 3 * want) ? cur : want;
      if (next[j] > 1) grouped += next[j];
      else any_single = true;
    }
    if (grouped + (any_single ? 1 : 0) > capacity)
      for (size_t j = 0; j < which.size(); ++j)
        next[j] = std::min(tmp[j].h.G, jobs[idx[which[j]]].h.G_cap);
    for (size_t j = 0; j < which.size(); ++j) {
      const size_t k = which[j];
      SolveJob& J = jobs[idx[k]];
      const int want = next[j];
      if (want == J.h.G) continue;
      J.h.G = want;
      static const int32_t zero_ptr[1] = {0};
      const int32_t* aptr = J.lp->m > 0 ? J.lp->row_ptr : zero_ptr;
      std::vector<int32_t> arb, trb;
      group_blocks(*J.h.M, aptr, J.h.n, J.h.m, J.h.nnz, J.h.G, arb, trb);
      const std::vector<int32_t> ad = block_descriptors(arb, aptr);
      const std::vector<int32_t> td = block_descriptors(trb, J.h.M->tptr.data());
      HIP_TRY(ctx, put(W.lay[k].a_rb, ad.data(), ad.size() * 4));
      HIP_TRY(ctx, put(W.lay[k].t_rb, td.data(), td.size() * 4));
      std::vector<int32_t> awrb = partition(arb, aptr, J.h.G);
      std::vector<int32_t> twrb = partition(trb, J.h.M->tptr.data(), J.h.G);
      HIP_TRY(ctx, put(W.lay[k].a_wrb, awrb.data(), awrb.size() * 4));
      HIP_TRY(ctx, put(W.lay[k].t_wrb, twrb.data(), twrb.size() * 4));
      W.views[k].G = J.h.G;
    }
    return PRIMA_OK;
  };
  auto target_us = [&]() {
    if (!timed_chunks) return 50000.0;
    // launches of roughly 50 ms, shorter under a short time limit
    return std::min(50000.0, std::max(1000.0, p->time_limit_ms * 1000.0 / 8.0)) / model_scale;
  };
  for (size_t k = 0; k < nlp; ++k) was_running[k] = running_lp(k);
  {
    const int rc = plan(target_us());
    if (rc) return rc;
  }
  for (;;) {
    auto t_l = clk::now();
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIP_TRY(ctx, launch_pdhg(prec, W.d_views, W.d_off, W.d_ent, W.grid, P, ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    const double launch_ms = ms_since(t_l);
    {
      float ev_ms = 0.f;
      HIP_TRY(ctx, hipEventElapsedTime(&ev_ms, ctx->ev0, ctx->ev1));
      ctx->stats.kernel_ms += ev_ms;
      ctx->stats.launches += 1;
      ctx->stats.kernel_ms_by_precision[prec == 64] += ev_ms;
      ctx->stats.launches_by_precision[prec == 64] += 1;
    }
    {
      const int rc = check_abort();
      if (rc) return rc;
    }
    {
      int64_t wave_iters = 0, newly = 0;
      HIP_TRY(ctx, hipMemcpy(states.data(), W.arena + st_base, nlp * sizeof(SolveState),
                             hipMemcpyDeviceToHost));
      // Warm passes with a cold twin: once the warm pass has run a quarter of
      // its cap while fewer than capacity/32 LPs run (workgroups to spare),
      // the twin starts; the first of the two to reach a conclusive status at
      // a launch boundary stands for the LP and the other stops where it is.
      // A warm pass that reaches its cap leaves the LP to the twin (started
      // then if it was not yet). Boundaries fall on the same iterations every
      // run, so the outcome is reproducible.
      {
        int running = 0;
        for (size_t k = 0; k < nlp; ++k)
          running += active[k] && !dormant[k] && states[k].status < 0;
        auto conclusive_status = [](int st) {
          return st == PRIMA_OPTIMAL || st == PRIMA_PRIMAL_INFEASIBLE ||
                 st == PRIMA_DUAL_INFEASIBLE;
        };
        auto stop = [&](size_t k) -> hipError_t {
          dropped[k] = 1;
          states[k].status = PRIMA_INTERRUPTED;
          return put(W.lay[k].st + offsetof(SolveState, status), &states[k].status,
                     sizeof(int32_t));
        };
        for (size_t k = 0; k < nlp; ++k) {
          SolveJob& J = jobs[idx[k]];
          if (J.shadow < 0 || dropped[k]) continue;
          const size_t t = (size_t)slot_of[J.shadow];
          if (states[k].status >= 0) {
            if (states[k].status == PRIMA_ITERATION_LIMIT) {
              J.capped_iterations = states[k].iteration;
              dropped[k] = 1;
              dormant[t] = 0;  // the twin stands in (waiting, if it never ran)
            } else if (!dropped[t]) {
              // the warm pass concluded: the twin is not needed (a finished
              // twin keeps its state, a running or dormant one is stopped)
              dormant[t] = 0;
              if (states[t].status >= 0) dropped[t] = 1;
              else HIP_TRY(ctx, stop(t));
            }
            continue;
          }
          if (dormant[t] && states[k].iteration >= J.max_iter / 4 &&
              running < capacity / 32) {
            dormant[t] = 0;
            ++running;
          }
          if (!dormant[t] && !dropped[t] && states[t].status >= 0 &&
              conclusive_status(states[t].status)) {
            J.capped_iterations = states[k].iteration;
            HIP_TRY(ctx, stop(k));
          }
        }
      }
      for (size_t k = 0; k < nlp; ++k) {
        // a capped pass with a cold budget restarts cold in place, keeping
        // its workgroups, instead of after the rest of the batch
        SolveJob& J = jobs[idx[k]];
        if (states[k].status != PRIMA_ITERATION_LIMIT || J.cold_budget <= 0 ||
            J.capped_iterations >= 0 || J.shadow >= 0)
          continue;
        account(k, states[k]);
        ctx->prog_base += states[k].iteration;
        J.capped_iterations = states[k].iteration;
        HIP_TRY(ctx, start_point(k, nullptr));
        W.views[k].max_iter = J.cold_budget;
        HIP_TRY(ctx, hipMemcpy(W.d_views + k, &W.views[k], sizeof(LpView),
                               hipMemcpyHostToDevice));
        states[k].status = -2;
        states[k].iteration = 0;
      }
      for (size_t k = 0; k < nlp; ++k) {
        if (states[k].status >= 0 && done_ms[k] < 0) {
          done_ms[k] = ms_since(t_solve);
          ++newly;
        }
        wave_iters += states[k].iteration;
      }
      ctx->prog_launches.fetch_add(1);
      ctx->prog_finished.fetch_add(newly);
      ctx->prog_iterations.store(ctx->prog_base + wave_iters);
    }
    bool all_done = true;
    for (auto& st : states) all_done &= st.status >= 0;
    if (all_done) break;
    int final_status = -1;
    if (ctx->cancel.load()) final_status = PRIMA_INTERRUPTED;
    else if (p->time_limit_ms >= 0 && ms_since(t_solve) >= (double)p->time_limit_ms)
      final_status = PRIMA_TIME_LIMIT;
    if (final_status >= 0) {
      // finalise the running LPs (each reports its current point)
      KParams F = P;
      F.mode = 1;
      F.final_status = final_status;
      {
        const int rc = plan(target_us());
        if (rc) return rc;
      }
      for (;;) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, launch_pdhg(prec, W.d_views, W.d_off, W.d_ent, W.grid, F, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        {
          float ev_ms = 0.f;
          HIP_TRY(ctx, hipEventElapsedTime(&ev_ms, ctx->ev0, ctx->ev1));
          ctx->stats.kernel_ms += ev_ms;
          ctx->stats.launches += 1;
          ctx->stats.kernel_ms_by_precision[prec == 64] += ev_ms;
          ctx->stats.launches_by_precision[prec == 64] += 1;
        }
        {
          const int rc = check_abort();
          if (rc) return rc;
        }
        HIP_TRY(ctx, hipMemcpy(states.data(), W.arena + st_base, nlp * sizeof(SolveState),
                               hipMemcpyDeviceToHost));
        break;
      }
      break;
    }
    if (timed_chunks && planned_us > 0.0) {
      const double r = launch_ms * 1000.0 / planned_us;
      model_scale = std::min(100.0, std::max(0.01, 0.5 * model_scale + 0.5 * r));
    }
    {
      int rc = regroup();
      if (rc) return rc;
      rc = plan(target_us());
      if (rc) return rc;
    }
  }
  const double solve_ms = ms_since(t_solve);
  {
    int64_t wave_iters = 0, newly = 0;
    for (size_t k = 0; k < nlp; ++k) {
      wave_iters += states[k].iteration;
      newly += done_ms[k] < 0 ? 1 : 0;
    }
    ctx->prog_base += wave_iters;
    ctx->prog_iterations.store(ctx->prog_base);
    ctx->prog_finished.fetch_add(newly);
  }
  for (size_t k = 0; k < nlp; ++k) {
    SolveJob& J = jobs[idx[k]];
    const LpLayout& L = W.lay[k];
    const SolveState& st = states[k];
    // a slot reports unless it was dropped, never had to run, or is a twin
    // whose warm pass still stands (an interrupted race)
    const bool twin_unneeded = J.shadow_of >= 0 && !dropped[(size_t)slot_of[J.shadow_of]];
    if (dropped[k] || dormant[k] || twin_unneeded) {
      if (!dormant[k]) account(k, st);
      continue;
    }
    if (st.status < 0) return fail(ctx, PRIMA_EDEVICE, "solve ended without a status");
    if (J.x_out && J.h.n)
      HIP_TRY(ctx, hipMemcpy(J.x_out, W.arena + L.x_out, (size_t)J.h.n * 8,
                             hipMemcpyDeviceToHost));
    if (J.rc_out && J.h.n)
      HIP_TRY(ctx, hipMemcpy(J.rc_out, W.arena + L.rc_out, (size_t)J.h.n * 8,
                             hipMemcpyDeviceToHost));
    if (J.y_out && J.h.m)
      HIP_TRY(ctx, hipMemcpy(J.y_out, W.arena + L.y_out, (size_t)J.h.m * 8,
                             hipMemcpyDeviceToHost));
    account(k, st);
    prima_result* r = J.res;
    r->status = st.status;
    r->iterations = st.iteration;
    r->restarts = st.restarts;
    r->used_average = st.used_average;
    r->primal_residual = st.kkt[0];
    r->dual_residual = st.kkt[1];
    r->objective = st.kkt[2];
    r->dual_objective = st.kkt[3];
    // a twin reports the LP's time from the warm pass's admission
    const size_t a = J.shadow_of >= 0 ? (size_t)slot_of[J.shadow_of] : k;
    r->solve_time_ms = (done_ms[k] >= 0 ? done_ms[k] : solve_ms) - admit_ms[a];
    r->total_time_ms = r->solve_time_ms + upload_ms + J.prep_ms;
    r->final_step_size = st.step;
    r->final_primal_weight = st.weight;
    r->trials = st.total_trials;
    r->evaluations = st.evals;
    r->workgroups = J.h.G;
  }
  charge_matrices();
  return PRIMA_OK;
}

// Solve a set of prepared jobs. Every job whose arrays fit in a quarter of
// free device memory runs in one wave, all admitted at once (LPs on one
// workgroup share workgroups); a batch too large for that runs in waves of
// consecutive jobs that do fit, and then a warm pass cannot race a cold twin
// in another wave: it restarts cold in place instead, and the twins are left
// out.
int run_jobs(prima_ctx* ctx, int prec, std::vector<SolveJob>& jobs, const prima_params* p,
             bool certificates = true, const Workspace* ws = nullptr) {
  const int occ = ctx->occ[prec == 32 ? 0 : 1];
  const int capacity = std::max(1, ctx->cu_count * occ);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
  const size_t budget = free_b / 4;
  // nothing to iterate on: the first evaluation of Pdhg.solve finds every
  // residual zero and both objectives at the offset (Pdhg.scala:209-212)
  std::vector<int> work;
  for (size_t j = 0; j < jobs.size(); ++j) {
    if (jobs[j].h.n == 0 && jobs[j].h.m == 0) {
      if (jobs[j].shadow_of >= 0) continue;
      prima_result& r = *jobs[j].res;
      std::memset(&r, 0, sizeof(r));
      r.status = PRIMA_OPTIMAL;
      r.objective = r.dual_objective = jobs[j].lp->offset;
      r.final_step_size = NAN;
      r.final_primal_weight = NAN;
      continue;
    }
    work.push_back((int)j);
  }
  // arena bytes of each job as the wave carves it (a matrix shared with an
  // earlier job of the wave is counted once)
  auto arena_of = [&](const std::vector<int>& w) {
    Carver cv;
    std::vector<LpLayout> lay;
    std::vector<int> owner(w.size(), -1);
    for (size_t a = 0; a < w.size(); ++a) {
      for (size_t e = 0; e < a; ++e)
        if (jobs[w[e]].h.M == jobs[w[a]].h.M && owner[e] < 0) {
          owner[a] = (int)e;
          break;
        }
      lay.push_back(carve(cv, jobs[w[a]].h, prec, jobs[w[a]].ylo != nullptr,
                          owner[a] >= 0 ? &lay[owner[a]] : nullptr));
    }
    return cv.off + w.size() * (sizeof(SolveState) + 8);
  };
  auto groups_for = [&](const std::vector<int>& w) {
    std::vector<SolveJob*> ptrs;
    for (int j : w) ptrs.push_back(&jobs[j]);
    // G_cap depends only on the LP, so twins and their warm passes agree
    choose_groups(p, ptrs, capacity, prec);
  };
  if (work.empty()) return PRIMA_OK;
  groups_for(work);
  if (ws) return run_wave(ctx, prec, jobs, work, p, certificates, capacity, ws);
  if (arena_of(work) <= budget) return run_wave(ctx, prec, jobs, work, p, certificates, capacity);
  for (auto& J : jobs)
    if (J.shadow >= 0) {
      J.cold_budget = jobs[(size_t)J.shadow].max_iter;
      J.shadow = -1;
    }
  std::vector<int> wave;
  for (int j : work) {
    if (jobs[j].shadow_of >= 0) continue;
    wave.push_back(j);
    if (wave.size() > 1 && arena_of(wave) > budget) {
      wave.pop_back();
      groups_for(wave);
      const int rc = run_wave(ctx, prec, jobs, wave, p, certificates, capacity);
      if (rc) return rc;
      wave.assign(1, j);
    }
  }
  if (wave.empty()) return PRIMA_OK;
  groups_for(wave);
  return run_wave(ctx, prec, jobs, wave, p, certificates, capacity);
}

// Runs body(i) for i in [0, count) on up to 16 host threads (the host
// preparation of a batch is independent per LP).
template <typename F>
void parallel_for(size_t count, F body) {
  const size_t hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), count);
  if (nt <= 1) {
    for (size_t i = 0; i < count; ++i) body(i);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> pool;
  for (size_t t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (size_t i = next.fetch_add(1); i < count; i = next.fetch_add(1)) body(i);
    });
  for (auto& th : pool) th.join();
}

// Pdhg.scala:272-293: warm vectors must be finite.
std::string validate_warm(const prima_lp* lp, const prima_warm* warm) {
  if (!warm) return "";
  if (warm->primal)
    for (int32_t i = 0; i < lp->n; ++i)
      if (!std::isfinite(warm->primal[i])) return "warm start primal vector is not finite";
  if (warm->dual)
    for (int32_t r = 0; r < lp->m; ++r)
      if (!std::isfinite(warm->dual[r])) return "warm start dual vector is not finite";
  return "";
}

// Host preparation of a set of jobs: validation, matrix equilibration and
// transpose (once per distinct matrix, in parallel, through the context's
// content-addressed cache) and the scaled vectors. jobs[j] gets lps[j] and
// warms[j] (which may be null). The cache keeps every matrix of the call, so
// the two passes of a mixed-precision batch and branch-and-bound nodes
// prepare each matrix once.
int prepare_jobs(prima_ctx* ctx, std::vector<SolveJob>& jobs,
                 const std::vector<const prima_lp*>& lps, const prima_params* p,
                 const std::vector<const prima_warm*>& warms) {
  const size_t N = jobs.size();
  std::vector<std::string> err(N);
  std::vector<uint64_t> keys(N, 0);
  std::vector<clk::time_point> t0(N);
  parallel_for(N, [&](size_t j) {
    t0[j] = clk::now();
    err[j] = validate_lp(lps[j]);
    if (err[j].empty()) err[j] = validate_warm(lps[j], warms[j]);
    if (err[j].empty()) keys[j] = matrix_key(lps[j], p);
  });
  for (size_t j = 0; j < N; ++j)
    if (!err[j].empty()) return fail(ctx, PRIMA_EINVAL, err[j]);
  // distinct matrices not yet prepared
  std::vector<std::shared_ptr<const MatrixPrep>> prep(N);
  std::vector<size_t> todo;
  for (size_t j = 0; j < N; ++j) {
    for (const auto& c : ctx->prep_cache)
      if (c->key == keys[j]) prep[j] = c;
    if (prep[j]) continue;
    bool dup = false;
    for (size_t t : todo) dup |= keys[t] == keys[j];
    if (!dup) todo.push_back(j);
  }
  std::vector<std::shared_ptr<MatrixPrep>> fresh(todo.size());
  parallel_for(todo.size(), [&](size_t t) {
    fresh[t] = std::const_pointer_cast<MatrixPrep>(prepare_matrix(lps[todo[t]], p));
    fresh[t]->key = keys[todo[t]];
  });
  for (auto& M : fresh) ctx->prep_cache.push_back(M);
  for (size_t j = 0; j < N; ++j)
    if (!prep[j])
      for (const auto& M : fresh)
        if (M->key == keys[j]) prep[j] = M;
  // keep what this call uses, plus the most recent few others
  {
    std::vector<std::shared_ptr<const MatrixPrep>> keep;
    const size_t others = 4;
    size_t kept_others = 0;
    for (size_t i = ctx->prep_cache.size(); i-- > 0;) {
      const auto& c = ctx->prep_cache[i];
      bool used = false;
      for (size_t j = 0; j < N && !used; ++j) used = prep[j] == c;
      bool dup = false;
      for (const auto& k : keep) dup |= k == c;
      if (dup) continue;
      if (used || kept_others++ < others) keep.push_back(c);
    }
    std::reverse(keep.begin(), keep.end());
    ctx->prep_cache.swap(keep);
  }
  // settings prima_tune_solver chose for these matrices
  for (size_t j = 0; j < N; ++j) {
    prep[j]->tuned_G = 0;
    prep[j]->tuned_bpw = -1;
    for (const auto& t : ctx->tuned)
      if (t.key == prep[j]->key) {
        prep[j]->tuned_G = t.G;
        prep[j]->tuned_bpw = t.bpw;
      }
  }
  parallel_for(N, [&](size_t j) {
    SolveJob& J = jobs[j];
    const prima_lp* lp = lps[j];
    J.lp = lp;
    J.warm = warms[j];
    HostLp& h = J.h;
    h.src = lp;
    h.n = lp->n;
    h.m = lp->m;
    h.n_eq = lp->n_eq;
    h.nnz = lp->nnz;
    h.M = prep[j];
    scale_vectors(lp, *h.M, h);
    h.rhs_norm = norm2(lp->q, lp->m);
    h.obj_norm = norm2(lp->c, lp->n);
    J.prep_ms = ms_since(t0[j]);
  });
  return PRIMA_OK;
}

prima_params reduced_params(const prima_params* p, const prima_mixed_params* mx) {
  // MixedPrecision.scala:115-130
  prima_params r = *p;
  r.eps_abs = std::max(p->eps_abs, mx->reduced_tolerance);
  r.eps_rel = std::max(p->eps_rel, mx->reduced_tolerance);
  r.max_iterations = std::min(p->max_iterations, mx->reduced_max_iterations);
  if (p->time_limit_ms >= 0)
    r.time_limit_ms = std::max<int64_t>(1, (int64_t)(p->time_limit_ms * mx->reduced_time_share));
  r.infeasibility_tolerance = std::max(p->infeasibility_tolerance, mx->reduced_tolerance);
  return r;
}

std::string validate_mixed(const prima_mixed_params* mx) {
  if (mx->refinement != 0 && mx->refinement != 1) return "refinement must be 0 or 1";
  if (mx->refinement == 1 && (mx->max_refinements < 0 || !(mx->inner_tolerance > 0.0) ||
                              mx->inner_max_iterations <= 0))
    return "iterative refinement needs max_refinements >= 0, inner_tolerance > 0 and "
           "inner_max_iterations > 0";
  if (!(mx->fix_threshold >= 0.0) || std::isinf(mx->fix_threshold))
    return "fix_threshold must be finite and >= 0";
  if (!(mx->reduced_tolerance > 0.0)) return "reducedTolerance must be positive";
  if (mx->reduced_max_iterations <= 0) return "reducedMaxIterations must be positive";
  if (!(mx->reduced_time_share > 0.0 && mx->reduced_time_share < 1.0))
    return "reducedTimeShare must lie in (0, 1)";
  return "";
}

bool conclusive(int s) {
  return s == PRIMA_OPTIMAL || s == PRIMA_PRIMAL_INFEASIBLE || s == PRIMA_DUAL_INFEASIBLE;
}

// The two-pass contract of MixedPrecision.solve (MixedPrecision.scala:114-199)
// for a set of LPs: float32 device pass for all, then a float64 device pass
// warm-started from each usable reduced result that is not a confirmed
// certificate.
int mixed_batch(prima_ctx* ctx, int32_t count, const prima_lp* lps, const prima_params* p,
                const prima_mixed_params* mx, const prima_warm* warms, prima_result* results,
                prima_mixed_info* infos, double** xo, double** yo, double** rco) {
  const prima_params rp = reduced_params(p, mx);
  std::vector<std::vector<double>> rx(count), ry(count), rrc(count);
  std::vector<prima_result> red(count);
  {
    std::vector<SolveJob> jobs(count);
    {
      std::vector<const prima_lp*> lp_of(count);
      std::vector<const prima_warm*> warm_of(count);
      for (int32_t k = 0; k < count; ++k) {
        lp_of[k] = &lps[k];
        warm_of[k] = warms ? &warms[k] : nullptr;
      }
      int rc = prepare_jobs(ctx, jobs, lp_of, &rp, warm_of);
      if (rc) return rc;
    }
    for (int32_t k = 0; k < count; ++k) {
      rx[k].assign((size_t)lps[k].n, 0.0);
      ry[k].assign((size_t)lps[k].m, 0.0);
      rrc[k].assign((size_t)lps[k].n, 0.0);
      jobs[k].res = &red[k];
      jobs[k].x_out = rx[k].data();
      jobs[k].y_out = ry[k].data();
      jobs[k].rc_out = rrc[k].data();
    }
    int rc = run_jobs(ctx, 32, jobs, &rp);
    if (rc) return rc;
  }
  std::vector<SolveJob> refine;
  std::vector<prima_warm> rw(count);
  std::vector<int> which;
  for (int32_t k = 0; k < count; ++k) {
    prima_mixed_info& info = infos[k];
    std::memset(&info, 0, sizeof(info));
    info.reduced_iterations = red[k].iterations;
    info.reduced_restarts = red[k].restarts;
    info.reduced_ms = red[k].solve_time_ms;
    const int s = red[k].status;
    const bool holds = conclusive(s) && s != PRIMA_OPTIMAL &&
                       host_classify(&lps[k], rx[k].data(), ry[k].data(),
                                     p->infeasibility_tolerance) == s;
    if (!mx->always_refine && holds) {
      results[k] = red[k];
      if (xo && xo[k]) std::memcpy(xo[k], rx[k].data(), rx[k].size() * 8);
      if (yo && yo[k]) std::memcpy(yo[k], ry[k].data(), ry[k].size() * 8);
      if (rco && rco[k]) std::memcpy(rco[k], rrc[k].data(), rrc[k].size() * 8);
      continue;
    }
    bool usable = true;
    for (double v : rx[k]) usable &= std::isfinite(v);
    for (double v : ry[k]) usable &= std::isfinite(v);
    if (usable) {
      rw[k].primal = rx[k].data();
      rw[k].dual = ry[k].data();
      rw[k].step_size = red[k].final_step_size;
      rw[k].primal_weight = red[k].final_primal_weight;
    } else {
      rw[k] = prima_warm{nullptr, nullptr, NAN, NAN};
    }
    which.push_back(k);
  }
  if (which.empty()) return PRIMA_OK;
  // one time limit for the whole refinement batch: what is left after the
  // slowest reduced pass (MixedPrecision.scala charges each LP its own)
  prima_params fp = *p;
  if (p->time_limit_ms >= 0) {
    double spent = 0;
    for (int k : which) spent = std::max(spent, red[k].solve_time_ms);
    fp.time_limit_ms = std::max<int64_t>(1, p->time_limit_ms - (int64_t)spent);
  }
  // A warm-started pass still unconverged after stall_iterations restarts
  // cold with the rest of the iteration budget, in place (see run_wave).
  const bool stall = mx->stall_iterations > 0 && mx->stall_iterations < p->max_iterations;
  // Without a time limit each warm pass gets a cold twin that races it once
  // it runs long (run_wave); under one, the warm pass restarts in place.
  const size_t N = which.size();
  const bool twins = stall && fp.time_limit_ms < 0;
  std::vector<SolveJob> jobs(twins ? 2 * N : N);
  const prima_warm none{nullptr, nullptr, NAN, NAN};
  {
    std::vector<const prima_lp*> lp_of(jobs.size());
    std::vector<const prima_warm*> warm_of(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) {
      lp_of[j] = &lps[which[j % N]];
      warm_of[j] = j >= N ? &none : &rw[which[j % N]];
    }
    int rc = prepare_jobs(ctx, jobs, lp_of, &fp, warm_of);
    if (rc) return rc;
  }
  for (size_t j = 0; j < jobs.size(); ++j) {
    const int k = which[j % N];
    const bool twin = j >= N;
    jobs[j].res = &results[k];
    jobs[j].x_out = xo ? xo[k] : nullptr;
    jobs[j].y_out = yo ? yo[k] : nullptr;
    jobs[j].rc_out = rco ? rco[k] : nullptr;
    if (!stall) continue;
    if (twin) {
      jobs[j].max_iter = p->max_iterations - mx->stall_iterations;
      jobs[j].shadow_of = (int)(j - N);
    } else {
      jobs[j].max_iter = mx->stall_iterations;
      if (twins) jobs[j].shadow = (int)(j + N);
      else jobs[j].cold_budget = p->max_iterations - mx->stall_iterations;
    }
  }
  int rc = run_jobs(ctx, 64, jobs, &fp);
  if (rc) return rc;
  for (size_t j = 0; j < N; ++j) {
    const int k = which[j];
    const bool fell_back = jobs[j].capped_iterations >= 0;
    infos[k].refinement_iterations = fell_back ? jobs[j].capped_iterations : results[k].iterations;
    infos[k].fallback_iterations = fell_back ? results[k].iterations : 0;
    infos[k].refinement_ms = results[k].solve_time_ms;
    infos[k].refined = 1;
    results[k].total_time_ms += red[k].total_time_ms;
  }
  return PRIMA_OK;
}

// The seam SpMV in a given schedule: 0 picks a sub-wave width from the mean
// row length, 1..7 are sub-waves of 1, 2, ..., 64 lanes per row, 8 is the
// row-block (CSR-Adaptive) kernel the solver uses, 9 the column-panel tiles.
hipError_t spmv_scheduled(prima_ctx* ctx, const prima_mat* a, const void* x, void* out, int sch) {
  if (sch == 0)
    return op_spmv(ctx->precision, a->rows, a->nnz, a->ptr, a->col, a->val, x, out, ctx->stream);
  if (sch == PRIMA_SPMV_ROW_BLOCKS)
    return op_spmv_blocks(ctx->precision, a->nblocks, 1, a->ptr, a->col, a->val, a->rbd, x, out,
                          ctx->stream);
  if (sch == PRIMA_SPMV_ROW_BLOCKS_SMALL)
    return op_spmv_blocks_small(ctx->precision, a->nblocks_small, a->ptr, a->col, a->val,
                                a->rbd_small, x, out, ctx->stream);
  if (sch == PRIMA_SPMV_PANELS)
    return op_spmv_panels(ctx->precision, a->panels, a->nranges, a->range_row0, a->tile_off,
                          a->starts, a->pcol, a->pval, a->nlong, a->long_rows, a->ptr, a->col,
                          a->val, x, out, ctx->stream);
  return op_spmv_vector(ctx->precision, 1 << (sch - 1), a->rows, a->ptr, a->col, a->val, x, out,
                        ctx->stream);
}

// ---------------------------------------------------------------------------
// FP64 outer iterative refinement (Gleixner, Steffy & Wolter 2016) around
// float32 device solves of correction LPs on the same matrix. The refined
// point (x, y), the original LP's double-precision matrix, its transpose and
// data stay on the device; the residuals b - A x and c - A^T y, the KKT
// error, objectives and termination test, the correction problem's data and
// the correction step all run there in float64 (refine_kernels.hip). The
// host reads one result record per LP and round and decides the scale of the
// next correction.

struct RefineState {
  std::vector<double> x, y;  // host copy of the point when it is needed there
  double dp = 1.0, dd = 1.0, rhs_norm = 0.0, obj_norm = 0.0;
  bool active = true;
  int inner_iterations = 0, inner_restarts = 0, rounds = 0;
  double step = NAN, weight = NAN;
  int mat = -1;  // index of the device copy of the original matrix
  // device arrays (offsets into the refinement arena)
  size_t dx, dy, dc, dq, dl, du, drq, drc, dcx, dcy, dcc, dcl, dcu, dcq, dylo, dpart;
  // correction LP storage (host: the correction solve prepares it there)
  std::vector<double> cc, cq, cl, cu, ylo, cx, cy;
  prima_lp corr{};
  prima_result cres{};
};

// Device copy of an original (unscaled) constraint matrix and its transpose.
struct RefMat {
  const prima_lp* lp = nullptr;
  std::vector<int32_t> tptr, tcol;
  std::vector<double> tval;
  size_t aptr = 0, acol = 0, aval = 0, dtptr = 0, dtcol = 0, dtval = 0;
};

int refine_batch(prima_ctx* ctx, int32_t count, const prima_lp* lps, const prima_params* p,
                 const prima_mixed_params* mx, const prima_warm* warms, prima_result* results,
                 prima_mixed_info* infos, double** xo, double** yo, double** rco) {
  auto t0 = clk::now();
  const prima_params rp = reduced_params(p, mx);
  std::vector<RefineState> S(count);
  std::vector<prima_result> red(count);
  std::vector<uint64_t> mkey(count, 0);
  std::vector<std::vector<double>> crc(count);
  {
    std::vector<SolveJob> jobs(count);
    {
      std::vector<const prima_lp*> lp_of(count);
      std::vector<const prima_warm*> warm_of(count);
      for (int32_t k = 0; k < count; ++k) {
        lp_of[k] = &lps[k];
        warm_of[k] = warms ? &warms[k] : nullptr;
      }
      int rc = prepare_jobs(ctx, jobs, lp_of, &rp, warm_of);
      if (rc) return rc;
    }
    for (int32_t k = 0; k < count; ++k) {
      S[k].x.assign((size_t)lps[k].n, 0.0);
      S[k].y.assign((size_t)lps[k].m, 0.0);
      crc[k].assign((size_t)lps[k].n, 0.0);
      jobs[k].res = &red[k];
      jobs[k].x_out = S[k].x.data();
      jobs[k].y_out = S[k].y.data();
      jobs[k].rc_out = crc[k].data();
      mkey[k] = jobs[k].h.M ? jobs[k].h.M->key : 0;
    }
    int rc = run_jobs(ctx, 32, jobs, &rp);
    if (rc) return rc;
  }
  const double reduced_ms = ms_since(t0);
  std::vector<int> fallback;
  std::vector<RefMat> mats;
  for (int32_t k = 0; k < count; ++k) {
    prima_mixed_info& info = infos[k];
    std::memset(&info, 0, sizeof(info));
    info.reduced_iterations = red[k].iterations;
    info.reduced_restarts = red[k].restarts;
    info.reduced_ms = red[k].solve_time_ms;
    info.refined = 1;
    RefineState& R = S[k];
    const int st = red[k].status;
    const bool holds = conclusive(st) && st != PRIMA_OPTIMAL &&
                       host_classify(&lps[k], R.x.data(), R.y.data(),
                                     p->infeasibility_tolerance) == st;
    if (!mx->always_refine && holds) {  // MixedPrecision.scala:143-165
      results[k] = red[k];
      info.refined = 0;
      if (xo && xo[k]) std::memcpy(xo[k], R.x.data(), R.x.size() * 8);
      if (yo && yo[k]) std::memcpy(yo[k], R.y.data(), R.y.size() * 8);
      if (rco && rco[k]) std::memcpy(rco[k], crc[k].data(), crc[k].size() * 8);
      R.active = false;
      continue;
    }
    bool finite = true;
    for (double v : R.x) finite &= std::isfinite(v);
    for (double v : R.y) finite &= std::isfinite(v);
    if (!finite) {  // a diverged float32 pass: the float64 pass runs cold
      std::fill(R.x.begin(), R.x.end(), 0.0);
      std::fill(R.y.begin(), R.y.end(), 0.0);
      R.active = false;
      fallback.push_back(k);
      continue;
    }
    R.rhs_norm = norm2(lps[k].q, lps[k].m);
    R.obj_norm = norm2(lps[k].c, lps[k].n);
    R.step = red[k].final_step_size;
    R.weight = red[k].final_primal_weight;
    // one device copy per distinct matrix: same content key and, to rule
    // out a hash collision, the same arrays
    for (size_t j = 0; j < mats.size() && R.mat < 0; ++j) {
      const prima_lp* a = mats[j].lp;
      const prima_lp* b = &lps[k];
      if (mkey[a - lps] != mkey[k] || a->n != b->n || a->m != b->m || a->nnz != b->nnz) continue;
      const bool same =
          (a->m == 0 || std::memcmp(a->row_ptr, b->row_ptr, ((size_t)a->m + 1) * 4) == 0) &&
          (a->nnz == 0 || (std::memcmp(a->col, b->col, (size_t)a->nnz * 4) == 0 &&
                           std::memcmp(a->val, b->val, (size_t)a->nnz * 8) == 0));
      if (same) R.mat = (int)j;
    }
    if (R.mat < 0) {
      R.mat = (int)mats.size();
      mats.push_back(RefMat{});
      mats.back().lp = &lps[k];
    }
  }

  // device arena: one copy of each distinct original matrix, then the
  // float64 vectors of every refined LP, then all result records
  static const int32_t zero_ptr[1] = {0};
  parallel_for(mats.size(), [&](size_t j) {
    const prima_lp* lp = mats[j].lp;
    transpose_csr(lp->m, lp->n, lp->m > 0 ? lp->row_ptr : zero_ptr, lp->col, lp->val,
                  mats[j].tptr, mats[j].tcol, mats[j].tval);
  });
  Carver cv;
  for (RefMat& M : mats) {
    const size_t n = (size_t)M.lp->n, m = (size_t)M.lp->m, z = (size_t)M.lp->nnz;
    M.aptr = cv.take((m + 1) * 4);
    M.acol = cv.take(z * 4);
    M.aval = cv.take(z * 8);
    M.dtptr = cv.take((n + 1) * 4);
    M.dtcol = cv.take(z * 4);
    M.dtval = cv.take(z * 8);
  }
  const size_t part_bytes = (size_t)ref_partial_blocks() * 8 * 8;
  for (int32_t k = 0; k < count; ++k) {
    RefineState& R = S[k];
    if (!R.active) continue;
    const size_t n = (size_t)lps[k].n, m = (size_t)lps[k].m;
    R.dx = cv.take(n * 8);
    R.dc = cv.take(n * 8);
    R.dl = cv.take(n * 8);
    R.du = cv.take(n * 8);
    R.drc = cv.take(n * 8);
    R.dcx = cv.take(n * 8);
    R.dcc = cv.take(n * 8);
    R.dcl = cv.take(n * 8);
    R.dcu = cv.take(n * 8);
    R.dy = cv.take(m * 8);
    R.dq = cv.take(m * 8);
    R.drq = cv.take(m * 8);
    R.dcy = cv.take(m * 8);
    R.dcq = cv.take(m * 8);
    R.dylo = cv.take(m * 8);
    R.dpart = cv.take(part_bytes);
  }
  const int RF = ref_result_doubles();
  const size_t outs_off = cv.take((size_t)count * RF * 8);
  char* arena = nullptr;
  std::vector<double> outs((size_t)count * RF, 0.0);
  hipStream_t s = ctx->stream;
  auto D = [&](size_t off) { return (double*)(arena + off); };
  auto I = [&](size_t off) { return (int32_t*)(arena + off); };
  auto up = [&](size_t off, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(arena + off, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  auto down = [&](void* dst, size_t off, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, arena + off, bytes, hipMemcpyDeviceToHost, s) : hipSuccess;
  };
  struct ArenaGuard {
    char** a;
    ~ArenaGuard() {
      if (*a) (void)hipFree(*a);
    }
  } guard{&arena};
  bool any_active = false;
  for (int32_t k = 0; k < count; ++k) any_active |= S[k].active;
  if (any_active) {
    HIP_TRY(ctx, hipMalloc((void**)&arena, std::max<size_t>(cv.off, 256)));
    for (RefMat& M : mats) {
      const prima_lp* lp = M.lp;
      const size_t n = (size_t)lp->n, m = (size_t)lp->m, z = (size_t)lp->nnz;
      HIP_TRY(ctx, up(M.aptr, lp->m > 0 ? lp->row_ptr : zero_ptr, (m + 1) * 4));
      HIP_TRY(ctx, up(M.acol, lp->col, z * 4));
      HIP_TRY(ctx, up(M.aval, lp->val, z * 8));
      HIP_TRY(ctx, up(M.dtptr, M.tptr.data(), (n + 1) * 4));
      HIP_TRY(ctx, up(M.dtcol, M.tcol.data(), z * 4));
      HIP_TRY(ctx, up(M.dtval, M.tval.data(), z * 8));
    }
    for (int32_t k = 0; k < count; ++k) {
      RefineState& R = S[k];
      if (!R.active) continue;
      const prima_lp* lp = &lps[k];
      const size_t n = (size_t)lp->n, m = (size_t)lp->m;
      HIP_TRY(ctx, up(R.dx, R.x.data(), n * 8));
      HIP_TRY(ctx, up(R.dy, R.y.data(), m * 8));
      HIP_TRY(ctx, up(R.dc, lp->c, n * 8));
      HIP_TRY(ctx, up(R.dl, lp->l, n * 8));
      HIP_TRY(ctx, up(R.du, lp->u, n * 8));
      HIP_TRY(ctx, up(R.dq, lp->q, m * 8));
    }
  }
  // the device point back into the host copy (before a float64 fallback)
  auto point_to_host = [&](int k) -> int {
    RefineState& R = S[k];
    HIP_TRY(ctx, down(R.x.data(), R.dx, R.x.size() * 8));
    HIP_TRY(ctx, down(R.y.data(), R.dy, R.y.size() * 8));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return PRIMA_OK;
  };

  prima_params ip = *p;
  ip.eps_abs = mx->inner_tolerance;
  ip.eps_rel = mx->inner_tolerance;
  ip.max_iterations = mx->inner_max_iterations;
  ip.time_limit_ms = -1;
  const double growth = 1e4, cost_cap = 1e6;
  const double fix = mx->fix_threshold;
  for (int round = 0; round <= mx->max_refinements && any_active; ++round) {
    // float64 residuals, KKT error and termination test of every active LP
    for (int32_t k = 0; k < count; ++k) {
      RefineState& R = S[k];
      if (!R.active) continue;
      const prima_lp* lp = &lps[k];
      const RefMat& M = mats[(size_t)R.mat];
      HIP_TRY(ctx, ref_evaluate(lp->n, lp->m, lp->n_eq, lp->nnz, I(M.aptr), I(M.acol),
                                D(M.aval), I(M.dtptr), I(M.dtcol), D(M.dtval), D(R.dx),
                                D(R.dy), D(R.dc), D(R.dq), D(R.dl), D(R.du), p->eps_abs,
                                p->eps_rel, R.rhs_norm, R.obj_norm, lp->offset, D(R.drq),
                                D(R.drc), D(R.dpart), D(outs_off) + (size_t)k * RF, s));
    }
    HIP_TRY(ctx, down(outs.data(), outs_off, outs.size() * 8));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    std::vector<int> act;
    for (int32_t k = 0; k < count; ++k) {
      RefineState& R = S[k];
      if (!R.active) continue;
      const prima_lp* lp = &lps[k];
      const double* e = &outs[(size_t)k * RF];  // pres, dres, pobj, dobj, pv, dv, converged
      if (e[6] != 0.0) {
        prima_result& r = results[k];
        std::memset(&r, 0, sizeof(r));
        r.status = PRIMA_OPTIMAL;
        r.iterations = red[k].iterations + R.inner_iterations;
        r.restarts = red[k].restarts + R.inner_restarts;
        r.objective = e[2];
        r.dual_objective = e[3];
        r.primal_residual = e[0];
        r.dual_residual = e[1];
        r.final_step_size = R.step;
        r.final_primal_weight = R.weight;
        r.workgroups = red[k].workgroups;
        if (xo && xo[k]) HIP_TRY(ctx, down(xo[k], R.dx, (size_t)lp->n * 8));
        if (yo && yo[k]) HIP_TRY(ctx, down(yo[k], R.dy, (size_t)lp->m * 8));
        if (rco && rco[k]) HIP_TRY(ctx, down(rco[k], R.drc, (size_t)lp->n * 8));
        infos[k].rounds = R.rounds;
        infos[k].refinement_iterations = R.inner_iterations;
        R.active = false;
        continue;
      }
      if (round == mx->max_refinements) {
        R.active = false;
        fallback.push_back(k);
        int rc = point_to_host(k);
        if (rc) return rc;
        continue;
      }
      // scaled residual problem: primal and dual violations brought to O(1)
      R.dp = std::min(1.0 / std::max(e[4], 1e-300), growth * R.dp);
      R.dd = std::min(1.0 / std::max(e[5], 1e-300), growth * R.dd);
      const int32_t n = lp->n, m = lp->m;
      HIP_TRY(ctx, ref_build(n, m, lp->n_eq, R.dp, R.dd, cost_cap, fix, D(R.drc), D(R.drq), D(R.dx),
                             D(R.dy), D(R.dl), D(R.du), D(R.dcc), D(R.dcl), D(R.dcu), D(R.dcq),
                             D(R.dylo), s));
      R.cc.resize(n);
      R.cl.resize(n);
      R.cu.resize(n);
      R.cq.resize(m);
      R.ylo.resize(m);
      HIP_TRY(ctx, down(R.cc.data(), R.dcc, (size_t)n * 8));
      HIP_TRY(ctx, down(R.cl.data(), R.dcl, (size_t)n * 8));
      HIP_TRY(ctx, down(R.cu.data(), R.dcu, (size_t)n * 8));
      HIP_TRY(ctx, down(R.cq.data(), R.dcq, (size_t)m * 8));
      HIP_TRY(ctx, down(R.ylo.data(), R.dylo, (size_t)m * 8));
      R.corr = *lp;
      R.corr.c = R.cc.data();
      R.corr.q = R.cq.data();
      R.corr.l = R.cl.data();
      R.corr.u = R.cu.data();
      R.corr.offset = 0.0;
      R.cx.assign((size_t)n, 0.0);
      R.cy.assign((size_t)m, 0.0);
      act.push_back(k);
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (act.empty()) break;
    std::vector<SolveJob> jobs(act.size());
    {
      std::vector<const prima_lp*> lp_of(act.size());
      for (size_t j = 0; j < act.size(); ++j) lp_of[j] = &S[act[j]].corr;
      int rc = prepare_jobs(ctx, jobs, lp_of, &ip,
                            std::vector<const prima_warm*>(act.size(), nullptr));
      if (rc) return rc;
    }
    for (size_t j = 0; j < act.size(); ++j) {
      RefineState& R = S[act[j]];
      jobs[j].ylo = R.ylo.data();
      jobs[j].res = &R.cres;
      jobs[j].x_out = R.cx.data();
      jobs[j].y_out = R.cy.data();
    }
    int rc = run_jobs(ctx, 32, jobs, &ip, /*certificates=*/false);
    if (rc) return rc;
    for (int k : act) {
      RefineState& R = S[k];
      const prima_lp* lp = &lps[k];
      R.rounds++;
      R.inner_iterations += R.cres.iterations;
      R.inner_restarts += R.cres.restarts;
      // a correction that diverged (an infeasible or unbounded LP has an
      // infeasible or unbounded correction problem too) ends refinement:
      // the LP takes the float64 pass from the last finite point
      bool finite = R.cres.status == PRIMA_OPTIMAL || R.cres.status == PRIMA_ITERATION_LIMIT;
      for (double v : R.cx) finite &= std::isfinite(v);
      for (double v : R.cy) finite &= std::isfinite(v);
      if (!finite) {
        R.active = false;
        fallback.push_back(k);
        int rc2 = point_to_host(k);
        if (rc2) return rc2;
        continue;
      }
      HIP_TRY(ctx, up(R.dcx, R.cx.data(), (size_t)lp->n * 8));
      HIP_TRY(ctx, up(R.dcy, R.cy.data(), (size_t)lp->m * 8));
      HIP_TRY(ctx, ref_apply(lp->n, lp->m, lp->n_eq, R.dp, R.dd, D(R.dcx), D(R.dcy), D(R.dl),
                             D(R.du), D(R.dx), D(R.dy), s));
    }
  }
  HIP_TRY(ctx, hipStreamSynchronize(s));
  const double refine_ms = ms_since(t0) - reduced_ms;
  for (int32_t k = 0; k < count; ++k) {
    if (std::find(fallback.begin(), fallback.end(), k) != fallback.end()) continue;
    if (infos[k].refined) {
      results[k].solve_time_ms = ms_since(t0);
      results[k].total_time_ms = results[k].solve_time_ms;
      infos[k].refinement_ms = refine_ms;
    }
  }
  if (fallback.empty()) return PRIMA_OK;
  // the reference's float64 pass, warm-started from the refined point
  std::vector<prima_warm> fw(fallback.size());
  std::vector<SolveJob> jobs(fallback.size());
  std::vector<prima_result> fres(fallback.size());
  {
    std::vector<const prima_lp*> lp_of(fallback.size());
    std::vector<const prima_warm*> warm_of(fallback.size());
    for (size_t j = 0; j < fallback.size(); ++j) {
      const int k = fallback[j];
      RefineState& R = S[k];
      fw[j] = prima_warm{R.x.data(), R.y.data(), R.step, R.weight};
      lp_of[j] = &lps[k];
      warm_of[j] = &fw[j];
    }
    int rc = prepare_jobs(ctx, jobs, lp_of, p, warm_of);
    if (rc) return rc;
  }
  for (size_t j = 0; j < fallback.size(); ++j) {
    const int k = fallback[j];
    jobs[j].res = &fres[j];
    jobs[j].x_out = xo ? xo[k] : nullptr;
    jobs[j].y_out = yo ? yo[k] : nullptr;
    jobs[j].rc_out = rco ? rco[k] : nullptr;
  }
  int rc = run_jobs(ctx, 64, jobs, p);
  if (rc) return rc;
  for (size_t j = 0; j < fallback.size(); ++j) {
    const int k = fallback[j];
    results[k] = fres[j];
    infos[k].fallback_iterations = fres[j].iterations;
    infos[k].refinement_iterations = S[k].inner_iterations + fres[j].iterations;
    infos[k].rounds = S[k].rounds;
    infos[k].refinement_ms = ms_since(t0) - reduced_ms;
    results[k].solve_time_ms = ms_since(t0);
    results[k].total_time_ms = results[k].solve_time_ms;
  }
  return PRIMA_OK;
}


// ---------------------------------------------------------------------------
// Branch-and-bound (BranchAndBound.scala:168-469) with a batched frontier.

struct BnbWarm {
  std::vector<double> x, y;
  double step = NAN, weight = NAN;
};

struct BnbNode {
  std::vector<double> lower, upper;
  double safe_bound;
  std::shared_ptr<const BnbWarm> warm;
};

// skipsBeforeSolving (BranchAndBound.scala:407-413): the negation of >= so a
// NaN bound falls through to solving.
bool skips_before_solving(double safe_bound, double incumbent, bool have) {
  return have && safe_bound >= incumbent;
}

// safeBoundOf (BranchAndBound.scala:422-427).
double safe_bound_of(bool conclusive, double bound, double slack) {
  if (!conclusive) return -INFINITY;
  const double c = bound - slack;
  return std::isfinite(c) ? c : -INFINITY;
}

// Java's Math.round as a double: nearest whole number, ties toward +inf,
// NaN -> 0 (BranchAndBound.scala:447 and :461 call it).
double java_round(double v) {
  if (std::isnan(v)) return 0.0;
  if (!std::isfinite(v) || std::fabs(v) >= 4503599627370496.0) return v;  // 2^52: already whole
  const double f = std::floor(v);
  return (v - f >= 0.5) ? f + 1.0 : f;
}

// mostFractional (BranchAndBound.scala:441-453): -1 when all are integral.
int most_fractional(const double* x, const std::vector<int32_t>& ints, double tol) {
  int best = -1;
  double worst = tol;
  for (int32_t j : ints) {
    const double g = std::fabs(x[j] - java_round(x[j]));
    if (g > worst) {
      worst = g;
      best = j;
    }
  }
  return best;
}

}  // namespace

extern "C" void prima_default_bnb_params(prima_bnb_params* p) {
  std::memset(p, 0, sizeof(*p));
  prima_default_params(&p->lp);
  p->integrality_tolerance = 1e-6;
  p->gap_abs = 1e-6;
  p->gap_rel = 1e-6;
  p->pruning_safety_factor = 10.0;
  p->max_nodes = 100000;
  p->warm_start = 1;
  p->time_limit_ms = -1;
  p->frontier = 1;
}

extern "C" int prima_bnb_skips_before_solving(double safe_bound, double incumbent,
                                              int32_t have_incumbent) {
  return skips_before_solving(safe_bound, incumbent, have_incumbent != 0) ? 1 : 0;
}

extern "C" double prima_bnb_safe_bound_of(int32_t conclusive, double bound, double slack) {
  return safe_bound_of(conclusive != 0, bound, slack);
}

extern "C" int prima_milp_solve(prima_ctx* ctx, const prima_lp* lp, int32_t n_int,
                                const int32_t* integers, const prima_bnb_params* bp,
                                const prima_mixed_params* mixed, prima_milp_result* res,
                                double* x_out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!lp || !bp || !res || n_int < 0 || (n_int > 0 && !integers))
    return fail(ctx, PRIMA_EINVAL, "bad branch-and-bound arguments");
  std::string e = validate_lp(lp);
  if (!e.empty()) return fail(ctx, PRIMA_EINVAL, e);
  for (int32_t k = 0; k < n_int; ++k)
    if (integers[k] < 0 || integers[k] >= lp->n)
      return fail(ctx, PRIMA_EINVAL, "integer column " + std::to_string(integers[k]) +
                                         " is out of range");
  if (bp->frontier < 1 || bp->max_nodes < 1) return fail(ctx, PRIMA_EINVAL, "bad frontier");
  const auto t0 = clk::now();
  std::memset(res, 0, sizeof(*res));
  const int32_t n = lp->n, m = lp->m;
  auto solve_nodes = [&](std::vector<prima_lp>& subs, std::vector<prima_warm>& warms,
                         std::vector<prima_result>& out, std::vector<std::vector<double>>& xs,
                         std::vector<std::vector<double>>& ys) -> int {
    const int32_t k = (int32_t)subs.size();
    out.assign(k, prima_result{});
    xs.assign(k, std::vector<double>((size_t)n));
    ys.assign(k, std::vector<double>((size_t)m));
    std::vector<double*> px(k), py(k), prc(k, nullptr);
    for (int32_t i = 0; i < k; ++i) {
      px[i] = xs[i].data();
      py[i] = ys[i].data();
    }
    std::vector<prima_mixed_info> infos(k);
    res->batches++;
    res->relaxations += k;
    return prima_solve_batch(ctx, k, subs.data(), &bp->lp, mixed, warms.data(), out.data(),
                             mixed ? infos.data() : nullptr, px.data(), py.data(), prc.data());
  };

  if (n_int == 0) {  // BranchAndBound.scala:181-207: the relaxation is the answer
    std::vector<prima_lp> subs{*lp};
    std::vector<prima_warm> warms{prima_warm{nullptr, nullptr, NAN, NAN}};
    std::vector<prima_result> out;
    std::vector<std::vector<double>> xs, ys;
    rc = solve_nodes(subs, warms, out, xs, ys);
    if (rc) return rc;
    const int st = out[0].status;
    res->status = st == PRIMA_OPTIMAL            ? PRIMA_MILP_OPTIMAL
                  : st == PRIMA_PRIMAL_INFEASIBLE ? PRIMA_MILP_INFEASIBLE
                  : st == PRIMA_DUAL_INFEASIBLE   ? PRIMA_MILP_UNBOUNDED
                                                  : PRIMA_MILP_NO_SOLUTION;
    const bool solved = res->status == PRIMA_MILP_OPTIMAL;
    res->objective = solved ? out[0].objective : INFINITY;
    res->best_bound = solved ? out[0].objective : -INFINITY;
    res->nodes_explored = 1;
    res->unproven_nodes = conclusive(st) ? 0 : 1;
    res->has_incumbent = solved ? 1 : 0;
    if (solved && x_out) std::memcpy(x_out, xs[0].data(), (size_t)n * 8);
    res->solve_time_ms = ms_since(t0);
    return PRIMA_OK;
  }

  std::vector<int32_t> ordered(integers, integers + n_int);
  std::sort(ordered.begin(), ordered.end());
  ordered.erase(std::unique(ordered.begin(), ordered.end()), ordered.end());
  std::vector<BnbNode> stack;
  stack.push_back(BnbNode{std::vector<double>(lp->l, lp->l + n),
                          std::vector<double>(lp->u, lp->u + n), -INFINITY, nullptr});
  std::vector<double> incumbent;
  double inc_val = INFINITY;
  int nodes = 0, unproven = 0;
  bool root_unbounded = false;
  auto out_of_time = [&]() {
    return bp->time_limit_ms >= 0 && ms_since(t0) > (double)bp->time_limit_ms;
  };
  while (!stack.empty() && nodes < bp->max_nodes && !out_of_time()) {
    // pop the frontier, applying the pre-solve skip per node as the
    // sequential search would at its pop
    std::vector<BnbNode> batch;
    while (!stack.empty() && (int)batch.size() < bp->frontier &&
           nodes + (int)batch.size() < bp->max_nodes) {
      BnbNode node = std::move(stack.back());
      stack.pop_back();
      if (skips_before_solving(node.safe_bound, inc_val, !incumbent.empty())) continue;
      batch.push_back(std::move(node));
    }
    if (batch.empty()) continue;
    std::vector<prima_lp> subs(batch.size(), *lp);
    std::vector<prima_warm> warms(batch.size(), prima_warm{nullptr, nullptr, NAN, NAN});
    for (size_t i = 0; i < batch.size(); ++i) {
      subs[i].l = batch[i].lower.data();
      subs[i].u = batch[i].upper.data();
      if (bp->warm_start && batch[i].warm) {
        const BnbWarm& w = *batch[i].warm;
        warms[i] = prima_warm{w.x.data(), w.y.data(), w.step, w.weight};
      }
    }
    std::vector<prima_result> out;
    std::vector<std::vector<double>> xs, ys;
    rc = solve_nodes(subs, warms, out, xs, ys);
    if (rc) return rc;
    for (size_t i = 0; i < batch.size(); ++i) {
      const BnbNode& node = batch[i];
      const prima_result& rel = out[i];
      // an earlier node of this batch may have raised the incumbent past this
      // node's parent bound: the sequential search would not have solved it
      if (skips_before_solving(node.safe_bound, inc_val, !incumbent.empty())) continue;
      nodes++;
      if (rel.status == PRIMA_PRIMAL_INFEASIBLE) continue;
      if (rel.status == PRIMA_DUAL_INFEASIBLE) {
        if (nodes == 1) root_unbounded = true;
        else unproven++;
        continue;
      }
      const bool concl = conclusive(rel.status);
      if (!concl) unproven++;
      const double bound = rel.objective;
      const double slack = bp->pruning_safety_factor *
                           std::max(std::fabs(rel.objective - rel.dual_objective), bp->lp.eps_abs);
      const double safe = safe_bound_of(concl, bound, slack);
      const bool prunable = concl && !incumbent.empty() && safe >= inc_val;
      if (prunable) continue;
      const int j = most_fractional(xs[i].data(), ordered, bp->integrality_tolerance);
      if (j < 0) {
        if (!concl) continue;  // integral but unconverged: never an incumbent
        std::vector<double> snapped = xs[i];
        for (int32_t c : ordered) snapped[c] = java_round(snapped[c]);
        double value = lp->offset;
        for (int32_t c = 0; c < n; ++c) value += lp->c[c] * snapped[c];
        if (value < inc_val) {
          inc_val = value;
          incumbent = std::move(snapped);
        }
        continue;
      }
      std::shared_ptr<const BnbWarm> warm;
      if (bp->warm_start) {
        bool ok = true;  // Pdhg.WarmStart's preconditions: finite vectors
        for (double v : xs[i]) ok &= std::isfinite(v);
        for (double v : ys[i]) ok &= std::isfinite(v);
        if (ok) {
          auto w = std::make_shared<BnbWarm>();
          w->x = xs[i];
          w->y = ys[i];
          w->step = rel.final_step_size;
          w->weight = rel.final_primal_weight;
          warm = w;
        }
      }
      const double value = xs[i][j];
      const double fl = std::floor(value), ce = std::ceil(value);
      // ceiling first so the floor branch is explored first
      if (ce <= node.upper[j]) {
        BnbNode child{node.lower, node.upper, safe, warm};
        child.lower[j] = std::max(child.lower[j], ce);
        stack.push_back(std::move(child));
      }
      if (fl >= node.lower[j]) {
        BnbNode child{node.lower, node.upper, safe, warm};
        child.upper[j] = std::min(child.upper[j], fl);
        stack.push_back(std::move(child));
      }
    }
  }
  double open_bound = inc_val;
  for (const auto& nd : stack) open_bound = std::min(open_bound, nd.safe_bound);
  const double bound = stack.empty() ? inc_val : open_bound;
  auto closed = [&](double inc, double b) {
    const double a = std::fabs(inc - b);
    return a <= bp->gap_abs || a <= bp->gap_rel * std::max(1.0, std::fabs(inc));
  };
  int status;
  if (root_unbounded) status = PRIMA_MILP_UNBOUNDED;
  else if (incumbent.empty() && stack.empty() && unproven == 0) status = PRIMA_MILP_INFEASIBLE;
  else if (incumbent.empty()) status = PRIMA_MILP_NO_SOLUTION;
  else if (unproven > 0) status = PRIMA_MILP_FEASIBLE;
  else if (stack.empty()) status = PRIMA_MILP_OPTIMAL;
  else if (closed(inc_val, bound)) status = PRIMA_MILP_OPTIMAL;
  else status = PRIMA_MILP_FEASIBLE;
  res->status = status;
  res->objective = inc_val;
  res->best_bound = incumbent.empty() ? bound : std::min(bound, inc_val);
  res->nodes_explored = nodes;
  res->unproven_nodes = unproven;
  res->has_incumbent = incumbent.empty() ? 0 : 1;
  res->solve_time_ms = ms_since(t0);
  if (!incumbent.empty() && x_out) std::memcpy(x_out, incumbent.data(), (size_t)n * 8);
  return PRIMA_OK;
}

namespace {
}  // namespace

// ===========================================================================
// C ABI

extern "C" {

void prima_default_params(prima_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->eps_abs = 1e-8;
  p->eps_rel = 1e-8;
  p->max_iterations = 100000;
  p->evaluation_frequency = 64;
  p->time_limit_ms = -1;
  p->ruiz_iterations = 10;
  p->apply_pock_chambolle = 1;
  p->pock_chambolle_alpha = 1.0;
  p->infeasibility_tolerance = 1e-8;
  p->restart_sufficient = 0.2;
  p->restart_necessary = 0.8;
  p->restart_artificial = 0.36;
  p->primal_weight_smoothing = 0.5;
  p->workgroups_per_lp = 0;
  p->chunk_iterations = 0;
}

void prima_default_mixed_params(prima_mixed_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->reduced_tolerance = 1e-5;
  p->reduced_max_iterations = 20000;
  p->always_refine = 0;
  p->reduced_time_share = 0.5;
  p->refinement = 0;
  p->max_refinements = 4;
  p->inner_tolerance = 1e-4;
  p->inner_max_iterations = 20000;
}

int prima_open(int32_t device, int32_t precision, prima_ctx** out) {
  if (!out) return PRIMA_EINVAL;
  *out = nullptr;
  if (precision != 64 && precision != 32) return PRIMA_EINVAL;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return PRIMA_EDEVICE;
  if (hipSetDevice(device) != hipSuccess) return PRIMA_EDEVICE;
  prima_ctx* ctx = new prima_ctx();
  ctx->device = device;
  ctx->precision = precision;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    delete ctx;
    return PRIMA_EDEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->cu_count = prop.multiProcessorCount;
    ctx->device_name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
  }
  int coop = 0, xccs = 0;
  (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device);
  if (hipDeviceGetAttribute(&xccs, hipDeviceAttributeNumberOfXccs, device) == hipSuccess &&
      xccs > 0)
    ctx->xcds = xccs;
  ctx->occ[0] = occupancy_blocks_per_cu(32);
  ctx->occ[1] = occupancy_blocks_per_cu(64);
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipMalloc((void**)&ctx->scratch, 2048 * sizeof(double)) != hipSuccess) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return PRIMA_ENOMEM;
  }
  ctx->coop_ok = coop && ctx->occ[0] > 0 && ctx->occ[1] > 0 && ctx->cu_count > 0;
  // the allocation-free solve's workspace on this device (kept per device)
  ctx->scratch_ok = scratch_init(device) == hipSuccess;
  *out = ctx;
  return PRIMA_OK;
}

int prima_close(prima_ctx* ctx) {
  if (!ctx) return PRIMA_EINVAL;
  if (ctx->closed) return PRIMA_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  ctx->ev0 = ctx->ev1 = nullptr;
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->scratch = nullptr;
  ctx->stream = nullptr;
  ctx->closed = true;
  return PRIMA_OK;
}

void prima_destroy(prima_ctx* ctx) {
  if (!ctx) return;
  prima_close(ctx);
  delete ctx;
}

int prima_capabilities_of(const prima_ctx* ctx, prima_capabilities* out) {
  if (!ctx || !out) return PRIMA_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::snprintf(out->name, sizeof(out->name), "%s",
                ctx->precision == 64 ? "prima-hip-f64" : "prima-hip-f32");
  std::snprintf(out->device, sizeof(out->device), "%s", ctx->device_name.c_str());
  out->supports_float64 = ctx->precision == 64 ? 1 : 0;
  return PRIMA_OK;
}

const char* prima_last_error(const prima_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int prima_stats_get(const prima_ctx* ctx, prima_stats* out) {
  if (!ctx || !out) return PRIMA_EINVAL;
  *out = ctx->stats;
  return PRIMA_OK;
}

int prima_stats_reset(prima_ctx* ctx) {
  if (!ctx) return PRIMA_EINVAL;
  std::memset(&ctx->stats, 0, sizeof(ctx->stats));
  return PRIMA_OK;
}

int prima_progress(const prima_ctx* ctx, int64_t* launches, int64_t* accepted_iterations,
                   int64_t* finished_lps) {
  if (!ctx) return PRIMA_EINVAL;
  if (launches) *launches = ctx->prog_launches.load();
  if (accepted_iterations) *accepted_iterations = ctx->prog_iterations.load();
  if (finished_lps) *finished_lps = ctx->prog_finished.load();
  return PRIMA_OK;
}

int prima_cancel(prima_ctx* ctx) {
  if (!ctx) return PRIMA_EINVAL;
  ctx->cancel.store(1);
  return PRIMA_OK;
}

int prima_test_stall_barriers(prima_ctx* ctx, int32_t enable) {
  int rc = check_open(ctx);
  if (rc) return rc;
  ctx->test_stall = enable ? 1 : 0;
  return PRIMA_OK;
}

int prima_vec_alloc(prima_ctx* ctx, int64_t n, prima_vec** out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!out || n < 0) return fail(ctx, PRIMA_EINVAL, "bad vector length");
  prima_vec* v = new prima_vec();
  v->n = n;
  const size_t bytes = std::max<size_t>(1, (size_t)n * esize(ctx->precision));
  if (hipMalloc(&v->d, bytes) != hipSuccess) {
    delete v;
    return fail(ctx, PRIMA_ENOMEM, "vector allocation");
  }
  hipError_t e = hipMemsetAsync(v->d, 0, bytes, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(v->d);
    delete v;
    return dev_fail(ctx, e, "zero-fill");
  }
  *out = v;
  return PRIMA_OK;
}

int prima_vec_upload(prima_ctx* ctx, const double* data, int64_t n, prima_vec** out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (n > 0 && !data) return fail(ctx, PRIMA_EINVAL, "null data");
  prima_vec* v = nullptr;
  rc = prima_vec_alloc(ctx, n, &v);
  if (rc) return rc;
  if (n > 0) {
    std::vector<char> tmp((size_t)n * esize(ctx->precision));
    to_prec(ctx->precision, data, n, tmp.data());
    const hipError_t e = hipMemcpy(v->d, tmp.data(), tmp.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)prima_vec_free(ctx, v);
      return dev_fail(ctx, e, "vector upload");
    }
  }
  *out = v;
  return PRIMA_OK;
}

int prima_vec_download(prima_ctx* ctx, const prima_vec* v, double* into, int64_t n) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!v || n != v->n) return fail(ctx, PRIMA_EINVAL, "length mismatch");
  if (n == 0) return PRIMA_OK;
  if (!into) return fail(ctx, PRIMA_EINVAL, "null destination");
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->precision == 64) {
    HIP_TRY(ctx, hipMemcpy(into, v->d, (size_t)n * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp((size_t)n);
    HIP_TRY(ctx, hipMemcpy(tmp.data(), v->d, (size_t)n * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; ++i) into[i] = (double)tmp[i];
  }
  return PRIMA_OK;
}

int64_t prima_vec_length(const prima_vec* v) { return v ? v->n : -1; }

int prima_vec_free(prima_ctx* ctx, prima_vec* v) {
  if (!v) return PRIMA_OK;
  if (ctx && !ctx->closed) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  if (v->d) (void)hipFree(v->d);
  delete v;
  return PRIMA_OK;
}

// Column-panel layout of a CSR matrix (PRIMA_SPMV_PANELS). Panels are 512 KiB
// of x wide; rows are grouped greedily into ranges of at most panel_rows()
// rows whose tiles each hold at most panel_tile() entries; a range's entries
// are stored panel by panel, each tile row by row in the row's stored order.
// Rows longer than 512 entries stay out of the tiles.
static int build_panels(prima_ctx* ctx, prima_mat* M, const int32_t* row_ptr, const int32_t* col,
                        const double* vals) {
  const int prec = ctx->precision;
#ifndef PRIMA_PANEL_BYTES
#define PRIMA_PANEL_BYTES (1 << 19)
#endif
  const int64_t W = PRIMA_PANEL_BYTES / (prec == 32 ? 4 : 8);
  if ((int64_t)M->cols * (prec == 32 ? 4 : 8) < 4 * (1 << 20)) return PRIMA_OK;
  const int P = (int)((M->cols + W - 1) / W);
  const int R = panel_rows(), TP = panel_tile();
  // The layout pays for scattered columns; it costs 2 bytes per row per
  // panel of row starts and P steps per row to build. A matrix whose row
  // starts would outweigh its column indices (many panels, few entries per
  // row, as in a block-diagonal batch of LPs whose gathers are already
  // local) keeps CSR only.
  if (P > 256 || (double)P * ((double)M->rows / R + 1) * (R + 1) * 2 > 4.0 * (double)M->nnz)
    return PRIMA_OK;
  const int LONG = 512;
  const int rows = M->rows;
  std::vector<int32_t> range_row0{0}, long_rows;
  std::vector<int32_t> cnt(P), rc(P);
  std::vector<char> is_long((size_t)rows, 0);
  int r = 0;
  while (r < rows) {
    std::fill(cnt.begin(), cnt.end(), 0);
    int nr = 0;
    while (r < rows && nr < R) {
      const int len = row_ptr[r + 1] - row_ptr[r];
      bool lg = len > LONG;
      if (!lg) {
        std::fill(rc.begin(), rc.end(), 0);
        for (int32_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) rc[col[q] / W]++;
        bool fits = true;
        for (int p = 0; p < P; ++p) fits &= cnt[p] + rc[p] <= TP;
        if (!fits && nr > 0) break;  // close the range; the row opens the next one
        if (!fits) return fail(ctx, PRIMA_EINVAL, "row does not fit a panel tile");
        for (int p = 0; p < P; ++p) cnt[p] += rc[p];
      } else {
        long_rows.push_back(r);
        is_long[r] = 1;
      }
      ++r;
      ++nr;
    }
    range_row0.push_back(r);
  }
  const int G = (int)range_row0.size() - 1;
  std::vector<int32_t> tile_off((size_t)G * P + 1);
  std::vector<uint16_t> starts((size_t)G * P * (R + 1), 0);
  std::vector<int32_t> pcol;
  std::vector<double> pv;
  pcol.reserve((size_t)M->nnz);
  pv.reserve((size_t)M->nnz);
  std::vector<int32_t> counts, pos;
  int64_t total = 0;
  for (int g = 0; g < G; ++g) {
    const int r0 = range_row0[g], nr = range_row0[g + 1] - r0;
    counts.assign((size_t)P * nr, 0);
    for (int rl = 0; rl < nr; ++rl) {
      if (is_long[r0 + rl]) continue;
      for (int32_t q = row_ptr[r0 + rl]; q < row_ptr[r0 + rl + 1]; ++q)
        counts[(size_t)(col[q] / W) * nr + rl]++;
    }
    pos.assign((size_t)P * nr, 0);
    for (int p = 0; p < P; ++p) {
      const size_t t = (size_t)g * P + p;
      tile_off[t] = (int32_t)total;
      uint16_t* st = &starts[t * (R + 1)];
      int run = 0;
      for (int rl = 0; rl < nr; ++rl) {
        st[rl] = (uint16_t)run;
        pos[(size_t)p * nr + rl] = (int32_t)(total + run);
        run += counts[(size_t)p * nr + rl];
      }
      st[nr] = (uint16_t)run;
      total += run;
    }
    if (total > INT32_MAX) return fail(ctx, PRIMA_EINVAL, "matrix too large for panels");
    pcol.resize((size_t)total);
    pv.resize((size_t)total);
    for (int rl = 0; rl < nr; ++rl) {
      if (is_long[r0 + rl]) continue;
      for (int32_t q = row_ptr[r0 + rl]; q < row_ptr[r0 + rl + 1]; ++q) {
        int32_t& at = pos[(size_t)(col[q] / W) * nr + rl];
        pcol[at] = col[q];
        pv[at] = vals[q];
        ++at;
      }
    }
  }
  tile_off[(size_t)G * P] = (int32_t)total;
  const size_t s = esize(prec);
  std::vector<char> pval_t((size_t)total * s);
  to_prec(prec, pv.data(), total, pval_t.data());
  auto up = [&](void** d, const void* h, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(d, std::max<size_t>(bytes, 4));
    if (e == hipSuccess && bytes) e = hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice);
    return e;
  };
  hipError_t e = up((void**)&M->range_row0, range_row0.data(), range_row0.size() * 4);
  if (e == hipSuccess) e = up((void**)&M->tile_off, tile_off.data(), tile_off.size() * 4);
  if (e == hipSuccess) e = up((void**)&M->starts, starts.data(), starts.size() * 2);
  if (e == hipSuccess) e = up((void**)&M->pcol, pcol.data(), pcol.size() * 4);
  if (e == hipSuccess) e = up(&M->pval, pval_t.data(), pval_t.size());
  if (e == hipSuccess) e = up((void**)&M->long_rows, long_rows.data(), long_rows.size() * 4);
  if (e != hipSuccess) return dev_fail(ctx, e, "panel layout");
  M->panels = P;
  M->nranges = G;
  M->nlong = (int32_t)long_rows.size();
  return PRIMA_OK;
}

int prima_mat_upload_csr(prima_ctx* ctx, int32_t rows, int32_t cols, int64_t nnz,
                         const int32_t* row_ptr, const int32_t* col, const double* vals,
                         prima_mat** out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!out || rows < 0 || cols < 0 || nnz < 0 || nnz > INT32_MAX)
    return fail(ctx, PRIMA_EINVAL, "bad matrix shape");
  if (!row_ptr || (nnz > 0 && (!col || !vals))) return fail(ctx, PRIMA_EINVAL, "null CSR");
  if (row_ptr[0] != 0 || row_ptr[rows] != nnz)
    return fail(ctx, PRIMA_EINVAL, "row_ptr does not span nnz");
  for (int32_t r = 0; r < rows; ++r) {
    if (row_ptr[r + 1] < row_ptr[r]) return fail(ctx, PRIMA_EINVAL, "row_ptr not monotone");
    for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
      if (col[p] < 0 || col[p] >= cols) return fail(ctx, PRIMA_EINVAL, "column out of range");
      if (!std::isfinite(vals[p])) return fail(ctx, PRIMA_EINVAL, "entry is not finite");
    }
  }
  prima_mat* M = new prima_mat();
  M->rows = rows;
  M->cols = cols;
  M->nnz = nnz;
  const size_t s = esize(ctx->precision);
  bool ok = hipMalloc((void**)&M->ptr, ((size_t)rows + 1) * 4) == hipSuccess &&
            hipMalloc((void**)&M->col, std::max<size_t>(1, (size_t)nnz * 4)) == hipSuccess &&
            hipMalloc(&M->val, std::max<size_t>(1, (size_t)nnz * s)) == hipSuccess;
  if (!ok) {
    prima_mat_free(ctx, M);
    return fail(ctx, PRIMA_ENOMEM, "matrix allocation");
  }
  std::vector<char> tmp((size_t)nnz * s);
  to_prec(ctx->precision, vals, nnz, tmp.data());
  hipError_t e = hipMemcpy(M->ptr, row_ptr, ((size_t)rows + 1) * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(M->col, col, (size_t)nnz * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(M->val, tmp.data(), tmp.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    prima_mat_free(ctx, M);
    return dev_fail(ctx, e, "matrix upload");
  }
  {
    const std::vector<int32_t> rb = row_blocks(rows, row_ptr);
    const std::vector<int32_t> d = block_descriptors(rb, row_ptr);
    M->nblocks = (int32_t)rb.size() - 1;
    e = hipMalloc((void**)&M->rbd, d.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(M->rbd, d.data(), d.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      prima_mat_free(ctx, M);
      return dev_fail(ctx, e, "row-block descriptors");
    }
    const std::vector<int32_t> rbs =
        row_blocks(rows, row_ptr, small_block_tile(), small_block_rows());
    const std::vector<int32_t> ds = block_descriptors(rbs, row_ptr);
    M->nblocks_small = (int32_t)rbs.size() - 1;
    e = hipMalloc((void**)&M->rbd_small, ds.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(M->rbd_small, ds.data(), ds.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      prima_mat_free(ctx, M);
      return dev_fail(ctx, e, "row-block descriptors");
    }
  }
  {
    const int prc = build_panels(ctx, M, row_ptr, col, vals);
    if (prc) {
      prima_mat_free(ctx, M);
      return prc;
    }
  }
  *out = M;
  return PRIMA_OK;
}

int prima_mat_free(prima_ctx* ctx, prima_mat* M) {
  if (!M) return PRIMA_OK;
  if (ctx && !ctx->closed) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
  }
  if (M->ptr) (void)hipFree(M->ptr);
  if (M->col) (void)hipFree(M->col);
  if (M->val) (void)hipFree(M->val);
  if (M->rbd) (void)hipFree(M->rbd);
  if (M->rbd_small) (void)hipFree(M->rbd_small);
  for (void* p : {(void*)M->range_row0, (void*)M->tile_off, (void*)M->starts, (void*)M->pcol,
                  M->pval, (void*)M->long_rows})
    if (p) (void)hipFree(p);
  delete M;
  return PRIMA_OK;
}

int prima_spmv(prima_ctx* ctx, const prima_mat* a, const prima_vec* x, prima_vec* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!a || !x || !out || x->n != a->cols || out->n != a->rows)
    return fail(ctx, PRIMA_EINVAL, "spmv shape mismatch");
  if (x->d == out->d) return fail(ctx, PRIMA_EINVAL, "spmv output aliases its input");
  HIP_TRY(ctx, spmv_scheduled(ctx, a, x->d, out->d, a->schedule));
  return PRIMA_OK;
}

int prima_mat_set_schedule(prima_ctx* ctx, prima_mat* a, int32_t schedule) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!a || schedule < 0 || schedule > PRIMA_SPMV_ROW_BLOCKS_SMALL)
    return fail(ctx, PRIMA_EINVAL, "unknown SpMV schedule");
  if (schedule == PRIMA_SPMV_PANELS && a->panels == 0)
    return fail(ctx, PRIMA_EINVAL, "matrix has no column-panel layout (x fits in L2)");
  a->schedule = schedule;
  return PRIMA_OK;
}

int prima_mat_autotune(prima_ctx* ctx, prima_mat* a, int32_t reps, int32_t* chosen,
                       double* best_ms) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!a || reps <= 0) return fail(ctx, PRIMA_EINVAL, "bad autotune arguments");
  const size_t s = esize(ctx->precision);
  // owns both buffers from before the first allocation, so an early return
  // frees whichever exists
  struct Buffers {
    void* x = nullptr;
    void* y = nullptr;
    ~Buffers() {
      if (x) (void)hipFree(x);
      if (y) (void)hipFree(y);
    }
  } buf;
  HIP_TRY(ctx, hipMalloc(&buf.x, std::max<size_t>(1, (size_t)a->cols * s)));
  HIP_TRY(ctx, hipMalloc(&buf.y, std::max<size_t>(1, (size_t)a->rows * s)));
  void* const x = buf.x;
  void* const y = buf.y;
  HIP_TRY(ctx, hipMemsetAsync(x, 0, std::max<size_t>(1, (size_t)a->cols * s), ctx->stream));
  double best = INFINITY;
  int pick = PRIMA_SPMV_ROW_BLOCKS;
  for (int sch = PRIMA_SPMV_VECTOR_1; sch <= PRIMA_SPMV_ROW_BLOCKS_SMALL; ++sch) {
    if (sch == PRIMA_SPMV_PANELS && a->panels == 0) continue;
    HIP_TRY(ctx, spmv_scheduled(ctx, a, x, y, sch));  // warm
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps; ++r) HIP_TRY(ctx, spmv_scheduled(ctx, a, x, y, sch));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    const double per = ms / reps;
    if (per < best) {
      best = per;
      pick = sch;
    }
  }
  a->schedule = pick;
  if (chosen) *chosen = pick;
  if (best_ms) *best_ms = best;
  return PRIMA_OK;
}

int prima_axpby(prima_ctx* ctx, double alpha, const prima_vec* x, double beta, const prima_vec* y,
                prima_vec* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!x || !y || !out || x->n != out->n || y->n != out->n)
    return fail(ctx, PRIMA_EINVAL, "axpby length mismatch");
  HIP_TRY(ctx, op_axpby(ctx->precision, out->n, alpha, x->d, beta, y->d, out->d, ctx->stream));
  return PRIMA_OK;
}

int prima_scale(prima_ctx* ctx, double alpha, const prima_vec* x, prima_vec* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!x || !out || x->n != out->n) return fail(ctx, PRIMA_EINVAL, "scale length mismatch");
  HIP_TRY(ctx, op_scale(ctx->precision, out->n, alpha, x->d, out->d, ctx->stream));
  return PRIMA_OK;
}

int prima_copy(prima_ctx* ctx, const prima_vec* src, prima_vec* dst) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!src || !dst || src->n != dst->n) return fail(ctx, PRIMA_EINVAL, "copy length mismatch");
  if (src->n == 0 || src->d == dst->d) return PRIMA_OK;
  HIP_TRY(ctx, hipMemcpyAsync(dst->d, src->d, (size_t)src->n * esize(ctx->precision),
                              hipMemcpyDeviceToDevice, ctx->stream));
  return PRIMA_OK;
}

int prima_dot(prima_ctx* ctx, const prima_vec* x, const prima_vec* y, double* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!x || !y || !out || x->n != y->n) return fail(ctx, PRIMA_EINVAL, "dot length mismatch");
  if (x->n == 0) {
    *out = 0.0;
    return PRIMA_OK;
  }
  HIP_TRY(ctx, op_dot(ctx->precision, x->n, x->d, y->d, ctx->scratch, ctx->stream));
  HIP_TRY(ctx, hipMemcpyAsync(out, ctx->scratch + 1024, sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return PRIMA_OK;
}

int prima_squared_norm(prima_ctx* ctx, const prima_vec* x, double* out) {
  return prima_dot(ctx, x, x, out);
}

int prima_primal_step(prima_ctx* ctx, const prima_vec* x, const prima_vec* kty,
                      const prima_vec* cost, const prima_vec* lower, const prima_vec* upper,
                      double tau, prima_vec* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!x || !kty || !cost || !lower || !upper || !out) return fail(ctx, PRIMA_EINVAL, "null");
  const int64_t n = out->n;
  if (x->n != n || kty->n != n || cost->n != n || lower->n != n || upper->n != n)
    return fail(ctx, PRIMA_EINVAL, "primalStep length mismatch");
  HIP_TRY(ctx, op_primal(ctx->precision, n, x->d, kty->d, cost->d, lower->d, upper->d, tau,
                         out->d, ctx->stream));
  return PRIMA_OK;
}

int prima_dual_step(prima_ctx* ctx, const prima_vec* y, const prima_vec* kx_bar,
                    const prima_vec* rhs, double sigma, int32_t n_eq, prima_vec* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!y || !kx_bar || !rhs || !out) return fail(ctx, PRIMA_EINVAL, "null");
  const int64_t m = out->n;
  if (y->n != m || kx_bar->n != m || rhs->n != m || n_eq < 0 || n_eq > m)
    return fail(ctx, PRIMA_EINVAL, "dualStep shape mismatch");
  HIP_TRY(ctx, op_dual(ctx->precision, m, y->d, kx_bar->d, rhs->d, sigma, n_eq, out->d,
                       ctx->stream));
  return PRIMA_OK;
}

int prima_solve_batch(prima_ctx* ctx, int32_t count, const prima_lp* lps,
                      const prima_params* params, const prima_mixed_params* mixed,
                      const prima_warm* warms, prima_result* results, prima_mixed_info* infos,
                      double** x_outs, double** y_outs, double** rc_outs) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (count < 0 || (count > 0 && (!lps || !results)) || !params)
    return fail(ctx, PRIMA_EINVAL, "bad batch arguments");
  std::string e = validate_params(params);
  if (!e.empty()) return fail(ctx, PRIMA_EINVAL, e);
  if (!ctx->coop_ok)
    return fail(ctx, PRIMA_EDEVICE, "device lacks cooperative launch support for the solver");
  ctx->cancel.store(0);
  if (count == 0) return PRIMA_OK;
  if (mixed) {
    e = validate_mixed(mixed);
    if (!e.empty()) return fail(ctx, PRIMA_EINVAL, e);
    std::vector<prima_mixed_info> tmp(count);
    if (ctx->precision == 64) {
      // an fp64 device skips the reduced pass (MixedPrecision.scala:111-113)
      rc = prima_solve_batch(ctx, count, lps, params, nullptr, warms, results, nullptr, x_outs,
                             y_outs, rc_outs);
      if (rc) return rc;
      for (int32_t k = 0; k < count; ++k) {
        std::memset(&tmp[k], 0, sizeof(prima_mixed_info));
        tmp[k].refinement_iterations = results[k].iterations;
        tmp[k].refinement_ms = results[k].solve_time_ms;
      }
    } else if (mixed->refinement == 1) {
      rc = refine_batch(ctx, count, lps, params, mixed, warms, results, tmp.data(), x_outs, y_outs,
                        rc_outs);
      if (rc) return rc;
    } else {
      rc = mixed_batch(ctx, count, lps, params, mixed, warms, results, tmp.data(), x_outs, y_outs,
                       rc_outs);
      if (rc) return rc;
    }
    if (infos) std::memcpy(infos, tmp.data(), count * sizeof(prima_mixed_info));
    return PRIMA_OK;
  }
  std::vector<SolveJob> jobs(count);
  {
    std::vector<const prima_lp*> lp_of(count);
    std::vector<const prima_warm*> warm_of(count);
    for (int32_t k = 0; k < count; ++k) {
      lp_of[k] = &lps[k];
      warm_of[k] = warms ? &warms[k] : nullptr;
    }
    rc = prepare_jobs(ctx, jobs, lp_of, params, warm_of);
    if (rc) return rc;
  }
  for (int32_t k = 0; k < count; ++k) {
    jobs[k].res = &results[k];
    jobs[k].x_out = x_outs ? x_outs[k] : nullptr;
    jobs[k].y_out = y_outs ? y_outs[k] : nullptr;
    jobs[k].rc_out = rc_outs ? rc_outs[k] : nullptr;
  }
  return run_jobs(ctx, ctx->precision, jobs, params);
}

int prima_solve(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                const prima_warm* warm, prima_result* res, double* x_out, double* y_out,
                double* rc_out) {
  if (!lp || !res) return ctx ? fail(ctx, PRIMA_EINVAL, "null argument") : PRIMA_EINVAL;
  return prima_solve_batch(ctx, 1, lp, params, nullptr, warm, res, nullptr, &x_out, &y_out,
                           &rc_out);
}

int prima_tune_solver(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                      int32_t iterations, prima_tune_result* out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!lp || !params || !out || iterations <= 0)
    return fail(ctx, PRIMA_EINVAL, "bad tuning arguments");
  std::string e = validate_lp(lp);
  if (e.empty()) e = validate_params(params);
  if (!e.empty()) return fail(ctx, PRIMA_EINVAL, e);
  std::memset(out, 0, sizeof(*out));
  const uint64_t key = matrix_key(lp, params);
  // forget an earlier choice for this matrix while the candidates run
  ctx->tuned.erase(std::remove_if(ctx->tuned.begin(), ctx->tuned.end(),
                                  [&](const prima_ctx::Tuned& t) { return t.key == key; }),
                   ctx->tuned.end());
  const int occ = ctx->occ[ctx->precision == 32 ? 0 : 1];
  const int gmax = std::max(1, std::min(ctx->cu_count * std::max(occ, 1), kMaxGroup));
  const int64_t nnz = std::max<int64_t>(lp->nnz, 1);
  const int rule = (int)std::max<int64_t>(1, std::min<int64_t>((nnz + 1023) / 1024, gmax));
  std::vector<int> gs;
  for (int g : {std::max(1, rule / 2), rule, std::min(gmax, 2 * rule)})
    if (std::find(gs.begin(), gs.end(), g) == gs.end()) gs.push_back(g);
  // every candidate runs exactly `iterations` accepted iterations: no
  // tolerance is ever met, certificates are not tested
  prima_params tp = *params;
  tp.eps_abs = 0.0;
  tp.eps_rel = 0.0;
  tp.infeasibility_tolerance = 0.0;
  tp.max_iterations = iterations;
  tp.time_limit_ms = -1;
  int best = -1;
  double best_us = INFINITY;
  for (int g : gs) {
    for (int bpw : {0, 1, 2}) {
      if (g == 1 && bpw != 0) continue;  // one workgroup always takes the coarse cut
      if (out->candidates >= PRIMA_TUNE_MAX_CANDIDATES) break;
      ctx->tuned.push_back(prima_ctx::Tuned{key, g, bpw});
      tp.workgroups_per_lp = g;
      const prima_stats before = ctx->stats;
      prima_result r{};
      rc = prima_solve(ctx, lp, &tp, nullptr, &r, nullptr, nullptr, nullptr);
      ctx->tuned.pop_back();
      if (rc) return rc;
      const double ms = ctx->stats.kernel_ms - before.kernel_ms;
      const double us = r.iterations > 0 ? 1e3 * ms / r.iterations : INFINITY;
      const int c = out->candidates++;
      out->cand_workgroups[c] = g;
      out->cand_blocks_per_workgroup[c] = bpw;
      out->cand_us_per_iteration[c] = us;
      if (us < best_us) {
        best_us = us;
        best = c;
      }
    }
  }
  if (best < 0) return fail(ctx, PRIMA_EDEVICE, "no tuning candidate ran");
  out->workgroups = out->cand_workgroups[best];
  out->blocks_per_workgroup = out->cand_blocks_per_workgroup[best];
  out->us_per_iteration = best_us;
  ctx->tuned.push_back(prima_ctx::Tuned{key, out->workgroups, out->blocks_per_workgroup});
  return PRIMA_OK;
}

int prima_solve_noalloc(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                        const prima_warm* warm, prima_result* res, double* x_out, double* y_out,
                        double* rc_out) {
  int rc = check_open(ctx);
  if (rc) return rc;
  if (!lp || !res || !params) return fail(ctx, PRIMA_EINVAL, "null argument");
  std::string e = validate_params(params);
  if (!e.empty()) return fail(ctx, PRIMA_EINVAL, e);
  if (!ctx->coop_ok)
    return fail(ctx, PRIMA_EDEVICE, "device lacks cooperative launch support for the solver");
  size_t bytes = 0;
  void* base = scratch_of(ctx->device, &bytes);
  if (!ctx->scratch_ok || !base)
    return fail(ctx, PRIMA_EDEVICE, "the allocation-free workspace of this device is unknown");
  ctx->cancel.store(0);
  std::vector<SolveJob> jobs(1);
  rc = prepare_jobs(ctx, jobs, {lp}, params, {warm});
  if (rc) return rc;
  jobs[0].res = res;
  jobs[0].x_out = x_out;
  jobs[0].y_out = y_out;
  jobs[0].rc_out = rc_out;
  const Workspace ws{(char*)base, bytes};
  return run_jobs(ctx, ctx->precision, jobs, params, true, &ws);
}

int prima_mixed_solve(prima_ctx* ctx, const prima_lp* lp, const prima_params* params,
                      const prima_mixed_params* mixed, prima_result* res, prima_mixed_info* info,
                      double* x_out, double* y_out, double* rc_out) {
  if (!lp || !res || !mixed) return ctx ? fail(ctx, PRIMA_EINVAL, "null argument") : PRIMA_EINVAL;
  return prima_solve_batch(ctx, 1, lp, params, mixed, nullptr, res, info, &x_out, &y_out,
                           &rc_out);
}

}  // extern "C"
