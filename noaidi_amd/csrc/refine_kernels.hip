// FP64 side of the outer iterative refinement (Gleixner, Steffy & Wolter
// 2016) around float32 correction solves: the residuals b - A x and
// c - A^T y of the original LP in double precision, its KKT error and
// objectives (Kkt.scala:68-125) with the termination test of
// Pdhg.scala:396-417, the scaled correction problem built from them, and
// the correction step x += dx / dp, y += dy / dd with the box and dual cone
// re-imposed. Everything is float64 on the device; the host reads back one
// small result record per LP.
//
// Reductions are fixed-order trees over a fixed grid, so the results are
// identical run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace prima {

constexpr int REF_NT = 256;
constexpr int REF_GRID = 512;  // workgroups of the residual passes (partials per LP)

// result record of one evaluation, in doubles
enum {
  RF_PRES = 0, RF_DRES, RF_POBJ, RF_DOBJ, RF_PV, RF_DV, RF_CONVERGED, RF_COUNT
};

// Per-block partials of the residual passes.
enum { RP_PSQ = 0, RP_QY, RP_PV, RP_DSQ, RP_DLIN, RP_CX, RP_DV, RP_COUNT };

__device__ inline double ref_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double ref_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// K sums then one max over the workgroup, written by thread 0 to out.
template <int K>
__device__ inline void ref_block_reduce(double (&s)[K], double mx, double* out) {
  __shared__ double sh[REF_NT / 64][K + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) s[k] = ref_wave_sum(s[k]);
  mx = ref_wave_max(mx);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[w][k] = s[k];
    sh[w][K] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[K + 1];
#pragma unroll
    for (int k = 0; k <= K; ++k) t[k] = 0.0;
    for (int j = 0; j < REF_NT / 64; ++j) {
#pragma unroll
      for (int k = 0; k < K; ++k) t[k] += sh[j][k];
      t[K] = fmax(t[K], sh[j][K]);
    }
#pragma unroll
    for (int k = 0; k <= K; ++k) out[k] = t[k];
  }
}

// Rows of A: ax = A x; rq = q - ax; partials of |primal residual|^2
// (equalities exact, inequalities one-sided: Kkt.scala:80-95), q'y and the
// largest primal violation. TPR lanes per row.
template <int TPR>
__global__ __launch_bounds__(REF_NT) void k_ref_rows(int m, int n_eq,
                                                    const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ col,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ x,
                                                    const double* __restrict__ q,
                                                    const double* __restrict__ y,
                                                    double* __restrict__ rq,
                                                    double* __restrict__ part) {
  const int per_block = REF_NT / TPR;
  const int li = threadIdx.x % TPR;
  double s[2] = {0.0, 0.0};  // psq, qy
  double pv = 0.0;
  for (int r = blockIdx.x * per_block + threadIdx.x / TPR; r < m; r += gridDim.x * per_block) {
    double acc = 0.0;
    for (int p = ptr[r] + li; p < ptr[r + 1]; p += TPR) acc += val[p] * x[col[p]];
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (li == 0) {
      const double res = q[r] - acc;
      rq[r] = res;
      if (r < n_eq) {
        s[0] += res * res;
        pv = fmax(pv, fabs(res));
      } else if (res > 0.0) {
        s[0] += res * res;
        pv = fmax(pv, res);
      }
      s[1] += q[r] * y[r];
    }
  }
  ref_block_reduce<2>(s, pv, part + (size_t)blockIdx.x * RP_COUNT + RP_PSQ);
}

// Rows of A^T: aty = A^T y; rc = c - aty; partials of |dual residual|^2 and
// the bound terms of the dual objective (Kkt.scala:97-120), c'x and the
// largest reduced-cost violation of complementarity at the current x.
template <int TPR>
__global__ __launch_bounds__(REF_NT) void k_ref_cols(int n, const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ col,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ c,
                                                    const double* __restrict__ l,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ rc_out,
                                                    double* __restrict__ part) {
  const int per_block = REF_NT / TPR;
  const int li = threadIdx.x % TPR;
  double s[3] = {0.0, 0.0, 0.0};  // dsq, dlin, cx
  double dv = 0.0;
  for (int i = blockIdx.x * per_block + threadIdx.x / TPR; i < n; i += gridDim.x * per_block) {
    double acc = 0.0;
    for (int p = ptr[i] + li; p < ptr[i + 1]; p += TPR) acc += val[p] * y[col[p]];
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (li == 0) {
      const double rc = c[i] - acc;
      const double lo = l[i], hi = u[i], xi = x[i];
      rc_out[i] = rc;
      if (rc > 0.0) {
        if (isinf(lo) && lo < 0.0) s[0] += rc * rc;
        else s[1] += lo * rc;
      } else if (rc < 0.0) {
        if (isinf(hi) && hi > 0.0) s[0] += rc * rc;
        else s[1] += hi * rc;
      }
      s[2] += c[i] * xi;
      const bool atl = isfinite(lo) && xi - lo <= 1e-12 * (1.0 + fabs(lo));
      const bool atu = isfinite(hi) && hi - xi <= 1e-12 * (1.0 + fabs(hi));
      double v;
      if (atl && atu) v = 0.0;
      else if (atl) v = fmax(-rc, 0.0);
      else if (atu) v = fmax(rc, 0.0);
      else v = fabs(rc);
      dv = fmax(dv, v);
    }
  }
  ref_block_reduce<3>(s, dv, part + (size_t)blockIdx.x * RP_COUNT + RP_DSQ);
}

struct RefTol {
  double eps_abs, eps_rel, rhs_norm, obj_norm, offset;
};

// One workgroup: the partials of both passes in block order, then the KKT
// error, objectives and the termination test (Pdhg.scala:406-412).
__global__ __launch_bounds__(REF_NT) void k_ref_finish(int blocks, const double* __restrict__ part,
                                                      RefTol t, double* __restrict__ out) {
  __shared__ double sh[REF_NT / 64][RP_COUNT];
  double s[RP_COUNT];
#pragma unroll
  for (int k = 0; k < RP_COUNT; ++k) s[k] = 0.0;
  for (int b = threadIdx.x; b < blocks; b += REF_NT) {
#pragma unroll
    for (int k = 0; k < RP_COUNT; ++k) {
      const double v = part[(size_t)b * RP_COUNT + k];
      s[k] = (k == RP_PV || k == RP_DV) ? fmax(s[k], v) : s[k] + v;
    }
  }
#pragma unroll
  for (int k = 0; k < RP_COUNT; ++k)
    s[k] = (k == RP_PV || k == RP_DV) ? ref_wave_max(s[k]) : ref_wave_sum(s[k]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < RP_COUNT; ++k) sh[w][k] = s[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double a[RP_COUNT];
#pragma unroll
    for (int k = 0; k < RP_COUNT; ++k) a[k] = 0.0;
    for (int j = 0; j < REF_NT / 64; ++j)
#pragma unroll
      for (int k = 0; k < RP_COUNT; ++k)
        a[k] = (k == RP_PV || k == RP_DV) ? fmax(a[k], sh[j][k]) : a[k] + sh[j][k];
    const double pres = sqrt(a[RP_PSQ]), dres = sqrt(a[RP_DSQ]);
    const double pobj = t.offset + a[RP_CX];
    const double dobj = t.offset + a[RP_QY] + a[RP_DLIN];
    const bool finite = isfinite(pres) && isfinite(dres) && isfinite(pobj) && isfinite(dobj);
    const bool ok = finite && pres <= t.eps_abs + t.eps_rel * t.rhs_norm &&
                    dres <= t.eps_abs + t.eps_rel * t.obj_norm &&
                    fabs(pobj - dobj) <= t.eps_abs + t.eps_rel * (fabs(pobj) + fabs(dobj));
    out[RF_PRES] = pres;
    out[RF_DRES] = dres;
    out[RF_POBJ] = pobj;
    out[RF_DOBJ] = dobj;
    out[RF_PV] = a[RP_PV];
    out[RF_DV] = a[RP_DV];
    out[RF_CONVERGED] = ok ? 1.0 : 0.0;
  }
}

// The scaled correction problem at (x, y): costs dd (c - A^T y) clipped to
// +-cap, bounds dp (l - x) <= dx <= dp (u - x) opened to contain 0, rows
// dp (q - A x), and the shifted dual cone y_r >= -dd y_r of inequality rows.
// With fix > 0, a variable at a bound whose scaled reduced cost holds that
// bound by more than `fix` (dd rc > fix at the lower bound, < -fix at the
// upper) is fixed in the correction (both bounds 0, cost 0): its cost would
// otherwise be of order dd |rc| >> 1 and dominate the correction problem's
// norms, so that a relative tolerance on it would say nothing about the
// violations being corrected.
__global__ void k_ref_build(int n, int m, int n_eq, double dp, double dd, double cap, double fix,
                            const double* __restrict__ rc, const double* __restrict__ rq,
                            const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ l, const double* __restrict__ u,
                            double* __restrict__ cc, double* __restrict__ cl,
                            double* __restrict__ cu, double* __restrict__ cq,
                            double* __restrict__ ylo) {
  const int64_t total = (int64_t)n + m;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k < n) {
      const int i = (int)k;
      const double li = l[i], ui = u[i], xi = x[i];
      const double cv = dd * rc[i];
      const double lo = dp * (li - xi), hi = dp * (ui - xi);
      const bool atl = isfinite(li) && xi - li <= 1e-12 * (1.0 + fabs(li));
      const bool atu = isfinite(ui) && ui - xi <= 1e-12 * (1.0 + fabs(ui));
      if (fix > 0.0 && ((atl && cv > fix) || (atu && cv < -fix))) {
        cc[i] = 0.0;
        cl[i] = 0.0;
        cu[i] = 0.0;
        continue;
      }
      cc[i] = cv < -cap ? -cap : (cv > cap ? cap : cv);
      cl[i] = lo > 0.0 ? 0.0 : lo;
      cu[i] = hi < 0.0 ? 0.0 : hi;
    } else {
      const int r = (int)(k - n);
      cq[r] = dp * rq[r];
      ylo[r] = r >= n_eq ? -dd * y[r] : 0.0;
    }
  }
}

// x <- clamp(x + dx / dp, l, u); y <- y + dy / dd, inequality rows kept >= 0.
__global__ void k_ref_apply(int n, int m, int n_eq, double dp, double dd,
                            const double* __restrict__ cx, const double* __restrict__ cy,
                            const double* __restrict__ l, const double* __restrict__ u,
                            double* __restrict__ x, double* __restrict__ y) {
  const int64_t total = (int64_t)n + m;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    if (k < n) {
      const int i = (int)k;
      const double v = x[i] + cx[i] / dp;
      x[i] = v < l[i] ? l[i] : (v > u[i] ? u[i] : v);
    } else {
      const int r = (int)(k - n);
      const double v = y[r] + cy[r] / dd;
      y[r] = (r >= n_eq && v < 0.0) ? 0.0 : v;
    }
  }
}

// ---------------------------------------------------------------------------
// launchers

static int ref_grid_elem(int64_t total) {
  int64_t g = (total + REF_NT - 1) / REF_NT;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

int ref_partial_blocks() { return REF_GRID; }
int ref_result_doubles() { return RF_COUNT; }

// Evaluates the original LP at (x, y): writes rq (m), rc (n), the partials
// (REF_GRID * RP_COUNT doubles of scratch: block b's row pass fills slots
// RP_PSQ..RP_PV of record b and its column pass RP_DSQ..RP_DV, so every slot
// is written and none needs clearing) and the result record out.
hipError_t ref_evaluate(int n, int m, int n_eq, int64_t nnz, const int32_t* aptr,
                        const int32_t* acol, const double* aval, const int32_t* tptr,
                        const int32_t* tcol, const double* tval, const double* x,
                        const double* y, const double* c, const double* q, const double* l,
                        const double* u, double eps_abs, double eps_rel, double rhs_norm,
                        double obj_norm, double offset, double* rq, double* rc, double* part,
                        double* out, hipStream_t s) {
  const double mean_r = m > 0 ? (double)nnz / m : 0.0;
  const double mean_c = n > 0 ? (double)nnz / n : 0.0;
  if (mean_r <= 4.0)
    k_ref_rows<2><<<REF_GRID, REF_NT, 0, s>>>(m, n_eq, aptr, acol, aval, x, q, y, rq, part);
  else if (mean_r <= 16.0)
    k_ref_rows<8><<<REF_GRID, REF_NT, 0, s>>>(m, n_eq, aptr, acol, aval, x, q, y, rq, part);
  else
    k_ref_rows<32><<<REF_GRID, REF_NT, 0, s>>>(m, n_eq, aptr, acol, aval, x, q, y, rq, part);
  if (mean_c <= 4.0)
    k_ref_cols<2><<<REF_GRID, REF_NT, 0, s>>>(n, tptr, tcol, tval, y, c, l, u, x, rc, part);
  else if (mean_c <= 16.0)
    k_ref_cols<8><<<REF_GRID, REF_NT, 0, s>>>(n, tptr, tcol, tval, y, c, l, u, x, rc, part);
  else
    k_ref_cols<32><<<REF_GRID, REF_NT, 0, s>>>(n, tptr, tcol, tval, y, c, l, u, x, rc, part);
  RefTol t{eps_abs, eps_rel, rhs_norm, obj_norm, offset};
  k_ref_finish<<<1, REF_NT, 0, s>>>(REF_GRID, part, t, out);
  return hipGetLastError();
}

hipError_t ref_build(int n, int m, int n_eq, double dp, double dd, double cap, double fix,
                     const double* rc, const double* rq, const double* x, const double* y,
                     const double* l, const double* u, double* cc, double* cl, double* cu,
                     double* cq, double* ylo, hipStream_t s) {
  k_ref_build<<<ref_grid_elem((int64_t)n + m), REF_NT, 0, s>>>(n, m, n_eq, dp, dd, cap, fix, rc,
                                                              rq, x, y, l, u, cc, cl, cu, cq, ylo);
  return hipGetLastError();
}

hipError_t ref_apply(int n, int m, int n_eq, double dp, double dd, const double* cx,
                     const double* cy, const double* l, const double* u, double* x, double* y,
                     hipStream_t s) {
  k_ref_apply<<<ref_grid_elem((int64_t)n + m), REF_NT, 0, s>>>(n, m, n_eq, dp, dd, cx, cy, l, u,
                                                              x, y);
  return hipGetLastError();
}

}  // namespace prima
