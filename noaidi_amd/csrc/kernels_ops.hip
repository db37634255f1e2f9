// The eight numeric operations of the reference's kernel seam
// (kernels/Kernels.scala:70-123), one HIP kernel each, in the context's
// iteration precision T (double or float). These back the per-op C-ABI
// entry points; the solver itself runs the fused persistent kernel instead.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "spmv_blocks.h"

namespace prima {

constexpr int OPS_NT = 256;

static inline int grid_for(int64_t n) {
  int64_t g = (n + OPS_NT - 1) / OPS_NT;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

// CSR SpMV (SparseMatrix.scala:76-86): a power-of-two sub-wave per row,
// sub-wave width chosen from the mean row length; fixed reduction order.
template <typename T, int TPR>
__global__ __launch_bounds__(OPS_NT) void k_spmv(int rows, const int32_t* __restrict__ ptr,
                                                 const int32_t* __restrict__ col,
                                                 const T* __restrict__ val,
                                                 const T* __restrict__ x, T* __restrict__ out) {
  const int per_block = OPS_NT / TPR;
  const int li = threadIdx.x % TPR;
  for (int r = blockIdx.x * per_block + threadIdx.x / TPR; r < rows;
       r += gridDim.x * per_block) {
    T acc = T(0);
    // val and col are streamed once: non-temporal loads leave the cache to x
    for (int p = ptr[r] + li; p < ptr[r + 1]; p += TPR)
      acc += __builtin_nontemporal_load(&val[p]) * x[__builtin_nontemporal_load(&col[p])];
#pragma unroll
    for (int off = TPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (li == 0) out[r] = acc;
  }
}

template <typename T>
__global__ void k_axpby(int64_t n, T a, const T* x, T b, const T* y, T* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a * x[i] + b * y[i];
}

template <typename T>
__global__ void k_scale(int64_t n, T a, const T* x, T* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a * x[i];
}

// Primal step (ScalaKernels.scala:78-98): two comparisons, so infinite
// bounds pass the candidate through.
template <typename T>
__global__ void k_primal(int64_t n, const T* x, const T* kty, const T* c, const T* l, const T* u,
                         T tau, T* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T s = x[i] - tau * (c[i] - kty[i]);
    out[i] = s < l[i] ? l[i] : (s > u[i] ? u[i] : s);
  }
}

// Dual step (ScalaKernels.scala:100-118): rows below n_eq are free.
template <typename T>
__global__ void k_dual(int64_t m, const T* y, const T* kxb, const T* q, T sigma, int64_t n_eq,
                       T* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T s = y[i] + sigma * (q[i] - kxb[i]);
    out[i] = (i < n_eq || s > T(0)) ? s : T(0);
  }
}

// Dot product, stage 1: per-thread accumulation in T, fixed-order tree in
// double per block. Stage 2 sums the block partials in one block.
template <typename T>
__global__ __launch_bounds__(OPS_NT) void k_dot_partial(int64_t n, const T* x, const T* y,
                                                        double* part) {
  __shared__ double sh[OPS_NT / 64];
  T acc = T(0);
  for (int64_t i = blockIdx.x * (int64_t)OPS_NT + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * OPS_NT)
    acc += x[i] * y[i];
  double s = (double)acc;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < OPS_NT / 64; ++w) t += sh[w];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(OPS_NT) void k_sum_partials(int count, const double* part,
                                                         double* out) {
  __shared__ double sh[OPS_NT / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += OPS_NT) s += part[i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < OPS_NT / 64; ++w) t += sh[w];
    *out = t;
  }
}

// ---------------------------------------------------------------------------
// launchers (called by the C-ABI layer)

template <typename T>
static hipError_t spmv_t(int rows, int64_t nnz, const int32_t* ptr, const int32_t* col,
                         const void* val, const void* x, void* out, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const double mean = (double)nnz / rows;
  const T* v = (const T*)val;
  const T* xv = (const T*)x;
  T* o = (T*)out;
  if (mean <= 2.0) {
    k_spmv<T, 1><<<grid_for(rows), OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o);
  } else if (mean <= 8.0) {
    k_spmv<T, 4><<<grid_for((int64_t)rows * 4), OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o);
  } else if (mean <= 32.0) {
    k_spmv<T, 16><<<grid_for((int64_t)rows * 16), OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o);
  } else {
    k_spmv<T, 64><<<grid_for((int64_t)rows * 64), OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o);
  }
  return hipGetLastError();
}

hipError_t op_spmv(int prec, int rows, int64_t nnz, const int32_t* ptr, const int32_t* col,
                   const void* val, const void* x, void* out, hipStream_t s) {
  return prec == 32 ? spmv_t<float>(rows, nnz, ptr, col, val, x, out, s)
                    : spmv_t<double>(rows, nnz, ptr, col, val, x, out, s);
}

// Sub-wave-per-row schedule with an explicit width (the autotuner's "vector"
// candidates): tpr in {1, 2, 4, ..., 64}.
template <typename T>
static hipError_t spmv_vec_t(int tpr, int rows, const int32_t* ptr, const int32_t* col,
                             const void* val, const void* x, void* out, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const T* v = (const T*)val;
  const T* xv = (const T*)x;
  T* o = (T*)out;
  const int g = grid_for((int64_t)rows * tpr);
  switch (tpr) {
    case 1: k_spmv<T, 1><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    case 2: k_spmv<T, 2><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    case 4: k_spmv<T, 4><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    case 8: k_spmv<T, 8><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    case 16: k_spmv<T, 16><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    case 32: k_spmv<T, 32><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
    default: k_spmv<T, 64><<<g, OPS_NT, 0, s>>>(rows, ptr, col, v, xv, o); break;
  }
  return hipGetLastError();
}

hipError_t op_spmv_vector(int prec, int tpr, int rows, const int32_t* ptr, const int32_t* col,
                          const void* val, const void* x, void* out, hipStream_t s) {
  return prec == 32 ? spmv_vec_t<float>(tpr, rows, ptr, col, val, x, out, s)
                    : spmv_vec_t<double>(tpr, rows, ptr, col, val, x, out, s);
}

// Row-block (CSR-Adaptive) schedule: the solver's in-kernel SpMV as a plain
// launch, each workgroup owning `bpw` consecutive row blocks.
template <typename T>
struct EpiStore {
  T* out;
  struct Ops {};
  __device__ Ops load(int) const { return Ops{}; }
  __device__ void apply(int r, T s, const Ops&) { out[r] = s; }
};

template <typename T>
__global__ __launch_bounds__(NT) void k_spmv_blocks(const int32_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ col,
                                                    const T* __restrict__ val,
                                                    const T* __restrict__ x,
                                                    const int32_t* __restrict__ rbd, int nblocks,
                                                    int bpw, T* __restrict__ out) {
  __shared__ T tile[TILE];
  __shared__ T rsum[ROWCAP];
  __shared__ int rptr[ROWCAP + 1];
  const int b0 = blockIdx.x * bpw;
  const int b1 = b0 + bpw < nblocks ? b0 + bpw : nblocks;
  EpiStore<T> epi{out};
  spmv_blocks(ptr, col, val, x, rbd, b0, b1, reinterpret_cast<T*>(tile),
              reinterpret_cast<T*>(rsum), rptr, epi);
}

// The same schedule over small row blocks (SMALL_TILE nonzeros, SMALL_ROWS
// rows): 28 KB of LDS in float64 instead of 112 KB, so three workgroups fit a
// CU (the kernel's registers allow three) where the solver's block size fits
// one; for matrices of short rows with local gathers.
constexpr int SMALL_TILE = 2048;
constexpr int SMALL_ROWS = 1024;

template <typename T>
__global__ __launch_bounds__(NT) void k_spmv_blocks_small(const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val,
                                                          const T* __restrict__ x,
                                                          const int32_t* __restrict__ rbd,
                                                          int nblocks, T* __restrict__ out) {
  __shared__ T tile[SMALL_TILE];
  __shared__ T rsum[SMALL_ROWS];
  __shared__ int rptr[SMALL_ROWS + 1];
  const int b0 = blockIdx.x;
  const int b1 = b0 + 1 < nblocks ? b0 + 1 : nblocks;
  EpiStore<T> epi{out};
  spmv_blocks<T, EpiStore<T>, SMALL_TILE>(ptr, col, val, x, rbd, b0, b1, tile, rsum, rptr, epi);
}

int small_block_tile() { return SMALL_TILE; }
int small_block_rows() { return SMALL_ROWS; }

hipError_t op_spmv_blocks_small(int prec, int nblocks, const int32_t* ptr, const int32_t* col,
                                const void* val, const int32_t* rbd, const void* x, void* out,
                                hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  if (prec == 32)
    k_spmv_blocks_small<float><<<nblocks, NT, 0, s>>>(ptr, col, (const float*)val,
                                                      (const float*)x, rbd, nblocks, (float*)out);
  else
    k_spmv_blocks_small<double><<<nblocks, NT, 0, s>>>(ptr, col, (const double*)val,
                                                       (const double*)x, rbd, nblocks,
                                                       (double*)out);
  return hipGetLastError();
}

hipError_t op_spmv_blocks(int prec, int nblocks, int bpw, const int32_t* ptr, const int32_t* col,
                          const void* val, const int32_t* rbd, const void* x, void* out,
                          hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  const int g = (nblocks + bpw - 1) / bpw;
  if (prec == 32)
    k_spmv_blocks<float><<<g, NT, 0, s>>>(ptr, col, (const float*)val, (const float*)x, rbd,
                                          nblocks, bpw, (float*)out);
  else
    k_spmv_blocks<double><<<g, NT, 0, s>>>(ptr, col, (const double*)val, (const double*)x, rbd,
                                           nblocks, bpw, (double*)out);
  return hipGetLastError();
}

// Column-panel schedule for matrices whose x does not fit an XCD's L2. The
// columns are cut into panels of W columns (W·sizeof(T) = 512 KiB; a sweep of
// 512 KiB to 2 MiB panels and 512 to 2048-row ranges on a 1M x 2M matrix
// with 20M nonzeros found this the best overall) and the rows into ranges of
// at most PANEL_ROWS rows; the entries of range g that fall in
// panel p form tile (g, p), stored contiguously, with uint16 row starts
// relative to the tile. One workgroup per range sweeps the panels in order,
// so the workgroups resident at a time gather from one or two panels of x,
// which stay in L2, while the tiles stream from HBM. A row's sum accumulates
// panel by panel in registers, each panel's entries in stored order: the
// result is deterministic. Rows longer than PANEL_LONG are left out of the
// tiles (their sums stay 0 here) and summed afterwards by k_spmv_rows.
#ifndef PRIMA_PANEL_ROWS
#define PRIMA_PANEL_ROWS 512
#endif
#ifndef PRIMA_PANEL_TILE
#define PRIMA_PANEL_TILE 2048
#endif
constexpr int PANEL_NT = 512;
constexpr int PANEL_ROWS = PRIMA_PANEL_ROWS;
constexpr int PANEL_RPT = PANEL_ROWS / PANEL_NT;
constexpr int PANEL_TILE = PRIMA_PANEL_TILE;

constexpr int PANEL_EPT = PANEL_TILE / PANEL_NT;  // tile entries per thread

// One tile's operands for this thread: column indices and values of entries
// threadIdx.x + j*PANEL_NT, and the row starts of this thread's rows.
template <typename T>
struct PanelTile {
  int n;
  int c[PANEL_EPT];
  T v[PANEL_EPT];
  int s[PANEL_RPT], e[PANEL_RPT];
};

template <typename T>
__device__ inline void panel_fetch(PanelTile<T>& f, int t, int nrows,
                                   const int32_t* __restrict__ tile_off,
                                   const uint16_t* __restrict__ starts,
                                   const int32_t* __restrict__ pcol, const T* __restrict__ pval) {
  const int base = tile_off[t];
  f.n = tile_off[t + 1] - base;
#pragma unroll
  for (int j = 0; j < PANEL_EPT; ++j) {
    const int k = (int)threadIdx.x + j * PANEL_NT;
    if (k < f.n) {
      f.c[j] = __builtin_nontemporal_load(&pcol[base + k]);
      f.v[j] = __builtin_nontemporal_load(&pval[base + k]);
    }
  }
  const uint16_t* st = starts + (size_t)t * (PANEL_ROWS + 1);
#pragma unroll
  for (int j = 0; j < PANEL_RPT; ++j) {
    const int r = (int)threadIdx.x + j * PANEL_NT;
    f.s[j] = r < nrows ? (int)st[r] : 0;
    f.e[j] = r < nrows ? (int)st[r + 1] : 0;
  }
}

template <typename T>
__global__ __launch_bounds__(PANEL_NT) void k_spmv_panels(
    int P, const int32_t* __restrict__ range_row0, const int32_t* __restrict__ tile_off,
    const uint16_t* __restrict__ starts, const int32_t* __restrict__ pcol,
    const T* __restrict__ pval, const T* __restrict__ x, T* __restrict__ out) {
  __shared__ T prod[PANEL_TILE];
  const int g = blockIdx.x;
  const int row0 = range_row0[g], nrows = range_row0[g + 1] - row0;
  T acc[PANEL_RPT];
#pragma unroll
  for (int j = 0; j < PANEL_RPT; ++j) acc[j] = T(0);
  // software pipeline over the panels: the next tile's indices, values and
  // row starts are in flight while this tile's gathers and sums run
  PanelTile<T> cur, nxt;
  panel_fetch(cur, g * P, nrows, tile_off, starts, pcol, pval);
  for (int p = 0; p < P; ++p) {
    T pr[PANEL_EPT];
#pragma unroll
    for (int j = 0; j < PANEL_EPT; ++j)
      if ((int)threadIdx.x + j * PANEL_NT < cur.n) pr[j] = cur.v[j] * x[cur.c[j]];
    if (p + 1 < P) panel_fetch(nxt, g * P + p + 1, nrows, tile_off, starts, pcol, pval);
#pragma unroll
    for (int j = 0; j < PANEL_EPT; ++j) {
      const int k = (int)threadIdx.x + j * PANEL_NT;
      if (k < cur.n) prod[k] = pr[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PANEL_RPT; ++j)
      for (int q = cur.s[j]; q < cur.e[j]; ++q) acc[j] += prod[q];
    __syncthreads();
    cur = nxt;
  }
#pragma unroll
  for (int j = 0; j < PANEL_RPT; ++j) {
    const int r = (int)threadIdx.x + j * PANEL_NT;
    if (r < nrows) out[row0 + r] = acc[j];
  }
}

// One wave per listed row (the rows the panel tiles leave out): lanes stride
// the row, fixed-order shuffle tree.
template <typename T>
__global__ __launch_bounds__(OPS_NT) void k_spmv_rows(int count, const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ ptr,
                                                      const int32_t* __restrict__ col,
                                                      const T* __restrict__ val,
                                                      const T* __restrict__ x,
                                                      T* __restrict__ out) {
  const int w = blockIdx.x * (OPS_NT / 64) + (int)threadIdx.x / 64;
  const int li = threadIdx.x & 63;
  if (w >= count) return;
  const int r = list[w];
  T acc = T(0);
  for (int p = ptr[r] + li; p < ptr[r + 1]; p += 64) acc += val[p] * x[col[p]];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (li == 0) out[r] = acc;
}

template <typename T>
static hipError_t spmv_panels_t(int P, int nranges, const int32_t* range_row0,
                                const int32_t* tile_off, const uint16_t* starts,
                                const int32_t* pcol, const void* pval, int nlong,
                                const int32_t* long_rows, const int32_t* ptr, const int32_t* col,
                                const void* val, const void* x, void* out, hipStream_t s) {
  if (nranges > 0)
    k_spmv_panels<T><<<nranges, PANEL_NT, 0, s>>>(P, range_row0, tile_off, starts, pcol,
                                                  (const T*)pval, (const T*)x, (T*)out);
  if (nlong > 0)
    k_spmv_rows<T><<<(nlong + OPS_NT / 64 - 1) / (OPS_NT / 64), OPS_NT, 0, s>>>(
        nlong, long_rows, ptr, col, (const T*)val, (const T*)x, (T*)out);
  return hipGetLastError();
}

hipError_t op_spmv_panels(int prec, int P, int nranges, const int32_t* range_row0,
                          const int32_t* tile_off, const uint16_t* starts, const int32_t* pcol,
                          const void* pval, int nlong, const int32_t* long_rows,
                          const int32_t* ptr, const int32_t* col, const void* val, const void* x,
                          void* out, hipStream_t s) {
  return prec == 32 ? spmv_panels_t<float>(P, nranges, range_row0, tile_off, starts, pcol, pval,
                                           nlong, long_rows, ptr, col, val, x, out, s)
                    : spmv_panels_t<double>(P, nranges, range_row0, tile_off, starts, pcol,
                                            pval, nlong, long_rows, ptr, col, val, x, out, s);
}

int panel_rows() { return PANEL_ROWS; }
int panel_tile() { return PANEL_TILE; }

hipError_t op_axpby(int prec, int64_t n, double a, const void* x, double b, const void* y,
                    void* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (prec == 32)
    k_axpby<float><<<grid_for(n), OPS_NT, 0, s>>>(n, (float)a, (const float*)x, (float)b,
                                                  (const float*)y, (float*)out);
  else
    k_axpby<double><<<grid_for(n), OPS_NT, 0, s>>>(n, a, (const double*)x, b, (const double*)y,
                                                   (double*)out);
  return hipGetLastError();
}

hipError_t op_scale(int prec, int64_t n, double a, const void* x, void* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (prec == 32)
    k_scale<float><<<grid_for(n), OPS_NT, 0, s>>>(n, (float)a, (const float*)x, (float*)out);
  else
    k_scale<double><<<grid_for(n), OPS_NT, 0, s>>>(n, a, (const double*)x, (double*)out);
  return hipGetLastError();
}

hipError_t op_primal(int prec, int64_t n, const void* x, const void* kty, const void* c,
                     const void* l, const void* u, double tau, void* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (prec == 32)
    k_primal<float><<<grid_for(n), OPS_NT, 0, s>>>(n, (const float*)x, (const float*)kty,
                                                   (const float*)c, (const float*)l,
                                                   (const float*)u, (float)tau, (float*)out);
  else
    k_primal<double><<<grid_for(n), OPS_NT, 0, s>>>(n, (const double*)x, (const double*)kty,
                                                    (const double*)c, (const double*)l,
                                                    (const double*)u, tau, (double*)out);
  return hipGetLastError();
}

hipError_t op_dual(int prec, int64_t m, const void* y, const void* kxb, const void* q,
                   double sigma, int64_t n_eq, void* out, hipStream_t s) {
  if (m == 0) return hipSuccess;
  if (prec == 32)
    k_dual<float><<<grid_for(m), OPS_NT, 0, s>>>(m, (const float*)y, (const float*)kxb,
                                                 (const float*)q, (float)sigma, n_eq,
                                                 (float*)out);
  else
    k_dual<double><<<grid_for(m), OPS_NT, 0, s>>>(m, (const double*)y, (const double*)kxb,
                                                  (const double*)q, sigma, n_eq, (double*)out);
  return hipGetLastError();
}

// part must hold at least 1024 doubles; result lands in part[1024].
hipError_t op_dot(int prec, int64_t n, const void* x, const void* y, double* part,
                  hipStream_t s) {
  int g = grid_for(n);
  if (g > 1024) g = 1024;
  if (prec == 32)
    k_dot_partial<float><<<g, OPS_NT, 0, s>>>(n, (const float*)x, (const float*)y, part);
  else
    k_dot_partial<double><<<g, OPS_NT, 0, s>>>(n, (const double*)x, (const double*)y, part);
  k_sum_partials<<<1, OPS_NT, 0, s>>>(g, part, part + 1024);
  return hipGetLastError();
}

}  // namespace prima
