// CSR-Adaptive row-block SpMV shared by the persistent solver
// (pdhg_kernel.hip) and the per-operation seam (kernels_ops.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prima {

#ifndef PRIMA_NT
#define PRIMA_NT 512
#endif
#ifndef PRIMA_TILE
#define PRIMA_TILE 8192
#endif
#ifndef PRIMA_ROWCAP
#define PRIMA_ROWCAP 4096
#endif
#ifndef PRIMA_DEPTH
#define PRIMA_DEPTH 16
#endif
#ifndef PRIMA_EPI_ROWS
#define PRIMA_EPI_ROWS 4
#endif
constexpr int NT = PRIMA_NT;          // threads per workgroup
constexpr int NWAVE = NT / 64;
constexpr int TILE = PRIMA_TILE;      // nnz of a streamed row block (staged in LDS)
constexpr int ROWCAP = PRIMA_ROWCAP;  // rows of a streamed row block
constexpr int DEPTH = PRIMA_DEPTH;    // independent product loads in flight per thread
constexpr int EPI_ROWS = PRIMA_EPI_ROWS;  // rows whose epilogue operands are in flight

// ---------------------------------------------------------------------------
// CSR SpMV over a workgroup's row blocks (CSR-Adaptive style). A block of
// short rows streams val*x[col] into LDS with coalesced, 8-deep loads while
// its row pointers are staged in LDS; each row is then summed from LDS by a
// power-of-two sub-wave into an LDS row-sum array; finally a coalesced pass
// runs the epilogue with the operand loads of four rows in flight per thread.
// A block holding one row longer than the tile is summed by the whole
// workgroup. Epilogues are two-phase: load(r) fetches a row's operands,
// apply(r, sum, ops) consumes them. TL is the capacity of `tile` in entries
// (the row blocks must have been cut with it).
template <typename T, typename Epi, int TL = TILE>
__device__ inline void spmv_blocks(const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                                   const T* __restrict__ val, const T* __restrict__ xv,
                                   const int32_t* __restrict__ rbd, int b0, int b1, T* tile,
                                   T* rsum, int* rptr, Epi& epi) {
  // rbd holds {first row, first nnz} per block, so a block's extent is two
  // independent loads rather than a row-block -> row-pointer chain.
  const int2* __restrict__ bd = reinterpret_cast<const int2*>(rbd);
  for (int b = b0; b < b1; ++b) {
    const int2 d0 = bd[b], d1 = bd[b + 1];
    const int r0 = d0.x, r1 = d1.x;
    const int p0 = d0.y, p1 = d1.y;
    const int nnz = p1 - p0, nrows = r1 - r0;
    if (nnz > TL) {
      T acc = T(0);
      for (int p = p0 + (int)threadIdx.x; p < p1; p += NT) acc += val[p] * xv[col[p]];
      // fixed-order workgroup tree in the iteration precision
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if ((threadIdx.x & 63) == 0) tile[threadIdx.x >> 6] = acc;
      __syncthreads();
      if (threadIdx.x == 0) {
        T s = T(0);
        for (int w = 0; w < NWAVE; ++w) s += tile[w];
        const typename Epi::Ops o = epi.load(r0);
        epi.apply(r0, s, o);
      }
      __syncthreads();
      continue;
    }
    for (int r = threadIdx.x; r <= nrows; r += NT) rptr[r] = ptr[r0 + r] - p0;
    int k = threadIdx.x;
    for (; k + (DEPTH - 1) * NT < nnz; k += DEPTH * NT) {
      int c[DEPTH];
      T v[DEPTH], x[DEPTH];
#pragma unroll
      for (int j = 0; j < DEPTH; ++j) c[j] = __builtin_nontemporal_load(&col[p0 + k + j * NT]);
#pragma unroll
      for (int j = 0; j < DEPTH; ++j) v[j] = __builtin_nontemporal_load(&val[p0 + k + j * NT]);
#pragma unroll
      for (int j = 0; j < DEPTH; ++j) x[j] = xv[c[j]];
#pragma unroll
      for (int j = 0; j < DEPTH; ++j) tile[k + j * NT] = v[j] * x[j];
    }
    for (; k + 3 * NT < nnz; k += 4 * NT) {
      int c[4];
      T v[4], x[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = col[p0 + k + j * NT];
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = val[p0 + k + j * NT];
#pragma unroll
      for (int j = 0; j < 4; ++j) x[j] = xv[c[j]];
#pragma unroll
      for (int j = 0; j < 4; ++j) tile[k + j * NT] = v[j] * x[j];
    }
    for (; k + NT < nnz; k += 2 * NT) {
      const int c0 = col[p0 + k], c1 = col[p0 + k + NT];
      const T v0 = val[p0 + k], v1 = val[p0 + k + NT];
      const T x0 = xv[c0], x1 = xv[c1];
      tile[k] = v0 * x0;
      tile[k + NT] = v1 * x1;
    }
    if (k < nnz) tile[k] = val[p0 + k] * xv[col[p0 + k]];
    __syncthreads();
    int tpr = 1;
    if (nrows < NT) {
      tpr = NT / (nrows > 0 ? nrows : 1);
      tpr = 1 << (31 - __clz(tpr));
      if (tpr > 64) tpr = 64;
    }
    const int li = threadIdx.x & (tpr - 1);
    const int stride = NT / tpr;
    for (int r = threadIdx.x / tpr; r < nrows; r += stride) {
      const int s = rptr[r], e = rptr[r + 1];
      T acc = T(0);
      for (int q = s + li; q < e; q += tpr) acc += tile[q];
      for (int off = tpr >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if (li == 0) rsum[r] = acc;
    }
    __syncthreads();
    for (int base = threadIdx.x; base < nrows; base += EPI_ROWS * NT) {
      typename Epi::Ops o[EPI_ROWS];
#pragma unroll
      for (int j = 0; j < EPI_ROWS; ++j)
        if (base + j * NT < nrows) o[j] = epi.load(r0 + base + j * NT);
#pragma unroll
      for (int j = 0; j < EPI_ROWS; ++j)
        if (base + j * NT < nrows) epi.apply(r0 + base + j * NT, rsum[base + j * NT], o[j]);
    }
    __syncthreads();
  }
}

}  // namespace prima
