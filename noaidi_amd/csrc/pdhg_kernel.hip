// Persistent restarted-PDHG solver for CDNA4 (gfx950).
//
// One cooperative launch runs a chunk of the iteration for a batch of LPs.
// Each LP is owned by a group of G workgroups; inside a group every
// workgroup owns a contiguous, nnz-balanced range of constraint rows (rows
// of A) and of variables (rows of the stored transpose A^T). A trial step
// needs two group barriers:
//
//   phase 1 (own variables): [pending commit: kty = A^T yNext, x = xNext,
//            running sums]  then the projected primal step, extrapolation
//            and the |dx|^2 partial                 (Pdhg.scala:341-348)
//   -- barrier A --
//   phase 2 (own rows): kxBar = A xBar fused with the projected dual step,
//            |dy|^2 and dy.Kdx partials             (Pdhg.scala:345-360)
//   -- barrier B --
//   decide: every workgroup reduces the partials of all G workgroups in the
//            same fixed order, so the adaptive step-size decision
//            (Pdhg.scala:356-380) is bit-identical everywhere without a
//            third barrier; on acceptance the row-side commit
//            (kx = (kxBar + kx)/2, y = yNext, sums; Pdhg.scala:373-389) is
//            local to each workgroup's own rows.
//
// Every evaluationFrequency accepted iterations the checkpoint of
// Pdhg.scala:396-417 runs on the device: a fresh K x, the KKT error of the
// current and averaged iterates in original units (Kkt.scala:68-125), the
// Farkas certificates (Certificates.scala:60-219) and the restart rule
// (Pdhg.scala:463-502), all from one fused pass plus one reduction.
//
// No atomics touch numeric data: SpMV runs over an explicit transpose and
// all sums are fixed-order trees, so results are reproducible run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "pdhg_shared.h"
#include "spmv_blocks.h"

namespace prima {


enum {
  E_PSQ_C = 0, E_QY_C, E_YN_C, E_CONE_C, E_ROW_C, E_DYR_C,
  E_PSQ_A, E_QY_A, E_DYR_A,
  E_CX_C, E_DSQ_C, E_DB_C, E_CRSQ_C, E_CVAL_C, E_XN_C, E_BSQ_C, E_DXR_C,
  E_CX_A, E_DSQ_A, E_DB_A, E_DXR_A,
  E_COUNT
};
static_assert(E_COUNT <= kEvalParts, "eval partial slots");

// ---------------------------------------------------------------------------
// reductions

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum K per-thread values over the workgroup; every thread receives the
// totals. Fixed tree order: deterministic.
template <int K>
__device__ inline void block_sum(double (&v)[K], double* red) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[wv * K + k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int w = 0; w < NWAVE; ++w) s += red[w * K + threadIdx.x];
    red[NWAVE * K + threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = red[NWAVE * K + k];
  __syncthreads();
}

// Sum K partials over the G workgroups of a group, identically in each.
template <int K>
__device__ inline void group_sum(const double* part, int stride, int G, double (&out)[K],
                                 double* red) {
  if (threadIdx.x < 64) {
    double s[K];
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = 0.0;
    for (int r = threadIdx.x; r < G; r += 64) {
#pragma unroll
      for (int k = 0; k < K; ++k) s[k] += part[r * stride + k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) s[k] = wave_sum(s[k]);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < K; ++k) red[k] = s[k];
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = red[k];
  __syncthreads();
}

// Poll `flag` (one thread) until it reads `expect`. Every 1024 polls the
// group's abort word bar[2] is read: if another workgroup set it, give up.
// After `limit` polls set the abort word (so the rest of the group gives up
// too) and give up. Returns false when it gave up.
__device__ inline bool bounded_wait(uint32_t* flag, uint32_t expect, uint32_t* bar,
                                    uint64_t limit) {
  const uint32_t zero = (uint32_t)(limit >> 63);  // 0, opaque to the compiler
  uint64_t spins = 0;
  while (__hip_atomic_fetch_add(flag, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
         expect) {
    __builtin_amdgcn_s_sleep(2);
    if ((++spins & 1023u) == 0) {
      if (__hip_atomic_fetch_add(&bar[2], zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
          0u)
        return false;
      if (spins > limit) {
        __hip_atomic_exchange(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
    }
  }
  return true;
}

// Group barrier over the G workgroups of one LP. Words: bar[0] arrivals,
// bar[1] completed barriers (read at launch start into *epoch), bar[2]
// abort; bar[32 + 32 j] the release flag of rank j, each on its own 128-byte
// line. Arrival is one read-modify-write on the count; the last to arrive
// resets it and writes the barrier's number into every other rank's flag,
// one lane per rank, so each waiter polls only its own line. Every poll is a
// read-modify-write (with a zero the compiler cannot fold): those are what
// stays coherent across XCDs without per-poll cache maintenance. The release
// fence before arriving and the acquire fence after leaving order the data.
// A waiter gives up after 2^26 polls (or when another sets the abort word)
// and reports false.
__device__ inline bool group_barrier(uint32_t* bar, int G, int rank, uint32_t* epoch) {
  __syncthreads();
  if (G == 1) return true;
  __shared__ int ok_sh;
  if (threadIdx.x < 64) {
    int ok = 1;
    const uint32_t zero = (uint32_t)G >> 31;
    uint32_t e = 0;
    int last = 0;
    if (threadIdx.x == 0) {
      e = ++*epoch;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const uint32_t arrived =
          __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (arrived == (uint32_t)(G - 1)) {
        // last to arrive: take the others' releases, reset the count (its
        // return is waited for before any flag is written), count the barrier
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        const uint32_t was =
            __hip_atomic_exchange(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&bar[1], 1u + (was & zero), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        last = 1 + (int)(was & zero);
      }
    }
    e = __shfl(e, 0);
    last = __shfl(last, 0);
    if (last) {
      for (int j = threadIdx.x; j < G; j += 64)
        if (j != rank)
          __hip_atomic_exchange(&bar[32 + 32 * j], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
      ok = bounded_wait(&bar[32 + 32 * rank], e, bar, 1ull << 26) ? 1 : 0;
    }
    if (threadIdx.x == 0) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      ok_sh = ok;
    }
  }
  __syncthreads();
  return ok_sh != 0;
}

// ---------------------------------------------------------------------------
// Typed view of an LP: reinterprets the iteration-precision arrays.

template <typename T>
struct Lp {
  const LpView& v;
  __device__ Lp(const LpView& view) : v(view) {}
  __device__ const T* aval() const { return (const T*)v.a_val; }
  __device__ const T* tval() const { return (const T*)v.t_val; }
  __device__ const T* c() const { return (const T*)v.c; }
  __device__ const T* q() const { return (const T*)v.q; }
  __device__ const T* l() const { return (const T*)v.l; }
  __device__ const T* u() const { return (const T*)v.u; }
  __device__ T* x() const { return (T*)v.x; }
  __device__ T* y() const { return (T*)v.y; }
  __device__ T* kx() const { return (T*)v.kx; }
  __device__ T* kty() const { return (T*)v.kty; }
  __device__ T* xn() const { return (T*)v.xn; }
  __device__ T* yn() const { return (T*)v.yn; }
  __device__ T* xbar() const { return (T*)v.xbar; }
  __device__ T* kxbar() const { return (T*)v.kxbar; }
  __device__ T* sx() const { return (T*)v.sx; }
  __device__ T* sy() const { return (T*)v.sy; }
  __device__ T* skx() const { return (T*)v.skx; }
  __device__ T* skty() const { return (T*)v.skty; }
  __device__ T* xr() const { return (T*)v.xr; }
  __device__ T* yr() const { return (T*)v.yr; }
};

// Projected primal step, extrapolation and |dx|^2 (ScalaKernels.scala:78-98,
// Pdhg.scala:341-348) from the committed x_i and kty_i.
template <typename T>
__device__ inline void primal_update(const Lp<T>& L, int i, T xi, T ktyi, T ci, T li, T ui, T tau,
                                     T& dx2) {
  const T st = xi - tau * (ci - ktyi);
  const T xnew = st < li ? li : (st > ui ? ui : st);
  L.xn()[i] = xnew;
  L.xbar()[i] = T(2) * xnew - xi;
  const T d = xnew - xi;
  dx2 += d * d;
}

template <typename T>
struct EpiCommitPrimal {  // rows of A^T: pending column-side commit, then primal step
  const Lp<T>& L;
  T w, tau;
  bool primal;
  T dx2;
  struct Ops {
    T xn, sx, skty, c, l, u;
  };
  __device__ Ops load(int i) const {
    Ops o;
    o.xn = L.xn()[i];
    o.sx = L.sx()[i];
    o.skty = L.skty()[i];
    if (primal) {
      o.c = L.c()[i];
      o.l = L.l()[i];
      o.u = L.u()[i];
    }
    return o;
  }
  __device__ void apply(int i, T s, const Ops& o) {
    L.x()[i] = o.xn;
    L.kty()[i] = s;
    L.sx()[i] = w * o.xn + o.sx;
    L.skty()[i] = w * s + o.skty;
    if (primal) primal_update(L, i, o.xn, s, o.c, o.l, o.u, tau, dx2);
  }
};

template <typename T>
struct EpiDual {  // rows of A: kxBar and the projected dual step (ScalaKernels.scala:100-118)
  const Lp<T>& L;
  T sigma;
  int n_eq;
  T dy2, dykdx;
  // The row-side commit of the previous accepted iteration (Pdhg.scala:373-377,
  // 384-389), deferred into this pass: kx = (kxBar + kx)/2, y = yNext and the
  // running sums, applied before the dual step reads y and kx.
  bool commit;
  T w;
  struct Ops {
    T y, q, kx, lo, kb, sy, skx;
  };
  __device__ Ops load(int r) const {
    Ops o;
    if (commit) {
      o.y = L.yn()[r];
      o.kb = L.kxbar()[r];
      o.sy = L.sy()[r];
      o.skx = L.skx()[r];
    } else {
      o.y = L.y()[r];
    }
    o.q = L.q()[r];
    o.kx = L.kx()[r];
    o.lo = (r >= n_eq && L.v.ylo) ? ((const T*)L.v.ylo)[r] : T(0);
    return o;
  }
  __device__ void apply(int r, T s, const Ops& o) {
    T kxv = o.kx;
    if (commit) {
      kxv = T(0.5) * o.kb + T(0.5) * o.kx;
      L.kx()[r] = kxv;
      L.y()[r] = o.y;
      L.sy()[r] = w * o.y + o.sy;
      L.skx()[r] = w * kxv + o.skx;
    }
    T ynew = o.y + sigma * (o.q - s);
    if (r >= n_eq && !(ynew > o.lo)) ynew = o.lo;
    const T kdx = T(0.5) * s - T(0.5) * kxv;
    const T d = ynew - o.y;
    L.yn()[r] = ynew;
    L.kxbar()[r] = s;
    dy2 += d * d;
    dykdx += d * kdx;
  }
};

// Row-side evaluation partials in original units (Kkt.scala:68-125,
// Certificates.scala:60-195) for the current and, optionally, the averaged
// iterate. Operands are fetched by m_ops so that several rows' loads can be
// in flight together.
template <typename T>
struct MOps {
  T y, yr, sy, skx;
  double rs, q, ylo;
};

template <typename T>
__device__ inline MOps<T> m_ops(const Lp<T>& L, int r, bool avg) {
  const LpView& v = L.v;
  MOps<T> o;
  o.y = L.y()[r];
  o.yr = L.yr()[r];
  if (avg) {
    o.sy = L.sy()[r];
    o.skx = L.skx()[r];
  }
  o.rs = v.rs[r];
  o.q = v.oq[r];
  o.ylo = (v.oylo && r >= v.n_eq) ? v.oylo[r] : 0.0;
  return o;
}

template <typename T>
__device__ inline void m_eval(const LpView& v, int r, T kxv, const MOps<T>& o, bool avg, T inv,
                              double* e) {
  const double rs = o.rs, qv = o.q;
  const T yv = o.y;
  const double yo = (double)yv * rs, kxo = (double)kxv / rs;
  if (r < v.n_eq) {
    const double res = kxo - qv;
    e[E_PSQ_C] += res * res;
    e[E_ROW_C] += kxo * kxo;
  } else {
    const double sf = qv - kxo;
    if (sf > 0.0) e[E_PSQ_C] += sf * sf;
    if (kxo < 0.0) e[E_ROW_C] += kxo * kxo;
    if (yo < 0.0) e[E_CONE_C] += yo * yo;
    if (v.oylo) {
      // shifted cone y >= ylo: the correction problem's objectives carry
      // -ylo' (K x) on the primal side and -ylo' q on the dual side
      const double sh = -o.ylo;
      e[E_CX_C] += sh * kxo;
      e[E_QY_C] += sh * qv;
    }
  }
  e[E_QY_C] += qv * yo;
  e[E_YN_C] += yo * yo;
  const double yrv = (double)o.yr;
  const double dyr = (double)yv - yrv;
  e[E_DYR_C] += dyr * dyr;
  if (avg) {
    const T ya = inv * o.sy;
    const T ka = inv * o.skx;
    const double yao = (double)ya * rs, kao = (double)ka / rs;
    if (r < v.n_eq) {
      const double res = kao - qv;
      e[E_PSQ_A] += res * res;
    } else {
      const double sf = qv - kao;
      if (sf > 0.0) e[E_PSQ_A] += sf * sf;
      if (v.oylo) {
        const double sh = -o.ylo;
        e[E_CX_A] += sh * kao;
        e[E_QY_A] += sh * qv;
      }
    }
    e[E_QY_A] += qv * yao;
    const double d = (double)ya - yrv;
    e[E_DYR_A] += d * d;
  }
}

__device__ inline void dual_terms(double rc, double lo, double hi, double& sq, double& lin) {
  if (rc > 0.0) {
    if (isinf(lo) && lo < 0.0) sq += rc * rc;
    else lin += lo * rc;
  } else if (rc < 0.0) {
    if (isinf(hi) && hi > 0.0) sq += rc * rc;
    else lin += hi * rc;
  }
}

template <typename T>
__device__ inline void n_eval(const Lp<T>& L, int i, bool avg, T inv, double* e) {
  const LpView& v = L.v;
  const double cs = v.cs[i], cv = v.oc[i], lo = v.ol[i], hi = v.ou[i];
  const T xv = L.x()[i];
  const double xo = (double)xv * cs, ko = (double)L.kty()[i] / cs;
  e[E_CX_C] += cv * xo;
  dual_terms(cv - ko, lo, hi, e[E_DSQ_C], e[E_DB_C]);
  dual_terms(-ko, lo, hi, e[E_CRSQ_C], e[E_CVAL_C]);
  e[E_XN_C] += xo * xo;
  if (xo > 0.0 && !(isinf(hi) && hi > 0.0)) e[E_BSQ_C] += xo * xo;
  else if (xo < 0.0 && !(isinf(lo) && lo < 0.0)) e[E_BSQ_C] += xo * xo;
  const double xrv = (double)L.xr()[i];
  const double dxr = (double)xv - xrv;
  e[E_DXR_C] += dxr * dxr;
  if (avg) {
    const T xa = inv * L.sx()[i];
    const T ka = inv * L.skty()[i];
    const double xao = (double)xa * cs, kao = (double)ka / cs;
    e[E_CX_A] += cv * xao;
    dual_terms(cv - kao, lo, hi, e[E_DSQ_A], e[E_DB_A]);
    const double d = (double)xa - xrv;
    e[E_DXR_A] += d * d;
  }
}

template <typename T>
struct EpiFreshKx {  // rows of A: kx = A x recomputed, then row-side evaluation
  const Lp<T>& L;
  bool avg;
  T inv;
  double* e;
  using Ops = MOps<T>;
  __device__ Ops load(int r) const { return m_ops(L, r, avg); }
  __device__ void apply(int r, T s, const Ops& o) {
    L.kx()[r] = s;
    m_eval(L.v, r, s, o, avg, inv, e);
  }
};

template <typename T>
struct EpiInitKty {  // rows of A^T: kty = A^T y at initialisation
  const Lp<T>& L;
  struct Ops {
    T x;
  };
  __device__ Ops load(int i) const { return Ops{L.x()[i]}; }
  __device__ void apply(int i, T s, const Ops& o) {
    L.kty()[i] = s;
    L.xr()[i] = o.x;
    L.sx()[i] = T(0);
    L.skty()[i] = T(0);
  }
};

struct Kkt {
  double pres, dres, pobj, dobj;
};

__device__ inline bool kkt_finite(const Kkt& e) {
  return isfinite(e.pres) && isfinite(e.dres) && isfinite(e.pobj) && isfinite(e.dobj);
}

// Kkt.convergence (Kkt.scala:130-143) against the original data norms.
__device__ inline bool kkt_converged(const Kkt& e, const LpView& v, const KParams& P) {
  const double pt = P.eps_abs + P.eps_rel * v.rhs_norm;
  const double dt = P.eps_abs + P.eps_rel * v.obj_norm;
  const double gt = P.eps_abs + P.eps_rel * (fabs(e.pobj) + fabs(e.dobj));
  return kkt_finite(e) && e.pres <= pt && e.dres <= dt && fabs(e.pobj - e.dobj) <= gt;
}

// KktError.weighted (Kkt.scala:28-33).
__device__ inline double kkt_weighted(const Kkt& e, double w) {
  const double gap = e.pobj - e.dobj;
  return sqrt(w * e.pres * e.pres + (e.dres * e.dres) / w + gap * gap);
}

__device__ inline Kkt kkt_current(const double* E, double offset) {
  return Kkt{sqrt(E[E_PSQ_C]), sqrt(E[E_DSQ_C]), offset + E[E_CX_C],
             offset + E[E_QY_C] + E[E_DB_C]};
}
__device__ inline Kkt kkt_average(const double* E, double offset) {
  return Kkt{sqrt(E[E_PSQ_A]), sqrt(E[E_DSQ_A]), offset + E[E_CX_A],
             offset + E[E_QY_A] + E[E_DB_A]};
}

// Certificates.classify (Certificates.scala:207-219) on the current iterate:
// -1 when neither certificate holds.
__device__ inline int classify(const double* E, const LpView& v, double tol) {
  const double spec = v.spec_orig;
  double nsq = E[E_YN_C];
  if (nsq > 0.0 && isfinite(nsq)) {
    const double norm = sqrt(nsq);
    const double value = E[E_QY_C] + E[E_CVAL_C];
    if (value > 0.0 && isfinite(value)) {
      const double rt = spec > 0.0 ? sqrt(E[E_CRSQ_C]) / (spec * norm) : 0.0;
      const double ct = sqrt(E[E_CONE_C]) / norm;
      const double h = hypot(rt, ct);
      if (isfinite(h) && h <= tol) return 1;  // PrimalInfeasible
    }
  }
  nsq = E[E_XN_C];
  if (nsq > 0.0 && isfinite(nsq)) {
    const double norm = sqrt(nsq);
    const double value = E[E_CX_C];
    if (value < 0.0 && isfinite(value)) {
      const double rt = spec > 0.0 ? sqrt(E[E_ROW_C]) / (spec * norm) : 0.0;
      const double bt = sqrt(E[E_BSQ_C]) / norm;
      const double h = hypot(rt, bt);
      if (isfinite(h) && h <= tol) return 2;  // DualInfeasible
    }
  }
  return -1;
}

// ---------------------------------------------------------------------------
// The persistent kernel. Cold paths (checkpoint evaluation, restart, result
// write-back) are separate non-inlined functions so that their registers do
// not burden the trial loop; reduced totals live in LDS.

struct Ranges {
  int v0, v1, r0, r1;      // own variables and rows
  int tb0, tb1, ab0, ab1;  // own row blocks of A^T and A
  int rank, G;
};

// Workgroup scratch in LDS, typed on the iteration precision: the streamed
// products and row sums of a block (TILE and ROWCAP entries of T) and its
// row pointers, plus the fixed-order reduction slots.
template <typename T>
struct Smem {
  T tile[TILE];
  T rsum[ROWCAP];
  int rptr[ROWCAP + 1];
  double red[NWAVE * kEvalParts + kEvalParts];
  double E[kEvalParts];  // group totals of the last evaluation
  uint32_t epoch;        // group barriers completed (thread 0's copy)
};

// One evaluation pass (Kkt.evaluate + certificate inputs) over own rows and
// variables, reduced over the group into sm.E. fresh_kx recomputes K x first
// (Pdhg.scala:399). Returns false if a group barrier aborted.
template <typename T>
__device__ __noinline__ bool eval_pass(const LpView& v, const Ranges& R, bool fresh_kx, bool avg,
                                       double wsum, Smem<T>& sm) {
  const Lp<T> L(v);
  const T inv = avg ? (T)(1.0 / wsum) : T(0);
  double e[E_COUNT];
#pragma unroll
  for (int k = 0; k < E_COUNT; ++k) e[k] = 0.0;
  if (fresh_kx) {
    EpiFreshKx<T> epi{L, avg, inv, e};
    spmv_blocks(v.a_ptr, v.a_col, L.aval(), (const T*)L.x(), v.a_rb, R.ab0, R.ab1,
                sm.tile, sm.rsum, sm.rptr, epi);
  } else {
    for (int r = R.r0 + threadIdx.x; r < R.r1; r += NT)
      m_eval(v, r, L.kx()[r], m_ops(L, r, avg), avg, inv, e);
  }
  for (int i = R.v0 + threadIdx.x; i < R.v1; i += NT) n_eval(L, i, avg, inv, e);
  block_sum<E_COUNT>(e, sm.red);
  double* epart = v.part + 2 * R.G * kIterParts;
  if (threadIdx.x == 0) {
    for (int k = 0; k < E_COUNT; ++k) epart[R.rank * kEvalParts + k] = e[k];
  }
  if (!group_barrier(v.bar, R.G, R.rank, &sm.epoch)) return false;
  double tot[E_COUNT];
  group_sum<E_COUNT>(epart, kEvalParts, R.G, tot, sm.red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < E_COUNT; ++k) sm.E[k] = tot[k];
  }
  __syncthreads();
  return true;
}

// Result vectors in original units (Pdhg.scala:510-532): the current or
// the averaged iterate.
template <typename T>
__device__ __noinline__ void write_outputs(const LpView& v, const Ranges& R, bool avg,
                                          double wsum) {
  const Lp<T> L(v);
  const T inv = avg ? (T)(1.0 / wsum) : T(0);
  for (int i = R.v0 + threadIdx.x; i < R.v1; i += NT) {
    T xv = L.x()[i], kv = L.kty()[i];
    if (avg) {
      xv = inv * L.sx()[i];
      kv = inv * L.skty()[i];
    }
    const double cs = v.cs[i];
    v.x_out[i] = (double)xv * cs;
    v.rc_out[i] = v.oc[i] - (double)kv / cs;
  }
  for (int r = R.r0 + threadIdx.x; r < R.r1; r += NT) {
    T yv = L.y()[r];
    if (avg) yv = inv * L.sy()[r];
    v.y_out[r] = (double)yv * v.rs[r];
  }
}

// restartAt (Pdhg.scala:478-502), the array part: re-centre on the chosen
// iterate, record it as the restart point and clear the running sums. Every
// array touched is in the workgroup's own range.
template <typename T>
__device__ __noinline__ void restart_arrays(const LpView& v, const Ranges& R, bool use_avg,
                                           double wsum) {
  const Lp<T> L(v);
  const T inv = use_avg ? (T)(1.0 / wsum) : T(0);
  for (int i = R.v0 + threadIdx.x; i < R.v1; i += NT) {
    T xv = L.x()[i];
    if (use_avg) {
      xv = inv * L.sx()[i];
      L.x()[i] = xv;
      L.kty()[i] = inv * L.skty()[i];
    }
    L.xr()[i] = xv;
    L.sx()[i] = T(0);
    L.skty()[i] = T(0);
  }
  for (int r = R.r0 + threadIdx.x; r < R.r1; r += NT) {
    T yv = L.y()[r];
    if (use_avg) {
      yv = inv * L.sy()[r];
      L.y()[r] = yv;
      L.kx()[r] = inv * L.skx()[r];
    }
    L.yr()[r] = yv;
    L.sy()[r] = T(0);
    L.skx()[r] = T(0);
  }
  __syncthreads();
}

// Pending column-side commit (x = xNext, kty = A^T yNext, running sums),
// optionally followed by the next primal step; returns this thread's |dx|^2.
template <typename T>
__device__ inline T commit_columns(const LpView& v, const Ranges& R, T w, bool primal, T tau,
                                   Smem<T>& sm) {
  const Lp<T> L(v);
  EpiCommitPrimal<T> epi{L, w, tau, primal, T(0)};
  spmv_blocks(v.t_ptr, v.t_col, L.tval(), (const T*)L.yn(), v.t_rb, R.tb0, R.tb1,
              sm.tile, sm.rsum, sm.rptr, epi);
  return epi.dx2;
}

template <typename T>
__device__ __noinline__ void commit_columns_cold(const LpView& v, const Ranges& R, T w,
                                                 Smem<T>& sm) {
  (void)commit_columns<T>(v, R, w, false, T(0), sm);
}

// The deferred row-side commit on its own, before a checkpoint or the final
// evaluation: the same arithmetic as EpiDual's commit branch.
template <typename T>
__device__ __noinline__ void commit_rows_cold(const LpView& v, const Ranges& R, T w) {
  const Lp<T> L(v);
  for (int r = R.r0 + threadIdx.x; r < R.r1; r += NT) {
    const T kxv = T(0.5) * L.kxbar()[r] + T(0.5) * L.kx()[r];
    const T yv = L.yn()[r];
    L.kx()[r] = kxv;
    L.y()[r] = yv;
    L.sy()[r] = w * yv + L.sy()[r];
    L.skx()[r] = w * kxv + L.skx()[r];
  }
  __syncthreads();
}

// Initialisation on the device (Pdhg.scala:305-309): kty = A^T y, restart
// point, zero running sums.
template <typename T>
__device__ __noinline__ void init_arrays(const LpView& v, const Ranges& R, Smem<T>& sm) {
  const Lp<T> L(v);
  EpiInitKty<T> epi{L};
  spmv_blocks(v.t_ptr, v.t_col, L.tval(), (const T*)L.y(), v.t_rb, R.tb0, R.tb1,
              sm.tile, sm.rsum, sm.rptr, epi);
  for (int r = R.r0 + threadIdx.x; r < R.r1; r += NT) {
    L.yr()[r] = L.y()[r];
    L.sy()[r] = T(0);
    L.skx()[r] = T(0);
  }
}

__device__ __noinline__ void store_state(const LpView& v, const Ranges& R, const SolveState& S) {
  if (R.rank == 0 && threadIdx.x == 0) *v.st = S;
}

template <typename T>
__device__ __noinline__ void finish(const LpView& v, const Ranges& R, SolveState& S, int status,
                                   const Kkt& err, bool avg) {
  write_outputs<T>(v, R, avg, S.wsum);
  S.status = status;
  S.kkt[0] = err.pres;
  S.kkt[1] = err.dres;
  S.kkt[2] = err.pobj;
  S.kkt[3] = err.dobj;
  S.used_average = avg ? 1 : 0;
  store_state(v, R, S);
}

// A group barrier gave up (a workgroup of the group never arrived, or
// another gave up first). The LP's state is left as it was at the launch's
// start; the launch-wide abort word records the LP (index + 1) so that the
// host reports a device error instead of a solver outcome.
__device__ __noinline__ void abort_launch(const LpView& v, int lp) {
  if (threadIdx.x == 0)
    __hip_atomic_exchange(v.launch_abort, (uint32_t)lp + 1u, __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT);
}

// The checkpoint decision (Pdhg.scala:396-417, 463-502) from the group
// totals in sm.E. Returns a terminal status (and sets *avg_out) or -1; on a
// restart the arrays are re-centred.
template <typename T>
__device__ __noinline__ int checkpoint_decide(const LpView& v, const Ranges& R, const KParams& P,
                                              SolveState& S, bool have_avg, Smem<T>& sm,
                                              bool* avg_out, Kkt* err_out) {
  const double* E = sm.E;
  const Kkt cur = kkt_current(E, v.offset);
  *avg_out = false;
  *err_out = cur;
  if (kkt_converged(cur, v, P)) return 0;
  const int cls = P.certificates ? classify(E, v, P.infeas_tol) : -1;
  if (cls > 0) return cls;
  const double cur_err = kkt_weighted(cur, S.weight);
  Kkt avg = cur;
  double avg_err = INFINITY;
  if (have_avg) {
    avg = kkt_average(E, v.offset);
    if (kkt_converged(avg, v, P)) {
      *avg_out = true;
      *err_out = avg;
      return 0;
    }
    avg_err = kkt_weighted(avg, S.weight);
  }
  const bool use_avg = avg_err < cur_err;
  const double cand = (isnan(cur_err) || isnan(avg_err)) ? NAN : fmin(cur_err, avg_err);
  const bool sufficient = cand <= P.r_suff * S.err_at_restart;
  const bool necessary = cand <= P.r_nec * S.err_at_restart && cand > S.last_candidate;
  const bool artificial =
      (double)(S.iteration - S.last_restart_it) >= P.r_art * (double)S.iteration;
  if (sufficient || necessary || artificial) {
    restart_arrays<T>(v, R, use_avg, S.wsum);
    const double pm = sqrt(use_avg ? E[E_DXR_A] : E[E_DXR_C]);
    const double dm = sqrt(use_avg ? E[E_DYR_A] : E[E_DYR_C]);
    if (pm > 0.0 && dm > 0.0) {
      const double nx = exp(P.theta * log(dm / pm) + (1.0 - P.theta) * log(S.weight));
      if (isfinite(nx) && nx > 0.0) S.weight = nx;
    }
    S.wsum = 0.0;
    S.restarts++;
    S.last_restart_it = S.iteration;
    S.err_at_restart = kkt_weighted(use_avg ? avg : cur, S.weight);
    S.last_candidate = INFINITY;
  } else {
    S.last_candidate = cand;
  }
  return -1;
}

// One LP's chunk of a launch, run by one workgroup of its group: `me` is
// the LP index << 9 | the workgroup's rank in the group.
template <typename T>
__device__ inline void run_chunk(const LpView* __restrict__ views, int32_t me, const KParams& P,
                                 Smem<T>& sm) {
  T* tile = sm.tile;
  const int lp = me >> 9;
  const LpView& v = views[lp];
  const Lp<T> L(v);
  Ranges R;
  R.G = v.G;
  R.rank = me & 511;
  R.tb0 = v.t_wrb[R.rank];
  R.tb1 = v.t_wrb[R.rank + 1];
  R.ab0 = v.a_wrb[R.rank];
  R.ab1 = v.a_wrb[R.rank + 1];
  R.v0 = v.t_rb[2 * R.tb0];
  R.v1 = v.t_rb[2 * R.tb1];
  R.r0 = v.a_rb[2 * R.ab0];
  R.r1 = v.a_rb[2 * R.ab1];
  const int G = R.G;
  // barriers this LP's group has completed so far: the base of this launch's
  // barrier numbers (read before this workgroup first arrives, so every
  // workgroup of the group reads the same value)
  if (threadIdx.x == 0)
    sm.epoch = __hip_atomic_fetch_add(&v.bar[1], (uint32_t)G >> 31, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);

  SolveState S = *v.st;
  if (S.status >= 0) return;
  double* ipart = v.part;  // [2][G][kIterParts]

  if (P.test_stall && G > 1 && R.rank == G - 1) {
    // test hook: the last rank waits for bar[3] (a word nothing writes) to
    // read 1, with a short limit, so that its bounded wait gives up, sets
    // the group's abort word and the rest of the group gives up in turn
    if (threadIdx.x == 0) (void)bounded_wait(&v.bar[3], 1u, v.bar, 1u << 14);
    __syncthreads();
    return abort_launch(v, lp);
  }

  if (P.mode == 1) {
    // finalise: an LP that never ran (still waiting for workgroups) is
    // initialised first and reports its starting point
    const bool fresh = S.status == -2;
    if (fresh) {
      init_arrays<T>(v, R, sm);
      if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);
      S.wsum = 0.0;
    }
    if (!eval_pass<T>(v, R, fresh, false, S.wsum, sm)) return abort_launch(v, lp);
    return finish<T>(v, R, S, P.final_status, kkt_current(sm.E, v.offset), false);
  }

  if (S.status == -2) {
    init_arrays<T>(v, R, sm);
    if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);
    S.wsum = 0.0;
    if (!eval_pass<T>(v, R, true, false, 0.0, sm)) return abort_launch(v, lp);
    const Kkt cur = kkt_current(sm.E, v.offset);
    if (kkt_converged(cur, v, P)) return finish<T>(v, R, S, 0, cur, false);
    S.err_at_restart = kkt_weighted(cur, S.weight);
    S.status = -1;
  }

  const int it_start = S.iteration;
  bool pending = false;       // column-side commit owed (applied in phase 1)
  bool pending_rows = false;  // row-side commit owed (applied in phase 2)
  T wpend = T(0);

  for (;;) {
    const T tau = (T)(S.step / S.weight);
    const T sigma = (T)(S.step * S.weight);
    const int slot = S.trial_counter & 1;
    double* ip = ipart + (slot * G + R.rank) * kIterParts;

    // phase 1: own variables
    T dx2 = T(0);
    if (pending) {
      dx2 = commit_columns<T>(v, R, wpend, true, tau, sm);
      pending = false;
    } else {
      for (int i = R.v0 + threadIdx.x; i < R.v1; i += NT)
        primal_update(L, i, L.x()[i], L.kty()[i], L.c()[i], L.l()[i], L.u()[i], tau, dx2);
    }
    {
      double a[1] = {(double)dx2};
      block_sum<1>(a, sm.red);
      if (threadIdx.x == 0) ip[0] = a[0];
    }
    if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);

    // phase 2: own rows
    {
      EpiDual<T> epi{L, sigma, v.n_eq, T(0), T(0), pending_rows, wpend};
      spmv_blocks(v.a_ptr, v.a_col, L.aval(), (const T*)L.xbar(), v.a_rb, R.ab0, R.ab1, tile,
                  sm.rsum, sm.rptr, epi);
      pending_rows = false;
      double a[2] = {(double)epi.dy2, (double)epi.dykdx};
      block_sum<2>(a, sm.red);
      if (threadIdx.x == 0) {
        ip[1] = a[0];
        ip[2] = a[1];
      }
    }
    if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);

    // decide (identical in every workgroup)
    double tot[3];
    group_sum<3>(ipart + slot * G * kIterParts, kIterParts, G, tot, sm.red);
    S.trial_counter++;
    S.trials++;
    S.total_trials++;
    const double movement = 0.5 * (S.weight * tot[0] + tot[1] / S.weight);

    if (movement == 0.0) {
      // fixed point (Pdhg.scala:221-231): fresh K x, then the usual test
      S.iteration++;
      S.trials = 0;
      if (!eval_pass<T>(v, R, true, false, S.wsum, sm)) return abort_launch(v, lp);
      const Kkt cur = kkt_current(sm.E, v.offset);
      return finish<T>(v, R, S, kkt_converged(cur, v, P) ? 0 : 5, cur, false);
    }

    const double inter = fabs(tot[2]);
    const double limit = inter > 0.0 ? movement / inter : INFINITY;
    const double t = (double)(S.iteration + 1);
    const double by_rule = (1.0 - pow(t, -0.3)) * limit;
    const double by_cap = (1.0 + pow(t, -0.6)) * S.step;
    // math.min propagates NaN (0 * inf at the first iteration)
    const double proposed = (isnan(by_rule) || isnan(by_cap)) ? NAN : fmin(by_rule, by_cap);
    const bool accepted = S.step <= limit || S.trials >= 60;

    if (accepted) {
      S.wsum += S.step;
      pending = true;
      pending_rows = true;
      wpend = (T)S.step;
      S.iteration++;
      S.trials = 0;
    }
    if (isfinite(proposed) && proposed > 0.0) S.step = proposed;
    if (!accepted) continue;

    bool evaluated_now = false;
    if (S.iteration % P.eval_freq == 0) {
      // checkpoint (Pdhg.scala:396-417)
      commit_rows_cold<T>(v, R, wpend);
      pending_rows = false;
      commit_columns_cold<T>(v, R, wpend, sm);
      pending = false;
      if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);
      const bool have_avg = S.wsum > 0.0;
      if (!eval_pass<T>(v, R, true, have_avg, S.wsum, sm)) return abort_launch(v, lp);
      S.evals++;
      evaluated_now = true;
      bool avg = false;
      Kkt err;
      const int st = checkpoint_decide<T>(v, R, P, S, have_avg, sm, &avg, &err);
      if (st >= 0) return finish<T>(v, R, S, st, err, avg);
    }

    if (S.iteration >= v.max_iter) {
      if (pending_rows) {
        commit_rows_cold<T>(v, R, wpend);
        pending_rows = false;
      }
      if (pending) {
        commit_columns_cold<T>(v, R, wpend, sm);
        pending = false;
      }
      // The evaluation partials are single-buffered: when the checkpoint
      // above has just reduced them, every rank must be done reading its
      // group's slots before any rank overwrites its own again.
      if (evaluated_now) {
        if (!group_barrier(v.bar, G, R.rank, &sm.epoch)) return abort_launch(v, lp);
      } else {
        __syncthreads();
      }
      if (!eval_pass<T>(v, R, false, false, S.wsum, sm)) return abort_launch(v, lp);
      return finish<T>(v, R, S, 3, kkt_current(sm.E, v.offset), false);
    }

    if (evaluated_now && S.iteration - it_start >= v.iter_budget) {
      // end of this launch's chunk: state is fully committed
      return store_state(v, R, S);
    }
  }
}

// Every workgroup runs the chunks its list names, in order: first the LP
// whose group it belongs to, if any (so every group's workgroups start on it
// together), then any number of LPs it runs alone.
#ifndef PRIMA_WAVES_PER_EU
#define PRIMA_WAVES_PER_EU 2
#endif
template <typename T>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(PRIMA_WAVES_PER_EU)))
void pdhg_persistent(const LpView* __restrict__ views,
                                                      const int32_t* __restrict__ wg_off,
                                                      const int32_t* __restrict__ wg_ent,
                                                      KParams P) {
  __shared__ Smem<T> sm;
  const int e1 = wg_off[blockIdx.x + 1];
  for (int e = wg_off[blockIdx.x]; e < e1; ++e) {
    run_chunk<T>(views, wg_ent[e], P, sm);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host-side launchers

int occupancy_blocks_per_cu(int precision) {
  int nb = 0;
  hipError_t e;
  if (precision == 32)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pdhg_persistent<float>, NT, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pdhg_persistent<double>, NT, 0);
  return e == hipSuccess ? nb : 0;
}

hipError_t launch_pdhg(int precision, const LpView* views, const int32_t* wg_off,
                       const int32_t* wg_ent, int grid, const KParams& P, hipStream_t stream) {
  void* args[] = {(void*)&views, (void*)&wg_off, (void*)&wg_ent, (void*)&P};
  if (precision == 32)
    return hipLaunchCooperativeKernel((const void*)pdhg_persistent<float>, dim3(grid), dim3(NT),
                                      args, 0, stream);
  return hipLaunchCooperativeKernel((const void*)pdhg_persistent<double>, dim3(grid), dim3(NT),
                                    args, 0, stream);
}

// Device workspace of the allocation-free solve (prima_solve_noalloc): one
// LP's arena, its view, the launch plan and its state, carved from this
// array. Each device the library runs on holds its own copy until the
// process ends; calls that use it concurrently on one device race.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxDevices] = {};
}  // namespace

// Takes the address of this device's copy of `scratch` (the device must be
// current). Called once per device, when a context is opened on it; the name
// `scratch` itself is a host-side stand-in, never a device address.
hipError_t scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  scratch_ptr[device] = p;
  return hipSuccess;
}

// The device address and size of this device's workspace (null before
// scratch_init succeeded on it).
void* scratch_of(int device, size_t* bytes) {
  *bytes = kScratchBytes;
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(scratch_mu);
  return scratch_ptr[device];
}

int threads_per_block() { return NT; }
int block_row_cap() { return ROWCAP; }
int tile_nnz() { return TILE; }

}  // namespace prima
