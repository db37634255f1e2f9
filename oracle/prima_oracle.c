/* CPU oracle for the Prima restarted-PDHG path. TEST INFRASTRUCTURE ONLY:
 * linked by tests/ and by bench.py's CPU-baseline leg, never by the product.
 * Compile with -O2 -ffp-contract=off (the JVM never contracts to FMA).
 * Each function cites the reference file:line it restates; paths are
 * relative to modules/prima-core/src/main/scala/org/noaidi/prima/. */
#include "prima_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* fdlibm __ieee754_log, the algorithm behind java.lang.StrictMath.log, which
 * java.util.Random.nextGaussian uses. Restated from the published fdlibm
 * method: x = 2^k (1+f), log(1+f) = f - s(f - R(z)) with s = f/(2+f) and a
 * degree-14 even minimax polynomial R in s. */

static uint32_t hi_word(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return (uint32_t)(b >> 32);
}
static uint32_t lo_word(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return (uint32_t)b;
}
static double with_hi(double x, uint32_t hi) {
  uint64_t b;
  memcpy(&b, &x, 8);
  b = ((uint64_t)hi << 32) | (b & 0xffffffffull);
  memcpy(&x, &b, 8);
  return x;
}

double orc_fdlibm_log(double x) {
  static const double ln2_hi = 6.93147180369123816490e-01; /* 3fe62e42 fee00000 */
  static const double ln2_lo = 1.90821492927058770002e-10; /* 3dea39ef 35793c76 */
  static const double two54 = 1.80143985094819840000e+16;
  static const double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                      Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                      Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                      Lg7 = 1.479819860511658591e-01;
  volatile double zero = 0.0;
  int32_t hx = (int32_t)hi_word(x);
  uint32_t lx = lo_word(x);
  int32_t k = 0;
  if (hx < 0x00100000) {
    if (((hx & 0x7fffffff) | (int32_t)lx) == 0) return -two54 / zero;
    if (hx < 0) return (x - x) / zero;
    k -= 54;
    x *= two54;
    hx = (int32_t)hi_word(x);
  }
  if (hx >= 0x7ff00000) return x + x;
  k += (hx >> 20) - 1023;
  hx &= 0x000fffff;
  int32_t i = (hx + 0x95f64) & 0x100000;
  x = with_hi(x, (uint32_t)(hx | (i ^ 0x3ff00000)));
  k += (i >> 20);
  double f = x - 1.0;
  double dk, R;
  if ((0x000fffff & (2 + hx)) < 3) {
    if (f == 0.0) {
      if (k == 0) return 0.0;
      dk = (double)k;
      return dk * ln2_hi + dk * ln2_lo;
    }
    R = f * f * (0.5 - 0.33333333333333333 * f);
    if (k == 0) return f - R;
    dk = (double)k;
    return dk * ln2_hi - ((R - dk * ln2_lo) - f);
  }
  double s = f / (2.0 + f);
  dk = (double)k;
  double z = s * s;
  i = hx - 0x6147a;
  double w = z * z;
  int32_t j = 0x6b851 - hx;
  double t1 = w * (Lg2 + w * (Lg4 + w * Lg6));
  double t2 = z * (Lg1 + w * (Lg3 + w * (Lg5 + w * Lg7)));
  i |= j;
  R = t2 + t1;
  if (i > 0) {
    double hfsq = 0.5 * f * f;
    if (k == 0) return f - (hfsq - s * (hfsq + R));
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
  }
  if (k == 0) return f - s * (f - R);
  return dk * ln2_hi - ((s * (f - R) - dk * ln2_lo) - f);
}

/* ------------------------------------------------------------------------ */
/* java.util.Random: 48-bit LCG (multiplier 0x5DEECE66D, increment 0xB),
 * nextDouble from 53 bits, nextInt(bound) with rejection, and the polar
 * Marsaglia nextGaussian with StrictMath log/sqrt. scala.util.Random(seed)
 * wraps exactly this (LpFixtures.scala:181). */

#define JR_MULT 0x5DEECE66DLL
#define JR_MASK ((1LL << 48) - 1)

void orc_jrandom_init(orc_jrandom* r, int64_t seed) {
  r->seed = (seed ^ JR_MULT) & JR_MASK;
  r->have_next_next_gaussian = 0;
  r->next_next_gaussian = 0.0;
}

static int32_t jr_next(orc_jrandom* r, int bits) {
  r->seed = (int64_t)(((uint64_t)r->seed * (uint64_t)JR_MULT + 0xBull) & (uint64_t)JR_MASK);
  return (int32_t)(uint32_t)((uint64_t)r->seed >> (48 - bits));
}

double orc_jrandom_next_double(orc_jrandom* r) {
  int64_t a = (int64_t)jr_next(r, 26);
  int64_t b = (int64_t)jr_next(r, 27);
  return (double)((a << 27) + b) * 0x1.0p-53;
}

int32_t orc_jrandom_next_int(orc_jrandom* r, int32_t bound) {
  int32_t v = jr_next(r, 31);
  int32_t m = bound - 1;
  if ((bound & m) == 0) return (int32_t)(((int64_t)bound * (int64_t)v) >> 31);
  for (int32_t u = v;; u = jr_next(r, 31)) {
    v = u % bound;
    /* u - v + m < 0 in Java int arithmetic (wrapping). */
    int32_t t = (int32_t)((uint32_t)u - (uint32_t)v + (uint32_t)m);
    if (t >= 0) break;
  }
  return v;
}

double orc_jrandom_next_gaussian(orc_jrandom* r) {
  if (r->have_next_next_gaussian) {
    r->have_next_next_gaussian = 0;
    return r->next_next_gaussian;
  }
  double v1, v2, s;
  do {
    v1 = 2 * orc_jrandom_next_double(r) - 1;
    v2 = 2 * orc_jrandom_next_double(r) - 1;
    s = v1 * v1 + v2 * v2;
  } while (s >= 1 || s == 0);
  double multiplier = sqrt(-2 * orc_fdlibm_log(s) / s);
  r->next_next_gaussian = v2 * multiplier;
  r->have_next_next_gaussian = 1;
  return v1 * multiplier;
}

/* ------------------------------------------------------------------------ */
/* CSR matrix (SparseMatrix.scala:15-21). */

typedef struct {
  int32_t rows, cols;
  int32_t* ptr;
  int32_t* col;
  double* val;
} csr;

struct orc_lp {
  int32_t n, m, n_eq;
  csr a;
  csr at; /* transpose, materialised once (SparseMatrix.scala:36-64) */
  double *c, *q, *l, *u;
  double offset;
};

static void* xmalloc(size_t n) {
  void* p = malloc(n ? n : 1);
  if (!p) abort();
  return p;
}
static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

static void csr_free(csr* a) {
  free(a->ptr);
  free(a->col);
  free(a->val);
  memset(a, 0, sizeof(*a));
}

/* SparseMatrix.transpose (SparseMatrix.scala:36-64): counting sort by column,
 * rows visited ascending so each transposed row is sorted. */
static void csr_transpose(const csr* a, csr* t) {
  int32_t nnz = a->ptr[a->rows];
  t->rows = a->cols;
  t->cols = a->rows;
  t->ptr = (int32_t*)calloc((size_t)a->cols + 1, sizeof(int32_t));
  t->col = (int32_t*)xmalloc((size_t)nnz * sizeof(int32_t));
  t->val = (double*)xmalloc((size_t)nnz * sizeof(double));
  for (int32_t k = 0; k < nnz; k++) t->ptr[a->col[k] + 1]++;
  for (int32_t c = 0; c < a->cols; c++) t->ptr[c + 1] += t->ptr[c];
  int32_t* cursor = (int32_t*)xmalloc(((size_t)a->cols + 1) * sizeof(int32_t));
  memcpy(cursor, t->ptr, (size_t)a->cols * sizeof(int32_t));
  for (int32_t r = 0; r < a->rows; r++) {
    for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) {
      int32_t dest = cursor[a->col[p]]++;
      t->col[dest] = r;
      t->val[dest] = a->val[p];
    }
  }
  free(cursor);
}

/* multiplyInto (SparseMatrix.scala:76-86): sequential fp64, index order. */
static void csr_mul(const csr* a, const double* x, double* out) {
  for (int32_t r = 0; r < a->rows; r++) {
    double acc = 0.0;
    for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) acc += a->val[p] * x[a->col[p]];
    out[r] = acc;
  }
}

/* normInf (SparseMatrix.scala:144-156). */
static double csr_norm_inf(const csr* a) {
  double m = 0.0;
  for (int32_t r = 0; r < a->rows; r++) {
    double s = 0.0;
    for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) s += fabs(a->val[p]);
    if (s > m) m = s;
  }
  return m;
}

/* spectralNormBound = sqrt(norm1 * normInf) with norm1 = transpose.normInf
 * (SparseMatrix.scala:141, 181). */
static double csr_spectral_bound(const csr* a, const csr* at) {
  return sqrt(csr_norm_inf(at) * csr_norm_inf(a));
}

typedef struct {
  int32_t r, c;
  double v;
  int64_t idx;
} trip;

static int trip_cmp(const void* pa, const void* pb) {
  const trip* a = (const trip*)pa;
  const trip* b = (const trip*)pb;
  if (a->r != b->r) return a->r < b->r ? -1 : 1;
  if (a->c != b->c) return a->c < b->c ? -1 : 1;
  /* original position breaks ties: the reference's sort is stable */
  return a->idx < b->idx ? -1 : (a->idx > b->idx ? 1 : 0);
}

/* SparseMatrix.fromTriplets (SparseMatrix.scala:204-260): stable sort by
 * (row, col), duplicates summed in input order, exact zeros dropped.
 * Returns -1 on an out-of-range index or a non-finite value. */
static int csr_from_triplets(int32_t rows, int32_t cols, int64_t count, const int32_t* ri,
                             const int32_t* ci, const double* vi, csr* out) {
  trip* t = (trip*)xmalloc((size_t)count * sizeof(trip));
  for (int64_t k = 0; k < count; k++) {
    if (ri[k] < 0 || ri[k] >= rows || ci[k] < 0 || ci[k] >= cols || !isfinite(vi[k])) {
      free(t);
      return -1;
    }
    t[k].r = ri[k];
    t[k].c = ci[k];
    t[k].v = vi[k];
    t[k].idx = k;
  }
  qsort(t, (size_t)count, sizeof(trip), trip_cmp);
  out->rows = rows;
  out->cols = cols;
  out->ptr = (int32_t*)calloc((size_t)rows + 1, sizeof(int32_t));
  out->col = (int32_t*)xmalloc((size_t)count * sizeof(int32_t));
  out->val = (double*)xmalloc((size_t)count * sizeof(double));
  int64_t m = 0;
  for (int64_t i = 0; i < count;) {
    double v = t[i].v;
    int64_t j = i + 1;
    while (j < count && t[j].r == t[i].r && t[j].c == t[i].c) {
      v += t[j].v;
      j++;
    }
    if (v != 0.0) {
      out->ptr[t[i].r + 1]++;
      out->col[m] = t[i].c;
      out->val[m] = v;
      m++;
    }
    i = j;
  }
  for (int32_t r = 0; r < rows; r++) out->ptr[r + 1] += out->ptr[r];
  free(t);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* LpProblem (LpProblem.scala:62-104). */

static orc_lp* lp_alloc(int32_t n, int32_t m, int32_t n_eq) {
  orc_lp* lp = (orc_lp*)calloc(1, sizeof(orc_lp));
  lp->n = n;
  lp->m = m;
  lp->n_eq = n_eq;
  lp->c = dalloc(n);
  lp->l = dalloc(n);
  lp->u = dalloc(n);
  lp->q = dalloc(m);
  return lp;
}

static int lp_validate(const orc_lp* lp) {
  if (lp->n_eq < 0 || lp->n_eq > lp->m) return -1;
  for (int32_t i = 0; i < lp->n; i++) {
    if (!(lp->l[i] <= lp->u[i])) return -1;
    if (!isfinite(lp->c[i])) return -1;
  }
  for (int32_t r = 0; r < lp->m; r++)
    if (!isfinite(lp->q[r])) return -1;
  return 0;
}

void orc_lp_free(orc_lp* lp) {
  if (!lp) return;
  csr_free(&lp->a);
  csr_free(&lp->at);
  free(lp->c);
  free(lp->q);
  free(lp->l);
  free(lp->u);
  free(lp);
}

orc_lp* orc_lp_from_csr(int32_t n, int32_t m, int32_t n_eq, const int32_t* row_ptr,
                        const int32_t* col, const double* val, const double* c,
                        const double* q, const double* l, const double* u, double offset) {
  orc_lp* lp = lp_alloc(n, m, n_eq);
  int32_t nnz = row_ptr[m];
  lp->a.rows = m;
  lp->a.cols = n;
  lp->a.ptr = (int32_t*)xmalloc(((size_t)m + 1) * sizeof(int32_t));
  lp->a.col = (int32_t*)xmalloc((size_t)nnz * sizeof(int32_t));
  lp->a.val = (double*)xmalloc((size_t)nnz * sizeof(double));
  memcpy(lp->a.ptr, row_ptr, ((size_t)m + 1) * sizeof(int32_t));
  memcpy(lp->a.col, col, (size_t)nnz * sizeof(int32_t));
  memcpy(lp->a.val, val, (size_t)nnz * sizeof(double));
  memcpy(lp->c, c, (size_t)n * sizeof(double));
  memcpy(lp->l, l, (size_t)n * sizeof(double));
  memcpy(lp->u, u, (size_t)n * sizeof(double));
  memcpy(lp->q, q, (size_t)m * sizeof(double));
  lp->offset = offset;
  if (lp_validate(lp)) {
    orc_lp_free(lp);
    return NULL;
  }
  csr_transpose(&lp->a, &lp->at);
  return lp;
}

orc_lp* orc_lp_from_triplets(int32_t n, int32_t m, int32_t n_eq, int64_t count,
                             const int32_t* rows, const int32_t* cols, const double* vals,
                             const double* c, const double* q, const double* l, const double* u,
                             double offset) {
  orc_lp* lp = lp_alloc(n, m, n_eq);
  if (csr_from_triplets(m, n, count, rows, cols, vals, &lp->a)) {
    orc_lp_free(lp);
    return NULL;
  }
  memcpy(lp->c, c, (size_t)n * sizeof(double));
  memcpy(lp->l, l, (size_t)n * sizeof(double));
  memcpy(lp->u, u, (size_t)n * sizeof(double));
  memcpy(lp->q, q, (size_t)m * sizeof(double));
  lp->offset = offset;
  if (lp_validate(lp)) {
    orc_lp_free(lp);
    return NULL;
  }
  csr_transpose(&lp->a, &lp->at);
  return lp;
}

void orc_lp_dims(const orc_lp* lp, int32_t* n, int32_t* m, int32_t* n_eq, int64_t* nnz) {
  *n = lp->n;
  *m = lp->m;
  *n_eq = lp->n_eq;
  *nnz = lp->a.ptr[lp->m];
}

void orc_lp_export(const orc_lp* lp, int32_t* row_ptr, int32_t* col, double* val, double* c,
                   double* q, double* l, double* u, double* offset) {
  int32_t nnz = lp->a.ptr[lp->m];
  memcpy(row_ptr, lp->a.ptr, ((size_t)lp->m + 1) * sizeof(int32_t));
  memcpy(col, lp->a.col, (size_t)nnz * sizeof(int32_t));
  memcpy(val, lp->a.val, (size_t)nnz * sizeof(double));
  memcpy(c, lp->c, (size_t)lp->n * sizeof(double));
  memcpy(q, lp->q, (size_t)lp->m * sizeof(double));
  memcpy(l, lp->l, (size_t)lp->n * sizeof(double));
  memcpy(u, lp->u, (size_t)lp->n * sizeof(double));
  *offset = lp->offset;
}

void orc_spmv(const orc_lp* lp, int32_t transpose, const double* x, double* out) {
  csr_mul(transpose ? &lp->at : &lp->a, x, out);
}

double orc_spectral_norm_bound(const orc_lp* lp) { return csr_spectral_bound(&lp->a, &lp->at); }

/* ------------------------------------------------------------------------ */
/* LpBuilder (LpProblem.scala:154-248): two-sided rows into equalities-first
 * standard form; `<=` rows negated, ranges split into two `>=` rows. */

typedef struct {
  int32_t n;
  double *c, *l, *u;
  double offset;
  int32_t nrows, cap_rows;
  int32_t* row_start; /* into the coefficient arrays */
  int32_t* row_len;
  double *row_lo, *row_hi;
  int32_t ncoef, cap_coef;
  int32_t* coef_var;
  double* coef_val;
} builder;

static void b_init(builder* b, int32_t n) {
  memset(b, 0, sizeof(*b));
  b->n = n;
  b->c = dalloc(n);
  b->l = dalloc(n);
  b->u = dalloc(n);
  for (int32_t i = 0; i < n; i++) {
    b->l[i] = -INFINITY;
    b->u[i] = INFINITY;
  }
}
static void b_bounds(builder* b, int32_t v, double lo, double hi) {
  b->l[v] = lo;
  b->u[v] = hi;
}
static void b_row(builder* b, int32_t len, const int32_t* vars, const double* vals, double lo,
                  double hi) {
  if (b->nrows == b->cap_rows) {
    b->cap_rows = b->cap_rows ? 2 * b->cap_rows : 16;
    b->row_start = (int32_t*)realloc(b->row_start, (size_t)b->cap_rows * sizeof(int32_t));
    b->row_len = (int32_t*)realloc(b->row_len, (size_t)b->cap_rows * sizeof(int32_t));
    b->row_lo = (double*)realloc(b->row_lo, (size_t)b->cap_rows * sizeof(double));
    b->row_hi = (double*)realloc(b->row_hi, (size_t)b->cap_rows * sizeof(double));
  }
  while (b->ncoef + len > b->cap_coef) {
    b->cap_coef = b->cap_coef ? 2 * b->cap_coef : 64;
    b->coef_var = (int32_t*)realloc(b->coef_var, (size_t)b->cap_coef * sizeof(int32_t));
    b->coef_val = (double*)realloc(b->coef_val, (size_t)b->cap_coef * sizeof(double));
  }
  b->row_start[b->nrows] = b->ncoef;
  b->row_len[b->nrows] = len;
  b->row_lo[b->nrows] = lo;
  b->row_hi[b->nrows] = hi;
  b->nrows++;
  memcpy(b->coef_var + b->ncoef, vars, (size_t)len * sizeof(int32_t));
  memcpy(b->coef_val + b->ncoef, vals, (size_t)len * sizeof(double));
  b->ncoef += len;
}

typedef struct {
  int64_t count, cap;
  int32_t *r, *c;
  double* v;
  int32_t nq, capq;
  double* q;
} emitter;

static int32_t e_emit(emitter* e, const builder* b, int32_t row, double sign, double value) {
  int32_t r = e->nq;
  int32_t len = b->row_len[row];
  while (e->count + len > e->cap) {
    e->cap = e->cap ? 2 * e->cap : 64;
    e->r = (int32_t*)realloc(e->r, (size_t)e->cap * sizeof(int32_t));
    e->c = (int32_t*)realloc(e->c, (size_t)e->cap * sizeof(int32_t));
    e->v = (double*)realloc(e->v, (size_t)e->cap * sizeof(double));
  }
  for (int32_t k = 0; k < len; k++) {
    e->r[e->count] = r;
    e->c[e->count] = b->coef_var[b->row_start[row] + k];
    e->v[e->count] = sign * b->coef_val[b->row_start[row] + k];
    e->count++;
  }
  if (e->nq == e->capq) {
    e->capq = e->capq ? 2 * e->capq : 16;
    e->q = (double*)realloc(e->q, (size_t)e->capq * sizeof(double));
  }
  e->q[e->nq++] = sign * value;
  return r;
}

static orc_lp* b_build(builder* b) {
  emitter e;
  memset(&e, 0, sizeof(e));
  for (int32_t i = 0; i < b->nrows; i++)
    if (b->row_lo[i] == b->row_hi[i]) e_emit(&e, b, i, 1.0, b->row_lo[i]);
  int32_t n_eq = e.nq;
  for (int32_t i = 0; i < b->nrows; i++) {
    if (b->row_lo[i] == b->row_hi[i]) continue;
    int has_lo = !(isinf(b->row_lo[i]) && b->row_lo[i] < 0);
    int has_hi = !(isinf(b->row_hi[i]) && b->row_hi[i] > 0);
    if (has_lo && has_hi) {
      e_emit(&e, b, i, 1.0, b->row_lo[i]);
      e_emit(&e, b, i, -1.0, b->row_hi[i]);
    } else if (has_lo) {
      e_emit(&e, b, i, 1.0, b->row_lo[i]);
    } else {
      e_emit(&e, b, i, -1.0, b->row_hi[i]);
    }
  }
  orc_lp* lp = orc_lp_from_triplets(b->n, e.nq, n_eq, e.count, e.r, e.c, e.v, b->c, e.q, b->l,
                                    b->u, b->offset);
  free(e.r);
  free(e.c);
  free(e.v);
  free(e.q);
  free(b->c);
  free(b->l);
  free(b->u);
  free(b->row_start);
  free(b->row_len);
  free(b->row_lo);
  free(b->row_hi);
  free(b->coef_var);
  free(b->coef_val);
  return lp;
}

/* ------------------------------------------------------------------------ */
/* LpFixtures (LpFixtures.scala:25-165, 174-213). */

#define ROW(b, lo, hi, ...)                                                  \
  do {                                                                       \
    const double _pairs[] = {__VA_ARGS__};                                   \
    int32_t _len = (int32_t)(sizeof(_pairs) / sizeof(double) / 2);          \
    int32_t _v[16];                                                          \
    double _a[16];                                                           \
    for (int32_t _k = 0; _k < _len; _k++) {                                  \
      _v[_k] = (int32_t)_pairs[2 * _k];                                      \
      _a[_k] = _pairs[2 * _k + 1];                                           \
    }                                                                        \
    b_row(b, _len, _v, _a, lo, hi);                                          \
  } while (0)

orc_lp* orc_fixture(const char* name) {
  builder bb, *b = &bb;
  const double inf = INFINITY;
  if (!strcmp(name, "tiny-lower-bound")) {
    b_init(b, 1);
    b->c[0] = 1.0;
    b_bounds(b, 0, 0.0, 10.0);
    ROW(b, 3.0, inf, 0, 1.0);
  } else if (!strcmp(name, "product-mix")) {
    b_init(b, 2);
    b->c[0] = -3.0;
    b->c[1] = -5.0;
    b_bounds(b, 0, 0.0, inf);
    b_bounds(b, 1, 0.0, inf);
    ROW(b, -inf, 4.0, 0, 1.0);
    ROW(b, -inf, 12.0, 1, 2.0);
    ROW(b, -inf, 18.0, 0, 3.0, 1, 2.0);
  } else if (!strcmp(name, "equality-split")) {
    b_init(b, 2);
    b->c[0] = 2.0;
    b->c[1] = 3.0;
    b_bounds(b, 0, 0.0, 6.0);
    b_bounds(b, 1, 0.0, 10.0);
    ROW(b, 10.0, 10.0, 0, 1.0, 1, 1.0);
  } else if (!strcmp(name, "degenerate")) {
    b_init(b, 2);
    b->c[0] = 1.0;
    b->c[1] = 1.0;
    b_bounds(b, 0, 0.0, 5.0);
    b_bounds(b, 1, 0.0, 5.0);
    ROW(b, 2.0, inf, 0, 1.0, 1, 1.0);
    ROW(b, 0.0, 0.0, 0, 1.0, 1, -1.0);
  } else if (!strcmp(name, "free-sign")) {
    b_init(b, 2);
    b->c[0] = 1.0;
    b->c[1] = -1.0;
    b_bounds(b, 0, -1.0, 1.0);
    b_bounds(b, 1, -1.0, 1.0);
    ROW(b, 0.0, 0.0, 0, 1.0, 1, 1.0);
  } else if (!strcmp(name, "range-row")) {
    b_init(b, 2);
    b->c[0] = 1.0;
    b_bounds(b, 0, -10.0, 10.0);
    b_bounds(b, 1, 0.0, 0.0);
    ROW(b, 2.0, 5.0, 0, 1.0, 1, 1.0);
  } else if (!strcmp(name, "infeasible")) {
    b_init(b, 1);
    b->c[0] = 1.0;
    b_bounds(b, 0, -inf, inf);
    ROW(b, 1.0, inf, 0, 1.0);
    ROW(b, -inf, 0.0, 0, 1.0);
  } else if (!strcmp(name, "unbounded")) {
    b_init(b, 1);
    b->c[0] = -1.0;
    b_bounds(b, 0, 0.0, inf);
    ROW(b, 0.0, inf, 0, 1.0);
  } else if (!strcmp(name, "economic-dispatch")) {
    const double loadsA[2] = {100.0, 80.0}, loadsB[2] = {50.0, 70.0};
    const double costs[3] = {10.0, 30.0, 20.0}, caps[3] = {120.0, 100.0, 90.0};
    const double line = 30.0;
    b_init(b, 8);
    for (int t = 0; t < 2; t++) {
      int base = 4 * t;
      for (int g = 0; g < 3; g++) {
        b->c[base + g] = costs[g];
        b_bounds(b, base + g, 0.0, caps[g]);
      }
      b_bounds(b, base + 3, -line, line);
      ROW(b, loadsA[t], loadsA[t], base, 1.0, base + 1, 1.0, base + 3, -1.0);
      ROW(b, loadsB[t], loadsB[t], base + 2, 1.0, base + 3, 1.0);
    }
  } else {
    return NULL;
  }
  return b_build(b);
}

orc_lp* orc_random_feasible(int64_t seed, int32_t n, int32_t n_eq, int32_t n_ineq, double density) {
  orc_jrandom rng;
  orc_jrandom_init(&rng, seed);
  builder bb, *b = &bb;
  b_init(b, n);
  double* witness = dalloc(n);
  for (int32_t i = 0; i < n; i++) {
    double lo = -10.0 + orc_jrandom_next_double(&rng) * 5.0;
    double hi = lo + 1.0 + orc_jrandom_next_double(&rng) * 20.0;
    b_bounds(b, i, lo, hi);
    witness[i] = lo + orc_jrandom_next_double(&rng) * (hi - lo);
    b->c[i] = orc_jrandom_next_gaussian(&rng) * 5.0;
  }
  int32_t* vars = (int32_t*)xmalloc(((size_t)n + 1) * sizeof(int32_t));
  double* vals = (double*)xmalloc(((size_t)n + 1) * sizeof(double));
  for (int32_t k = 0; k < n_eq + n_ineq; k++) {
    int32_t len = 0;
    for (int32_t j = 0; j < n; j++) {
      if (orc_jrandom_next_double(&rng) < density) {
        vars[len] = j;
        vals[len] = orc_jrandom_next_gaussian(&rng) * 2.0;
        len++;
      }
    }
    if (len == 0) {
      vars[0] = orc_jrandom_next_int(&rng, n);
      vals[0] = 1.0;
      len = 1;
    }
    double value = 0.0;
    for (int32_t t = 0; t < len; t++) value += vals[t] * witness[vars[t]];
    if (k < n_eq) {
      b_row(b, len, vars, vals, value, value);
    } else {
      double lo = value - orc_jrandom_next_double(&rng) * 5.0;
      b_row(b, len, vars, vals, lo, INFINITY);
    }
  }
  free(vars);
  free(vals);
  free(witness);
  return b_build(b);
}

/* ------------------------------------------------------------------------ */
/* Scaling.apply (Scaling.scala:77-155). */

void orc_scaling(const orc_lp* lp, int32_t ruiz_iterations, int32_t apply_pc, double alpha,
                 double* row_scale, double* col_scale, double* scaled_vals) {
  const csr* a = &lp->a;
  int32_t m = a->rows, n = a->cols, nnz = a->ptr[m];
  double* vals = scaled_vals;
  memcpy(vals, a->val, (size_t)nnz * sizeof(double));
  double* rf = dalloc(m);
  double* cf = dalloc(n);
  for (int32_t r = 0; r < m; r++) row_scale[r] = 1.0;
  for (int32_t c = 0; c < n; c++) col_scale[c] = 1.0;
  for (int32_t it = 0; it < ruiz_iterations; it++) {
    for (int32_t r = 0; r < m; r++) {
      double mx = 0.0;
      for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) {
        double v = fabs(vals[p]);
        if (v > mx) mx = v;
      }
      rf[r] = mx > 0.0 ? 1.0 / sqrt(mx) : 1.0;
    }
    for (int32_t c = 0; c < n; c++) cf[c] = 0.0;
    for (int32_t p = 0; p < nnz; p++) {
      double v = fabs(vals[p]);
      if (v > cf[a->col[p]]) cf[a->col[p]] = v;
    }
    for (int32_t c = 0; c < n; c++) cf[c] = cf[c] > 0.0 ? 1.0 / sqrt(cf[c]) : 1.0;
    for (int32_t r = 0; r < m; r++)
      for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) vals[p] = vals[p] * rf[r] * cf[a->col[p]];
    for (int32_t r = 0; r < m; r++) row_scale[r] *= rf[r];
    for (int32_t c = 0; c < n; c++) col_scale[c] *= cf[c];
  }
  if (apply_pc) {
    for (int32_t r = 0; r < m; r++) {
      double s = 0.0;
      for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) s += pow(fabs(vals[p]), 2.0 - alpha);
      rf[r] = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
    }
    /* Column sums in ascending row order, the order the transpose's rows hold. */
    for (int32_t c = 0; c < n; c++) cf[c] = 0.0;
    for (int32_t r = 0; r < m; r++)
      for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) cf[a->col[p]] += pow(fabs(vals[p]), alpha);
    for (int32_t c = 0; c < n; c++) cf[c] = cf[c] > 0.0 ? 1.0 / sqrt(cf[c]) : 1.0;
    for (int32_t r = 0; r < m; r++)
      for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++) vals[p] = vals[p] * rf[r] * cf[a->col[p]];
    for (int32_t r = 0; r < m; r++) row_scale[r] *= rf[r];
    for (int32_t c = 0; c < n; c++) col_scale[c] *= cf[c];
  }
  free(rf);
  free(cf);
}

/* ------------------------------------------------------------------------ */
/* Kkt (Kkt.scala:68-189). */

static double norm2(const double* a, int32_t n) {
  double s = 0.0;
  for (int32_t i = 0; i < n; i++) s += a[i] * a[i];
  return sqrt(s);
}

static double primal_residual_sq(const orc_lp* lp, const double* ax) {
  double sum = 0.0;
  int32_t i = 0;
  for (; i < lp->n_eq; i++) {
    double r = ax[i] - lp->q[i];
    sum += r * r;
  }
  for (; i < lp->m; i++) {
    double s = lp->q[i] - ax[i];
    if (s > 0.0) sum += s * s;
  }
  return sum;
}

void orc_kkt_evaluate(const orc_lp* lp, const double* x, const double* y, const double* kx,
                      const double* kty, double* out) {
  double* ax = NULL;
  double* aty = NULL;
  if (!kx) {
    ax = dalloc(lp->m);
    csr_mul(&lp->a, x, ax);
    kx = ax;
  }
  double psq = primal_residual_sq(lp, kx);
  double pobj = lp->offset;
  for (int32_t i = 0; i < lp->n; i++) pobj += lp->c[i] * x[i];
  if (!kty) {
    aty = dalloc(lp->n);
    csr_mul(&lp->at, y, aty);
    kty = aty;
  }
  double dobj = lp->offset;
  for (int32_t i = 0; i < lp->m; i++) dobj += lp->q[i] * y[i];
  double dsq = 0.0;
  for (int32_t i = 0; i < lp->n; i++) {
    double rc = lp->c[i] - kty[i];
    if (rc > 0.0) {
      double l = lp->l[i];
      if (isinf(l) && l < 0) dsq += rc * rc;
      else dobj += l * rc;
    } else if (rc < 0.0) {
      double u = lp->u[i];
      if (isinf(u) && u > 0) dsq += rc * rc;
      else dobj += u * rc;
    }
  }
  out[0] = sqrt(psq);
  out[1] = sqrt(dsq);
  out[2] = pobj;
  out[3] = dobj;
  free(ax);
  free(aty);
}

static int kkt_finite(const double* e) {
  return isfinite(e[0]) && isfinite(e[1]) && isfinite(e[2]) && isfinite(e[3]);
}

int32_t orc_kkt_converged(const double* e, double rhs_norm, double obj_norm, double eps_abs,
                          double eps_rel) {
  double pt = eps_abs + eps_rel * rhs_norm;
  double dt = eps_abs + eps_rel * obj_norm;
  double gt = eps_abs + eps_rel * (fabs(e[2]) + fabs(e[3]));
  return kkt_finite(e) && e[0] <= pt && e[1] <= dt && fabs(e[2] - e[3]) <= gt;
}

/* KktError.weighted (Kkt.scala:28-33). */
static double kkt_weighted(const double* e, double w) {
  double gap = e[2] - e[3];
  return sqrt(w * e[0] * e[0] + (e[1] * e[1]) / w + gap * gap);
}

/* ------------------------------------------------------------------------ */
/* Certificates (Certificates.scala:60-219). */

int32_t orc_primal_infeasibility(const orc_lp* lp, const double* y, const double* kty,
                                 double* out) {
  int32_t m = lp->m, n = lp->n;
  double nsq = 0.0;
  for (int32_t i = 0; i < m; i++) nsq += y[i] * y[i];
  if (nsq <= 0.0 || !isfinite(nsq)) return 0;
  double norm = sqrt(nsq);
  double* aty = NULL;
  if (!kty) {
    aty = dalloc(n);
    csr_mul(&lp->at, y, aty);
    kty = aty;
  }
  double value = 0.0;
  for (int32_t i = 0; i < m; i++) value += lp->q[i] * y[i];
  double rsq = 0.0;
  for (int32_t i = 0; i < n; i++) {
    double rc = -kty[i];
    if (rc > 0.0) {
      double l = lp->l[i];
      if (isinf(l) && l < 0) rsq += rc * rc;
      else value += l * rc;
    } else if (rc < 0.0) {
      double u = lp->u[i];
      if (isinf(u) && u > 0) rsq += rc * rc;
      else value += u * rc;
    }
  }
  double csq = 0.0;
  for (int32_t i = lp->n_eq; i < m; i++)
    if (y[i] < 0.0) csq += y[i] * y[i];
  free(aty);
  if (!(value > 0.0) || !isfinite(value)) return 0;
  double mn = orc_spectral_norm_bound(lp);
  double rt = mn > 0.0 ? sqrt(rsq) / (mn * norm) : 0.0;
  double ct = sqrt(csq) / norm;
  double h = hypot(rt, ct);
  if (!isfinite(h)) return 0;
  *out = h;
  return 1;
}

int32_t orc_dual_infeasibility(const orc_lp* lp, const double* d, const double* kd0,
                               double* out) {
  int32_t m = lp->m, n = lp->n;
  double nsq = 0.0;
  for (int32_t i = 0; i < n; i++) nsq += d[i] * d[i];
  if (nsq <= 0.0 || !isfinite(nsq)) return 0;
  double norm = sqrt(nsq);
  double value = 0.0;
  for (int32_t i = 0; i < n; i++) value += lp->c[i] * d[i];
  if (!(value < 0.0) || !isfinite(value)) return 0;
  double* kdb = NULL;
  const double* kd = kd0;
  if (!kd) {
    kdb = dalloc(m);
    csr_mul(&lp->a, d, kdb);
    kd = kdb;
  }
  double rsq = 0.0;
  int32_t i = 0;
  for (; i < lp->n_eq; i++) rsq += kd[i] * kd[i];
  for (; i < m; i++)
    if (kd[i] < 0.0) rsq += kd[i] * kd[i];
  double bsq = 0.0;
  for (i = 0; i < n; i++) {
    double di = d[i];
    if (di > 0.0 && !(isinf(lp->u[i]) && lp->u[i] > 0)) bsq += di * di;
    else if (di < 0.0 && !(isinf(lp->l[i]) && lp->l[i] < 0)) bsq += di * di;
  }
  free(kdb);
  double mn = orc_spectral_norm_bound(lp);
  double rt = mn > 0.0 ? sqrt(rsq) / (mn * norm) : 0.0;
  double bt = sqrt(bsq) / norm;
  double h = hypot(rt, bt);
  if (!isfinite(h)) return 0;
  *out = h;
  return 1;
}

static int32_t classify_with(const orc_lp* lp, const double* x, const double* y, double tol,
                             const double* kx, const double* kty) {
  double v;
  if (orc_primal_infeasibility(lp, y, kty, &v) && v <= tol) return ORC_PRIMAL_INFEASIBLE;
  if (orc_dual_infeasibility(lp, x, kx, &v) && v <= tol) return ORC_DUAL_INFEASIBLE;
  return -1;
}

int32_t orc_classify(const orc_lp* lp, const double* x, const double* y, double tol) {
  return classify_with(lp, x, y, tol, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Kernels: ScalaKernels (kernels/ScalaKernels.scala:19-123) and
 * Float32Kernels (kernels/Float32Kernels.scala:28-158). Vectors are double
 * arrays in both; the float32 backend rounds every arithmetic result. */

typedef struct {
  int f32;
  const csr* a; /* matrices as "uploaded": values narrowed for f32 */
  const csr* at;
} kern;

static double rnd(double v) { return (double)(float)v; }

static void k_upload(const kern* k, const double* src, double* dst, int32_t n) {
  if (!k->f32) {
    memcpy(dst, src, (size_t)n * sizeof(double));
    return;
  }
  for (int32_t i = 0; i < n; i++) dst[i] = isinf(src[i]) ? src[i] : rnd(src[i]);
}

static void k_spmv(const kern* k, const csr* a, const double* x, double* out) {
  if (!k->f32) {
    csr_mul(a, x, out);
    return;
  }
  for (int32_t r = 0; r < a->rows; r++) {
    float acc = 0.0f;
    for (int32_t p = a->ptr[r]; p < a->ptr[r + 1]; p++)
      acc = acc + ((float)a->val[p] * (float)x[a->col[p]]);
    out[r] = (double)acc;
  }
}

static void k_axpby(const kern* k, double alpha, const double* x, double beta, const double* y,
                    double* out, int32_t n) {
  if (!k->f32) {
    for (int32_t i = 0; i < n; i++) out[i] = alpha * x[i] + beta * y[i];
    return;
  }
  float a = (float)alpha, b = (float)beta;
  for (int32_t i = 0; i < n; i++) out[i] = (double)((a * (float)x[i]) + (b * (float)y[i]));
}

static void k_scale(const kern* k, double alpha, const double* x, double* out, int32_t n) {
  if (!k->f32) {
    for (int32_t i = 0; i < n; i++) out[i] = alpha * x[i];
    return;
  }
  float a = (float)alpha;
  for (int32_t i = 0; i < n; i++) out[i] = (double)(a * (float)x[i]);
}

static double k_dot(const kern* k, const double* x, const double* y, int32_t n) {
  if (!k->f32) {
    double s = 0.0;
    for (int32_t i = 0; i < n; i++) s += x[i] * y[i];
    return s;
  }
  float s = 0.0f;
  for (int32_t i = 0; i < n; i++) s = s + ((float)x[i] * (float)y[i]);
  return (double)s;
}

static double k_sqnorm(const kern* k, const double* x, int32_t n) {
  if (!k->f32) {
    double s = 0.0;
    for (int32_t i = 0; i < n; i++) s += x[i] * x[i];
    return s;
  }
  float s = 0.0f;
  for (int32_t i = 0; i < n; i++) {
    float v = (float)x[i];
    s = s + (v * v);
  }
  return (double)s;
}

static void k_primal(const kern* k, const double* x, const double* kty, const double* c,
                     const double* l, const double* u, double tau, double* out, int32_t n) {
  if (!k->f32) {
    for (int32_t i = 0; i < n; i++) {
      double s = x[i] - tau * (c[i] - kty[i]);
      out[i] = s < l[i] ? l[i] : (s > u[i] ? u[i] : s);
    }
    return;
  }
  float t = (float)tau;
  for (int32_t i = 0; i < n; i++) {
    double s = (double)((float)x[i] - (t * ((float)c[i] - (float)kty[i])));
    out[i] = s < l[i] ? l[i] : (s > u[i] ? u[i] : s);
  }
}

static void k_dual(const kern* k, const double* y, const double* kxb, const double* q,
                   double sigma, int32_t n_eq, double* out, int32_t m) {
  int32_t i = 0;
  if (!k->f32) {
    for (; i < n_eq; i++) out[i] = y[i] + sigma * (q[i] - kxb[i]);
    for (; i < m; i++) {
      double s = y[i] + sigma * (q[i] - kxb[i]);
      out[i] = s > 0.0 ? s : 0.0;
    }
    return;
  }
  float s = (float)sigma;
  for (; i < n_eq; i++) out[i] = (double)((float)y[i] + (s * ((float)q[i] - (float)kxb[i])));
  for (; i < m; i++) {
    double st = (double)((float)y[i] + (s * ((float)q[i] - (float)kxb[i])));
    out[i] = st > 0.0 ? st : 0.0;
  }
}

/* ------------------------------------------------------------------------ */
/* Pdhg.Solve (Pdhg.scala:123-534). */

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

void orc_default_params(orc_params* p) {
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
}

void orc_default_mixed_params(orc_mixed_params* p) {
  p->reduced_tolerance = 1e-5;
  p->reduced_max_iterations = 20000;
  p->always_refine = 0;
  p->reduced_time_share = 0.5;
}

/* PdhgParams requirements (PdhgParams.scala:61-86). */
static int params_valid(const orc_params* p) {
  if (!(p->eps_abs >= 0.0) || !(p->eps_rel >= 0.0)) return 0;
  if (p->max_iterations <= 0 || p->evaluation_frequency <= 0) return 0;
  if (!(p->primal_weight_smoothing >= 0.0 && p->primal_weight_smoothing <= 1.0)) return 0;
  if (!(p->restart_sufficient > 0.0 && p->restart_sufficient <= p->restart_necessary)) return 0;
  if (!(p->restart_artificial > 0.0 && p->restart_artificial <= 1.0)) return 0;
  if (!(p->infeasibility_tolerance >= 0.0)) return 0;
  if (p->ruiz_iterations < 0) return 0;
  if (!(p->pock_chambolle_alpha >= 0.0 && p->pock_chambolle_alpha <= 2.0)) return 0;
  return 1;
}

typedef struct {
  const orc_lp* lp;
  const orc_params* p;
  kern k;
  csr sa, sat; /* scaled (and, for f32, narrowed) matrix and transpose */
  double *row_scale, *col_scale;
  double rhs_norm, obj_norm;
  double *cost, *rhs, *lower, *upper;
  double *x, *y, *kx, *kty, *xn, *yn, *xbar, *kxbar, *dx, *dy, *kdx;
  double *sx, *sy, *skx, *skty, *ax, *ay, *akx, *akty, *xr, *yr, *zn, *zm;
  double *hx, *hy, *hkx, *hkty;
  double step, weight, wsum;
  int32_t it, restarts, last_restart_it;
  double err_at_restart, last_candidate;
  double t0;
} solve_t;

static void s_evaluate(solve_t* s, const double* xv, const double* yv, const double* kxv,
                       const double* ktyv, double* err) {
  int32_t n = s->lp->n, m = s->lp->m;
  memcpy(s->hx, xv, (size_t)n * sizeof(double));
  memcpy(s->hy, yv, (size_t)m * sizeof(double));
  memcpy(s->hkx, kxv, (size_t)m * sizeof(double));
  memcpy(s->hkty, ktyv, (size_t)n * sizeof(double));
  for (int32_t i = 0; i < n; i++) {
    s->hx[i] *= s->col_scale[i];
    s->hkty[i] /= s->col_scale[i];
  }
  for (int32_t r = 0; r < m; r++) {
    s->hy[r] *= s->row_scale[r];
    s->hkx[r] /= s->row_scale[r];
  }
  orc_kkt_evaluate(s->lp, s->hx, s->hy, s->hkx, s->hkty, err);
}

static int s_converged(const solve_t* s, const double* e) {
  return orc_kkt_converged(e, s->rhs_norm, s->obj_norm, s->p->eps_abs, s->p->eps_rel);
}

static void s_reset_averages(solve_t* s) {
  int32_t n = s->lp->n, m = s->lp->m;
  s->wsum = 0.0;
  memcpy(s->sx, s->zn, (size_t)n * sizeof(double));
  memcpy(s->skty, s->zn, (size_t)n * sizeof(double));
  memcpy(s->sy, s->zm, (size_t)m * sizeof(double));
  memcpy(s->skx, s->zm, (size_t)m * sizeof(double));
}

static double s_step(solve_t* s) {
  const kern* k = &s->k;
  int32_t n = s->lp->n, m = s->lp->m;
  double movement = 0.0;
  int accepted = 0, trials = 0;
  while (!accepted) {
    trials++;
    double tau = s->step / s->weight;
    double sigma = s->step * s->weight;
    k_primal(k, s->x, s->kty, s->cost, s->lower, s->upper, tau, s->xn, n);
    k_axpby(k, 2.0, s->xn, -1.0, s->x, s->xbar, n);
    k_spmv(k, &s->sa, s->xbar, s->kxbar);
    k_dual(k, s->y, s->kxbar, s->rhs, sigma, s->lp->n_eq, s->yn, m);
    k_axpby(k, 1.0, s->xn, -1.0, s->x, s->dx, n);
    k_axpby(k, 1.0, s->yn, -1.0, s->y, s->dy, m);
    k_axpby(k, 0.5, s->kxbar, -0.5, s->kx, s->kdx, m);
    double dxsq = k_sqnorm(k, s->dx, n);
    double dysq = k_sqnorm(k, s->dy, m);
    movement = 0.5 * (s->weight * dxsq + dysq / s->weight);
    if (movement == 0.0) {
      accepted = 1;
    } else {
      double inter = fabs(k_dot(k, s->dy, s->kdx, m));
      double limit = inter > 0.0 ? movement / inter : INFINITY;
      double t = (double)(s->it + 1);
      double by_rule = (1.0 - pow(t, -0.3)) * limit;
      double by_cap = (1.0 + pow(t, -0.6)) * s->step;
      /* math.min propagates NaN (0 * inf at the first iteration) */
      double proposed = (isnan(by_rule) || isnan(by_cap)) ? NAN : (by_rule < by_cap ? by_rule : by_cap);
      if (s->step <= limit || trials >= 60) {
        k_axpby(k, 0.5, s->kxbar, 0.5, s->kx, s->kx, m);
        memcpy(s->x, s->xn, (size_t)n * sizeof(double));
        memcpy(s->y, s->yn, (size_t)m * sizeof(double));
        k_spmv(k, &s->sat, s->y, s->kty);
        double w = s->step;
        s->wsum += w;
        k_axpby(k, w, s->x, 1.0, s->sx, s->sx, n);
        k_axpby(k, w, s->y, 1.0, s->sy, s->sy, m);
        k_axpby(k, w, s->kx, 1.0, s->skx, s->skx, m);
        k_axpby(k, w, s->kty, 1.0, s->skty, s->skty, n);
        accepted = 1;
      }
      if (isfinite(proposed) && proposed > 0.0) s->step = proposed;
    }
  }
  return movement;
}

static int s_classify(solve_t* s) {
  return classify_with(s->lp, s->hx, s->hy, s->p->infeasibility_tolerance, s->hkx, s->hkty);
}

static void s_restart_at(solve_t* s, int use_avg, double* err) {
  const kern* k = &s->k;
  int32_t n = s->lp->n, m = s->lp->m;
  if (use_avg) {
    memcpy(s->x, s->ax, (size_t)n * sizeof(double));
    memcpy(s->y, s->ay, (size_t)m * sizeof(double));
    memcpy(s->kx, s->akx, (size_t)m * sizeof(double));
    memcpy(s->kty, s->akty, (size_t)n * sizeof(double));
  }
  k_axpby(k, 1.0, s->x, -1.0, s->xr, s->dx, n);
  k_axpby(k, 1.0, s->y, -1.0, s->yr, s->dy, m);
  double pm = sqrt(k_sqnorm(k, s->dx, n));
  double dm = sqrt(k_sqnorm(k, s->dy, m));
  if (pm > 0.0 && dm > 0.0) {
    double th = s->p->primal_weight_smoothing;
    double next = exp(th * log(dm / pm) + (1.0 - th) * log(s->weight));
    if (isfinite(next) && next > 0.0) s->weight = next;
  }
  memcpy(s->xr, s->x, (size_t)n * sizeof(double));
  memcpy(s->yr, s->y, (size_t)m * sizeof(double));
  s_reset_averages(s);
  s->restarts++;
  s->last_restart_it = s->it;
  s_evaluate(s, s->x, s->y, s->kx, s->kty, err);
  s->err_at_restart = kkt_weighted(err, s->weight);
  s->last_candidate = INFINITY;
}

/* Returns the terminal status, or -1 to keep going. On a terminal status
 * err holds the KktError of the iterate left in the host mirrors. */
static int s_checkpoint(solve_t* s, double* err) {
  const kern* k = &s->k;
  int32_t n = s->lp->n, m = s->lp->m;
  k_spmv(k, &s->sa, s->x, s->kx);
  double cur[4];
  s_evaluate(s, s->x, s->y, s->kx, s->kty, cur);
  if (s_converged(s, cur)) {
    memcpy(err, cur, sizeof(cur));
    return ORC_OPTIMAL;
  }
  int st = s_classify(s);
  if (st >= 0) {
    memcpy(err, cur, sizeof(cur));
    return st;
  }
  double cur_err = kkt_weighted(cur, s->weight);
  double avg[4];
  int have_avg = s->wsum > 0.0;
  double avg_err = INFINITY;
  if (have_avg) {
    double inv = 1.0 / s->wsum;
    k_scale(k, inv, s->sx, s->ax, n);
    k_scale(k, inv, s->sy, s->ay, m);
    k_scale(k, inv, s->skx, s->akx, m);
    k_scale(k, inv, s->skty, s->akty, n);
    s_evaluate(s, s->ax, s->ay, s->akx, s->akty, avg);
    if (s_converged(s, avg)) {
      memcpy(err, avg, sizeof(avg));
      return ORC_OPTIMAL;
    }
    avg_err = kkt_weighted(avg, s->weight);
  }
  int use_avg = avg_err < cur_err;
  /* math.min propagates NaN */
  double cand = (isnan(cur_err) || isnan(avg_err)) ? NAN : (cur_err < avg_err ? cur_err : avg_err);
  const orc_params* p = s->p;
  int sufficient = cand <= p->restart_sufficient * s->err_at_restart;
  int necessary = cand <= p->restart_necessary * s->err_at_restart && cand > s->last_candidate;
  int artificial = (double)(s->it - s->last_restart_it) >= p->restart_artificial * (double)s->it;
  if (sufficient || necessary || artificial) {
    double e[4];
    s_restart_at(s, use_avg, e);
  } else {
    s->last_candidate = cand;
  }
  return -1;
}

static int32_t pdhg_run(const orc_lp* lp, const orc_params* p, int f32, const double* warm_x,
                        const double* warm_y, double warm_step, double warm_weight,
                        orc_result* res, double* x_out, double* y_out, double* rc_out) {
  if (!params_valid(p)) return -1;
  int32_t n = lp->n, m = lp->m, nnz = lp->a.ptr[m];
  if (warm_x)
    for (int32_t i = 0; i < n; i++)
      if (!isfinite(warm_x[i])) return -1;
  if (warm_y)
    for (int32_t r = 0; r < m; r++)
      if (!isfinite(warm_y[r])) return -1;

  solve_t S, *s = &S;
  memset(s, 0, sizeof(S));
  s->lp = lp;
  s->p = p;
  s->k.f32 = f32;

  /* Scaling (Pdhg.scala:135) and the scaled matrix plus its transpose,
   * "uploaded" (narrowed for f32, Float32Kernels.scala:53-60). */
  s->row_scale = dalloc(m);
  s->col_scale = dalloc(n);
  csr scaled = lp->a;
  scaled.val = dalloc(nnz);
  orc_scaling(lp, p->ruiz_iterations, p->apply_pock_chambolle, p->pock_chambolle_alpha,
              s->row_scale, s->col_scale, scaled.val);
  csr scaled_t;
  csr_transpose(&scaled, &scaled_t);
  double bound = csr_spectral_bound(&scaled, &scaled_t);
  double* sc = dalloc(n);
  double* sl = dalloc(n);
  double* su = dalloc(n);
  double* sq = dalloc(m);
  for (int32_t i = 0; i < n; i++) {
    double d = s->col_scale[i];
    sc[i] = lp->c[i] * d;
    sl[i] = lp->l[i] / d;
    su[i] = lp->u[i] / d;
  }
  for (int32_t r = 0; r < m; r++) sq[r] = lp->q[r] * s->row_scale[r];

  s->rhs_norm = norm2(lp->q, m);
  s->obj_norm = norm2(lp->c, n);

  s->sa = scaled;
  s->sa.val = dalloc(nnz);
  s->sat = scaled_t;
  s->sat.val = dalloc(nnz);
  if (f32) {
    for (int32_t k2 = 0; k2 < nnz; k2++) {
      s->sa.val[k2] = rnd(scaled.val[k2]);
      s->sat.val[k2] = rnd(scaled_t.val[k2]);
    }
  } else {
    memcpy(s->sa.val, scaled.val, (size_t)nnz * sizeof(double));
    memcpy(s->sat.val, scaled_t.val, (size_t)nnz * sizeof(double));
  }
  s->cost = dalloc(n);
  s->lower = dalloc(n);
  s->upper = dalloc(n);
  s->rhs = dalloc(m);
  k_upload(&s->k, sc, s->cost, n);
  k_upload(&s->k, sq, s->rhs, m);
  k_upload(&s->k, sl, s->lower, n);
  k_upload(&s->k, su, s->upper, n);

  double** nvecs[] = {&s->x, &s->kty, &s->xn, &s->xbar, &s->dx, &s->sx, &s->skty, &s->ax,
                      &s->akty, &s->xr, &s->zn, &s->hx, &s->hkty};
  double** mvecs[] = {&s->y, &s->kx, &s->yn, &s->kxbar, &s->dy, &s->kdx, &s->sy, &s->skx,
                      &s->ay, &s->akx, &s->yr, &s->zm, &s->hy, &s->hkx};
  for (size_t i = 0; i < sizeof(nvecs) / sizeof(nvecs[0]); i++) *nvecs[i] = dalloc(n);
  for (size_t i = 0; i < sizeof(mvecs) / sizeof(mvecs[0]); i++) *mvecs[i] = dalloc(m);
  s->step = 0.0;
  s->weight = 1.0;
  s->err_at_restart = INFINITY;
  s->last_candidate = INFINITY;
  s->t0 = now_ms();

  /* initialise (Pdhg.scala:250-309) */
  s->step = (warm_step > 0.0 && isfinite(warm_step)) ? warm_step : (bound > 0.0 ? 1.0 / bound : 1.0);
  double cn = norm2(sc, n), qn = norm2(sq, m);
  s->weight = (warm_weight > 0.0 && isfinite(warm_weight))
                  ? warm_weight
                  : ((cn > 0.0 && qn > 0.0) ? cn / qn : 1.0);
  double* start = dalloc(n);
  if (warm_x)
    for (int32_t i = 0; i < n; i++) start[i] = warm_x[i] / s->col_scale[i];
  for (int32_t i = 0; i < n; i++) {
    double l = sl[i], u = su[i], v = start[i];
    start[i] = v < l ? l : (v > u ? u : v);
  }
  k_upload(&s->k, start, s->x, n);
  if (warm_y) {
    double* sd = dalloc(m);
    for (int32_t r = 0; r < m; r++) {
      double v = warm_y[r] / s->row_scale[r];
      sd[r] = (r >= lp->n_eq && v < 0.0) ? 0.0 : v;
    }
    k_upload(&s->k, sd, s->y, m);
    free(sd);
  }
  free(start);
  k_spmv(&s->k, &s->sa, s->x, s->kx);
  k_spmv(&s->k, &s->sat, s->y, s->kty);
  memcpy(s->xr, s->x, (size_t)n * sizeof(double));
  memcpy(s->yr, s->y, (size_t)m * sizeof(double));
  s_reset_averages(s);

  /* run (Pdhg.scala:206-246) */
  double err[4];
  int status = -1;
  s_evaluate(s, s->x, s->y, s->kx, s->kty, err);
  if (s_converged(s, err)) {
    status = ORC_OPTIMAL;
  } else {
    s->err_at_restart = kkt_weighted(err, s->weight);
    while (status < 0) {
      double movement = s_step(s);
      s->it++;
      if (movement == 0.0) {
        k_spmv(&s->k, &s->sa, s->x, s->kx);
        s_evaluate(s, s->x, s->y, s->kx, s->kty, err);
        status = s_converged(s, err) ? ORC_OPTIMAL : ORC_NUMERICAL_ERROR;
      } else if (s->it % p->evaluation_frequency == 0) {
        status = s_checkpoint(s, err);
        if (status < 0 && p->time_limit_ms >= 0 && now_ms() - s->t0 >= (double)p->time_limit_ms) {
          s_evaluate(s, s->x, s->y, s->kx, s->kty, err);
          status = ORC_TIME_LIMIT;
        }
      }
      if (status < 0 && s->it >= p->max_iterations) {
        s_evaluate(s, s->x, s->y, s->kx, s->kty, err);
        status = ORC_ITERATION_LIMIT;
      }
    }
  }

  /* finish (Pdhg.scala:510-532) */
  res->status = status;
  res->iterations = s->it;
  res->restarts = s->restarts;
  res->objective = err[2];
  res->dual_objective = err[3];
  res->primal_residual = err[0];
  res->dual_residual = err[1];
  res->solve_time_ms = now_ms() - s->t0;
  res->final_step_size = s->step;
  res->final_primal_weight = s->weight;
  if (x_out) memcpy(x_out, s->hx, (size_t)n * sizeof(double));
  if (y_out) memcpy(y_out, s->hy, (size_t)m * sizeof(double));
  if (rc_out)
    for (int32_t i = 0; i < n; i++) rc_out[i] = lp->c[i] - s->hkty[i];

  for (size_t i = 0; i < sizeof(nvecs) / sizeof(nvecs[0]); i++) free(*nvecs[i]);
  for (size_t i = 0; i < sizeof(mvecs) / sizeof(mvecs[0]); i++) free(*mvecs[i]);
  free(s->cost);
  free(s->lower);
  free(s->upper);
  free(s->rhs);
  free(s->sa.val);
  free(s->sat.val);
  free(scaled.val);
  csr_free(&scaled_t);
  free(sc);
  free(sl);
  free(su);
  free(sq);
  free(s->row_scale);
  free(s->col_scale);
  return 0;
}

int32_t orc_pdhg_solve(const orc_lp* lp, const orc_params* params, int32_t precision,
                       const double* warm_x, const double* warm_y, double warm_step,
                       double warm_weight, orc_result* res, double* x_out, double* y_out,
                       double* rc_out) {
  if (precision != 64 && precision != 32) return -1;
  return pdhg_run(lp, params, precision == 32, warm_x, warm_y, warm_step, warm_weight, res, x_out,
                  y_out, rc_out);
}

/* ------------------------------------------------------------------------ */
/* MixedPrecision.solve (MixedPrecision.scala:104-199). */

int32_t orc_mixed_solve(const orc_lp* lp, const orc_params* params, int32_t device_precision,
                        const orc_mixed_params* mixed, orc_result* res, orc_mixed_info* info,
                        double* x_out, double* y_out, double* rc_out) {
  int32_t n = lp->n, m = lp->m;
  memset(info, 0, sizeof(*info));
  if (!(mixed->reduced_tolerance > 0.0) || mixed->reduced_max_iterations <= 0 ||
      !(mixed->reduced_time_share > 0.0 && mixed->reduced_time_share < 1.0))
    return -1;
  if (device_precision == 64) {
    int32_t rc = orc_pdhg_solve(lp, params, 64, NULL, NULL, NAN, NAN, res, x_out, y_out, rc_out);
    info->refinement_iterations = res->iterations;
    info->refinement_ms = res->solve_time_ms;
    return rc;
  }
  orc_params rp = *params;
  rp.eps_abs = fmax(params->eps_abs, mixed->reduced_tolerance);
  rp.eps_rel = fmax(params->eps_rel, mixed->reduced_tolerance);
  rp.max_iterations =
      params->max_iterations < mixed->reduced_max_iterations ? params->max_iterations
                                                             : mixed->reduced_max_iterations;
  if (params->time_limit_ms >= 0) {
    int64_t t = (int64_t)((double)params->time_limit_ms * mixed->reduced_time_share);
    rp.time_limit_ms = t > 1 ? t : 1;
  }
  rp.infeasibility_tolerance = fmax(params->infeasibility_tolerance, mixed->reduced_tolerance);
  double* rx = dalloc(n);
  double* ry = dalloc(m);
  double* rrc = dalloc(n);
  orc_result red;
  int32_t rc = orc_pdhg_solve(lp, &rp, 32, NULL, NULL, NAN, NAN, &red, rx, ry, rrc);
  if (rc) {
    free(rx);
    free(ry);
    free(rrc);
    return rc;
  }
  info->reduced_iterations = red.iterations;
  info->reduced_restarts = red.restarts;
  info->reduced_ms = red.solve_time_ms;
  int conclusive = red.status == ORC_OPTIMAL || red.status == ORC_PRIMAL_INFEASIBLE ||
                   red.status == ORC_DUAL_INFEASIBLE;
  int holds = conclusive && red.status != ORC_OPTIMAL &&
              orc_classify(lp, rx, ry, params->infeasibility_tolerance) == red.status;
  if (!mixed->always_refine && holds) {
    *res = red;
    if (x_out) memcpy(x_out, rx, (size_t)n * sizeof(double));
    if (y_out) memcpy(y_out, ry, (size_t)m * sizeof(double));
    if (rc_out) memcpy(rc_out, rrc, (size_t)n * sizeof(double));
    free(rx);
    free(ry);
    free(rrc);
    return 0;
  }
  orc_params fp = *params;
  if (params->time_limit_ms >= 0) {
    int64_t t = params->time_limit_ms - (int64_t)red.solve_time_ms;
    fp.time_limit_ms = t > 1 ? t : 1;
  }
  int usable = 1;
  for (int32_t i = 0; i < n; i++) usable &= isfinite(rx[i]) != 0;
  for (int32_t r = 0; r < m; r++) usable &= isfinite(ry[r]) != 0;
  double ws = (usable && isfinite(red.final_step_size) && red.final_step_size > 0.0)
                  ? red.final_step_size
                  : NAN;
  double ww = (usable && isfinite(red.final_primal_weight) && red.final_primal_weight > 0.0)
                  ? red.final_primal_weight
                  : NAN;
  rc = orc_pdhg_solve(lp, &fp, 64, usable ? rx : NULL, usable ? ry : NULL, ws, ww, res, x_out,
                      y_out, rc_out);
  info->refinement_iterations = res->iterations;
  info->refinement_ms = res->solve_time_ms;
  info->refined = 1;
  free(rx);
  free(ry);
  free(rrc);
  return rc;
}
