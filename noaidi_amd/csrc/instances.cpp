// Seeded synthetic LP generators for the benchmark and parity workloads.
// Standard form as LpProblem (LpProblem.scala:22-104): equalities first,
// K x >= q after, canonical CSR. Randomness is SplitMix64 with a hand-written
// Box-Muller transform so an instance depends only on its seed.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

extern "C" {
typedef struct {
  int32_t n, m, n_eq;
  int32_t pad;
  int64_t nnz;
  int32_t* row_ptr;
  int32_t* col;
  double* val;
  double* c;
  double* q;
  double* l;
  double* u;
  double offset;
  double known_objective;  // NaN when no closed form exists
  double* x_star;          // known primal optimum (or NULL)
  double* y_star;          // known dual optimum (or NULL)
  int32_t n_int;           // integer columns (0 for a pure LP)
  int32_t pad2;
  int32_t* integers;       // ascending column indices (or NULL)
} prima_instance;
}

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0xD1B54A32D192ED03ull) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  int64_t below(int64_t k) { return (int64_t)(uniform() * (double)k); }
  bool has_spare = false;
  double spare = 0;
  double normal() {  // Box-Muller
    if (has_spare) {
      has_spare = false;
      return spare;
    }
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2.0 * M_PI * u2);
    has_spare = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
};

struct Trip {
  int32_t r, c;
  double v;
};

// Ring of N buses plus `chords` distinct random chords; reactance U[0.01,0.1],
// capacity U[100,1000] per line, drawn in that order.
struct Grid {
  std::vector<std::pair<int32_t, int32_t>> lines;
  std::vector<double> x, cap;
};

Grid make_grid(Rng& rng, int32_t N, int32_t chords) {
  Grid g;
  for (int32_t b = 0; b < N; ++b) g.lines.push_back({b, (b + 1) % N});
  std::vector<uint64_t> keys;
  for (auto& e : g.lines)
    keys.push_back((uint64_t)std::min(e.first, e.second) * N + std::max(e.first, e.second));
  std::sort(keys.begin(), keys.end());
  int32_t added = 0, guard = 0;
  while (added < chords && guard < chords * 100) {
    ++guard;
    int32_t a = (int32_t)rng.below(N), b = (int32_t)rng.below(N);
    if (a == b) continue;
    const uint64_t k = (uint64_t)std::min(a, b) * N + std::max(a, b);
    if (std::binary_search(keys.begin(), keys.end(), k)) continue;
    keys.insert(std::lower_bound(keys.begin(), keys.end(), k), k);
    g.lines.push_back({a, b});
    ++added;
  }
  const size_t L = g.lines.size();
  g.x.resize(L);
  g.cap.resize(L);
  for (size_t l = 0; l < L; ++l) {
    g.x[l] = rng.uniform(0.01, 0.1);
    g.cap[l] = rng.uniform(100.0, 1000.0);
  }
  return g;
}

prima_instance* alloc_instance(int32_t n, int32_t m, int64_t nnz) {
  prima_instance* I = (prima_instance*)std::calloc(1, sizeof(prima_instance));
  I->n = n;
  I->m = m;
  I->nnz = nnz;
  I->row_ptr = (int32_t*)std::calloc((size_t)m + 1, sizeof(int32_t));
  I->col = (int32_t*)std::malloc(std::max<size_t>(1, (size_t)nnz) * sizeof(int32_t));
  I->val = (double*)std::malloc(std::max<size_t>(1, (size_t)nnz) * sizeof(double));
  I->c = (double*)std::calloc(std::max(1, n), sizeof(double));
  I->l = (double*)std::calloc(std::max(1, n), sizeof(double));
  I->u = (double*)std::calloc(std::max(1, n), sizeof(double));
  I->q = (double*)std::calloc(std::max(1, m), sizeof(double));
  I->known_objective = std::numeric_limits<double>::quiet_NaN();
  return I;
}

// Canonical CSR from triplets (rows sorted, columns ascending, duplicates
// summed, zeros dropped), as SparseMatrix.fromTriplets.
prima_instance* from_triplets(int32_t n, int32_t m, std::vector<Trip>& t) {
  std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
    return a.r != b.r ? a.r < b.r : a.c < b.c;
  });
  std::vector<Trip> merged;
  merged.reserve(t.size());
  for (size_t i = 0; i < t.size();) {
    size_t j = i + 1;
    double v = t[i].v;
    while (j < t.size() && t[j].r == t[i].r && t[j].c == t[i].c) v += t[j++].v;
    if (v != 0.0) merged.push_back({t[i].r, t[i].c, v});
    i = j;
  }
  prima_instance* I = alloc_instance(n, m, (int64_t)merged.size());
  for (size_t k = 0; k < merged.size(); ++k) {
    I->row_ptr[merged[k].r + 1]++;
    I->col[k] = merged[k].c;
    I->val[k] = merged[k].v;
  }
  for (int32_t r = 0; r < m; ++r) I->row_ptr[r + 1] += I->row_ptr[r];
  return I;
}

}  // namespace

extern "C" {

void prima_instance_free(prima_instance* I) {
  if (!I) return;
  std::free(I->row_ptr);
  std::free(I->col);
  std::free(I->val);
  std::free(I->c);
  std::free(I->q);
  std::free(I->l);
  std::free(I->u);
  std::free(I->x_star);
  std::free(I->y_star);
  std::free(I->integers);
  std::free(I);
}

// Random equality-form LP built around a known optimum:
//  - row lengths: fixed_len if > 0, else round(exp(mu + sigma N(0,1)))
//    clamped to [1, max_len], columns drawn uniformly without replacement,
//    values N(0,1);
//  - x* ~ U[0,1] with exactly round(zero_frac n) entries set to 0, b = A x*;
//  - y* ~ N(0,1), s ~ U[0.1,1] on the zeroed entries (0 elsewhere),
//    c = A'y* + s; bounds 0 <= x <= 10.
// (x*, y*) satisfies the KKT conditions, so obj* = c'x* = b'y*.
prima_instance* prima_gen_random_known(uint64_t seed, int32_t n, int32_t m, int32_t fixed_len,
                                       double mu, double sigma, int32_t max_len,
                                       double zero_frac) {
  if (n <= 0 || m < 0) return nullptr;
  Rng rng(seed);
  std::vector<int32_t> len((size_t)m);
  int64_t nnz = 0;
  for (int32_t r = 0; r < m; ++r) {
    int64_t L = fixed_len > 0 ? fixed_len : (int64_t)std::llround(std::exp(mu + sigma * rng.normal()));
    L = std::max<int64_t>(1, std::min<int64_t>(L, std::min(max_len, n)));
    len[r] = (int32_t)L;
    nnz += L;
  }
  prima_instance* I = alloc_instance(n, m, nnz);
  std::vector<int32_t> pick;
  std::vector<uint8_t> mark((size_t)n, 0);
  int64_t p = 0;
  for (int32_t r = 0; r < m; ++r) {
    pick.clear();
    while ((int32_t)pick.size() < len[r]) {
      const int32_t c = (int32_t)rng.below(n);
      if (!mark[c]) {
        mark[c] = 1;
        pick.push_back(c);
      }
    }
    std::sort(pick.begin(), pick.end());
    for (int32_t c : pick) {
      mark[c] = 0;
      I->col[p] = c;
      double v = rng.normal();
      while (v == 0.0) v = rng.normal();
      I->val[p] = v;
      ++p;
    }
    I->row_ptr[r + 1] = (int32_t)p;
  }
  I->n_eq = m;
  std::vector<double> xs((size_t)n), ys((size_t)m), s((size_t)n, 0.0);
  for (int32_t i = 0; i < n; ++i) xs[i] = rng.uniform();
  const int32_t nz = (int32_t)std::llround(zero_frac * n);
  std::vector<int32_t> perm((size_t)n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int32_t k = 0; k < nz; ++k) {  // partial Fisher-Yates
    const int32_t j = k + (int32_t)rng.below(n - k);
    std::swap(perm[k], perm[j]);
    xs[perm[k]] = 0.0;
    s[perm[k]] = rng.uniform(0.1, 1.0);
  }
  for (int32_t r = 0; r < m; ++r) ys[r] = rng.normal();
  for (int32_t r = 0; r < m; ++r) {
    double b = 0.0;
    for (int32_t q = I->row_ptr[r]; q < I->row_ptr[r + 1]; ++q) b += I->val[q] * xs[I->col[q]];
    I->q[r] = b;
  }
  std::vector<double> aty((size_t)n, 0.0);
  for (int32_t r = 0; r < m; ++r)
    for (int32_t q = I->row_ptr[r]; q < I->row_ptr[r + 1]; ++q)
      aty[I->col[q]] += I->val[q] * ys[r];
  long double obj = 0.0L;
  for (int32_t i = 0; i < n; ++i) {
    I->c[i] = aty[i] + s[i];
    I->l[i] = 0.0;
    I->u[i] = 10.0;
    obj += (long double)I->c[i] * (long double)xs[i];
  }
  I->known_objective = (double)obj;
  I->x_star = (double*)std::malloc((size_t)n * sizeof(double));
  I->y_star = (double*)std::malloc(std::max<size_t>(1, (size_t)m) * sizeof(double));
  std::memcpy(I->x_star, xs.data(), (size_t)n * sizeof(double));
  if (m) std::memcpy(I->y_star, ys.data(), (size_t)m * sizeof(double));
  return I;
}

// PyPSA-shaped multi-period economic dispatch with DC power flow in angle
// form. Per snapshot t the variables are, in order:
//   generator output p (gens_per_bus * buses), line flow f (lines),
//   bus angle theta (buses; theta of bus 0 fixed to 0 by its bounds),
//   storage charge, discharge and state of charge (storage each),
//   load shedding (buses).
// Rows (equalities first): nodal balance, flow definition
// f = (theta_from - theta_to) / x, storage energy balance with efficiency
// eta on both directions; then the storage power limit ch + dis <= p_nom.
// Lines: a ring plus `chords` distinct random chords, reactance U[0.01,0.1],
// capacity U[100,1000]. Generators: cost U[10,100], p_nom U[50,500].
// Load: bus base U[50,250] x diurnal sinusoid x U[0.8,1.2] per bus, times a
// per-bus scenario factor U[1-perturb, 1+perturb]. Shedding costs 3000.
prima_instance* prima_gen_dispatch(uint64_t seed, int32_t buses, int32_t chords,
                                   int32_t snapshots, int32_t gens_per_bus, int32_t storage,
                                   double efficiency, double perturb, uint64_t scenario_seed) {
  if (buses < 3 || snapshots < 1 || gens_per_bus < 1 || storage < 0) return nullptr;
  Rng rng(seed);
  const int32_t N = buses, T = snapshots, S = storage;
  const Grid grid = make_grid(rng, N, chords);
  const auto& lines = grid.lines;
  const auto& xl = grid.x;
  const auto& cap = grid.cap;
  const int32_t L = (int32_t)lines.size();
  const int32_t G = N * gens_per_bus;
  std::vector<double> gcost(G), gnom(G);
  for (int32_t g = 0; g < G; ++g) {
    gcost[g] = rng.uniform(10.0, 100.0);
    gnom[g] = rng.uniform(50.0, 500.0);
  }
  std::vector<int32_t> sbus(S);
  std::vector<double> snom(S);
  for (int32_t k = 0; k < S; ++k) {
    sbus[k] = (int32_t)rng.below(N);
    snom[k] = rng.uniform(50.0, 200.0);
  }
  std::vector<double> base(N), noise(N), scen(N, 1.0);
  for (int32_t b = 0; b < N; ++b) {
    base[b] = rng.uniform(50.0, 250.0);
    noise[b] = rng.uniform(0.8, 1.2);
  }
  if (perturb > 0.0) {
    Rng srng(scenario_seed);
    for (int32_t b = 0; b < N; ++b) scen[b] = srng.uniform(1.0 - perturb, 1.0 + perturb);
  }
  auto load = [&](int32_t b, int32_t t) {
    const double hour = (double)(t % 24);
    const double diurnal = 1.0 + 0.25 * std::sin(2.0 * M_PI * (hour - 8.0) / 24.0);
    return base[b] * diurnal * noise[b] * scen[b];
  };

  const int32_t per_t = G + L + N + 3 * S + N;
  const int32_t n = per_t * T;
  const int32_t eq_per_t = N + L + S;
  const int32_t m_eq = eq_per_t * T;
  const int32_t m = m_eq + S * T;
  auto vp = [&](int32_t t, int32_t g) { return t * per_t + g; };
  auto vf = [&](int32_t t, int32_t l) { return t * per_t + G + l; };
  auto vth = [&](int32_t t, int32_t b) { return t * per_t + G + L + b; };
  auto vch = [&](int32_t t, int32_t k) { return t * per_t + G + L + N + k; };
  auto vdis = [&](int32_t t, int32_t k) { return t * per_t + G + L + N + S + k; };
  auto vsoc = [&](int32_t t, int32_t k) { return t * per_t + G + L + N + 2 * S + k; };
  auto vshed = [&](int32_t t, int32_t b) { return t * per_t + G + L + N + 3 * S + b; };

  std::vector<Trip> tr;
  tr.reserve((size_t)T * (G + 2 * L + 3 * L + 5 * S + N + 2 * S));
  std::vector<double> q((size_t)m, 0.0);
  for (int32_t t = 0; t < T; ++t) {
    const int32_t rb = t * N;                    // balance rows
    const int32_t rf = T * N + t * L;            // flow-definition rows
    const int32_t rs = T * (N + L) + t * S;      // storage-balance rows
    for (int32_t g = 0; g < G; ++g) tr.push_back({rb + g / gens_per_bus, vp(t, g), 1.0});
    for (int32_t l = 0; l < L; ++l) {
      const int32_t a = lines[l].first, b = lines[l].second;
      tr.push_back({rb + b, vf(t, l), 1.0});
      tr.push_back({rb + a, vf(t, l), -1.0});
      tr.push_back({rf + l, vf(t, l), 1.0});
      tr.push_back({rf + l, vth(t, a), -1.0 / xl[l]});
      tr.push_back({rf + l, vth(t, b), 1.0 / xl[l]});
    }
    for (int32_t k = 0; k < S; ++k) {
      tr.push_back({rb + sbus[k], vdis(t, k), 1.0});
      tr.push_back({rb + sbus[k], vch(t, k), -1.0});
      tr.push_back({rs + k, vsoc(t, k), 1.0});
      if (t > 0) tr.push_back({rs + k, vsoc(t - 1, k), -1.0});
      tr.push_back({rs + k, vch(t, k), -efficiency});
      tr.push_back({rs + k, vdis(t, k), 1.0 / efficiency});
      // ch + dis <= p_nom, written -ch - dis >= -p_nom
      const int32_t ri = m_eq + t * S + k;
      tr.push_back({ri, vch(t, k), -1.0});
      tr.push_back({ri, vdis(t, k), -1.0});
      q[ri] = -snom[k];
    }
    for (int32_t b = 0; b < N; ++b) {
      tr.push_back({rb + b, vshed(t, b), 1.0});
      q[rb + b] = load(b, t);
    }
  }
  prima_instance* I = from_triplets(n, m, tr);
  I->n_eq = m_eq;
  std::memcpy(I->q, q.data(), (size_t)m * sizeof(double));
  const double inf = std::numeric_limits<double>::infinity();
  for (int32_t t = 0; t < T; ++t) {
    for (int32_t g = 0; g < G; ++g) {
      I->c[vp(t, g)] = gcost[g];
      I->l[vp(t, g)] = 0.0;
      I->u[vp(t, g)] = gnom[g];
    }
    for (int32_t l = 0; l < L; ++l) {
      I->l[vf(t, l)] = -cap[l];
      I->u[vf(t, l)] = cap[l];
    }
    for (int32_t b = 0; b < N; ++b) {
      I->l[vth(t, b)] = b == 0 ? 0.0 : -inf;
      I->u[vth(t, b)] = b == 0 ? 0.0 : inf;
      I->c[vshed(t, b)] = 3000.0;
      I->l[vshed(t, b)] = 0.0;
      I->u[vshed(t, b)] = load(b, t);
    }
    for (int32_t k = 0; k < S; ++k) {
      I->l[vch(t, k)] = 0.0;
      I->u[vch(t, k)] = snom[k];
      I->l[vdis(t, k)] = 0.0;
      I->u[vdis(t, k)] = snom[k];
      I->c[vdis(t, k)] = 0.5;
      I->l[vsoc(t, k)] = 0.0;
      I->u[vsoc(t, k)] = 4.0 * snom[k];
    }
  }
  return I;
}

// Multi-bus unit commitment over the dispatch network (DC flows in angle
// form, load shedding at 3000/MWh), with a binary on/off variable per
// generator and snapshot. Per snapshot t and generator g: output p, status u
// (integer, [0,1]), start-up su and shut-down sd in [0,1]. Rows, equalities
// first: nodal balance and flow definition per snapshot; then
//   p - pmin*u >= 0,  pnom*u - p >= 0,
//   su[t] - u[t] + u[t-1] >= 0,  sd[t] - u[t-1] + u[t] >= 0 (t > 0),
//   u[t] - sum(su[t-min_up+1 .. t]) >= 0                     (minimum up time)
//   -u[t] - sum(sd[t-min_down+1 .. t]) >= -1                 (minimum down time)
// with every unit off before t = 0 (sd[0] fixed at 0). Costs: marginal
// U[10,100], start-up U[100,1000]; p_nom U[50,500], p_min/p_nom U[0.2,0.5].
prima_instance* prima_gen_uc(uint64_t seed, int32_t buses, int32_t chords, int32_t snapshots,
                             int32_t gens_per_bus, int32_t min_up, int32_t min_down) {
  if (buses < 3 || snapshots < 1 || gens_per_bus < 1 || min_up < 1 || min_down < 1)
    return nullptr;
  Rng rng(seed);
  const int32_t N = buses, T = snapshots;
  const Grid grid = make_grid(rng, N, chords);
  const int32_t L = (int32_t)grid.lines.size();
  const int32_t G = N * gens_per_bus;
  std::vector<double> gcost(G), gnom(G), gmin(G), gstart(G);
  for (int32_t g = 0; g < G; ++g) {
    gcost[g] = rng.uniform(10.0, 100.0);
    gnom[g] = rng.uniform(50.0, 500.0);
    gmin[g] = rng.uniform(0.2, 0.5) * gnom[g];
    gstart[g] = rng.uniform(100.0, 1000.0);
  }
  std::vector<double> base(N), noise(N);
  for (int32_t b = 0; b < N; ++b) {
    base[b] = rng.uniform(100.0, 400.0);
    noise[b] = rng.uniform(0.8, 1.2);
  }
  auto load = [&](int32_t b, int32_t t) {
    const double hour = (double)(t % 24);
    return base[b] * (1.0 + 0.25 * std::sin(2.0 * M_PI * (hour - 8.0) / 24.0)) * noise[b];
  };
  const int32_t per_t = 4 * G + L + 2 * N;
  const int32_t n = per_t * T;
  auto vp = [&](int32_t t, int32_t g) { return t * per_t + g; };
  auto vu = [&](int32_t t, int32_t g) { return t * per_t + G + g; };
  auto vsu = [&](int32_t t, int32_t g) { return t * per_t + 2 * G + g; };
  auto vsd = [&](int32_t t, int32_t g) { return t * per_t + 3 * G + g; };
  auto vf = [&](int32_t t, int32_t l) { return t * per_t + 4 * G + l; };
  auto vth = [&](int32_t t, int32_t b) { return t * per_t + 4 * G + L + b; };
  auto vshed = [&](int32_t t, int32_t b) { return t * per_t + 4 * G + L + N + b; };
  const int32_t m_eq = (N + L) * T;
  const int32_t ineq_per_t = 6 * G;
  const int32_t m = m_eq + ineq_per_t * T;
  std::vector<Trip> tr;
  std::vector<double> q((size_t)m, 0.0);
  for (int32_t t = 0; t < T; ++t) {
    const int32_t rb = t * N, rf = T * N + t * L;
    for (int32_t g = 0; g < G; ++g) tr.push_back({rb + g / gens_per_bus, vp(t, g), 1.0});
    for (int32_t l = 0; l < L; ++l) {
      const int32_t a = grid.lines[l].first, b = grid.lines[l].second;
      tr.push_back({rb + b, vf(t, l), 1.0});
      tr.push_back({rb + a, vf(t, l), -1.0});
      tr.push_back({rf + l, vf(t, l), 1.0});
      tr.push_back({rf + l, vth(t, a), -1.0 / grid.x[l]});
      tr.push_back({rf + l, vth(t, b), 1.0 / grid.x[l]});
    }
    for (int32_t b = 0; b < N; ++b) {
      tr.push_back({rb + b, vshed(t, b), 1.0});
      q[rb + b] = load(b, t);
    }
    const int32_t ri = m_eq + t * ineq_per_t;
    for (int32_t g = 0; g < G; ++g) {
      const int32_t r0 = ri + 6 * g;
      tr.push_back({r0, vp(t, g), 1.0});  // p - pmin u >= 0
      tr.push_back({r0, vu(t, g), -gmin[g]});
      tr.push_back({r0 + 1, vu(t, g), gnom[g]});  // pnom u - p >= 0
      tr.push_back({r0 + 1, vp(t, g), -1.0});
      tr.push_back({r0 + 2, vsu(t, g), 1.0});  // su - u[t] + u[t-1] >= 0
      tr.push_back({r0 + 2, vu(t, g), -1.0});
      if (t > 0) tr.push_back({r0 + 2, vu(t - 1, g), 1.0});
      tr.push_back({r0 + 3, vsd(t, g), 1.0});  // sd - u[t-1] + u[t] >= 0
      tr.push_back({r0 + 3, vu(t, g), 1.0});
      if (t > 0) tr.push_back({r0 + 3, vu(t - 1, g), -1.0});
      tr.push_back({r0 + 4, vu(t, g), 1.0});  // u[t] - sum su >= 0
      for (int32_t k = std::max(0, t - min_up + 1); k <= t; ++k)
        tr.push_back({r0 + 4, vsu(k, g), -1.0});
      tr.push_back({r0 + 5, vu(t, g), -1.0});  // -u[t] - sum sd >= -1
      for (int32_t k = std::max(0, t - min_down + 1); k <= t; ++k)
        tr.push_back({r0 + 5, vsd(k, g), -1.0});
      q[r0 + 5] = -1.0;
    }
  }
  prima_instance* I = from_triplets(n, m, tr);
  I->n_eq = m_eq;
  std::memcpy(I->q, q.data(), (size_t)m * sizeof(double));
  const double inf = std::numeric_limits<double>::infinity();
  I->n_int = G * T;
  I->integers = (int32_t*)std::malloc((size_t)I->n_int * sizeof(int32_t));
  int32_t k = 0;
  for (int32_t t = 0; t < T; ++t) {
    for (int32_t g = 0; g < G; ++g) {
      I->c[vp(t, g)] = gcost[g];
      I->u[vp(t, g)] = gnom[g];
      I->u[vu(t, g)] = 1.0;
      I->c[vsu(t, g)] = gstart[g];
      I->u[vsu(t, g)] = 1.0;
      I->u[vsd(t, g)] = t == 0 ? 0.0 : 1.0;
      I->integers[k++] = vu(t, g);
    }
    for (int32_t l = 0; l < L; ++l) {
      I->l[vf(t, l)] = -grid.cap[l];
      I->u[vf(t, l)] = grid.cap[l];
    }
    for (int32_t b = 0; b < N; ++b) {
      I->l[vth(t, b)] = b == 0 ? 0.0 : -inf;
      I->u[vth(t, b)] = b == 0 ? 0.0 : inf;
      I->c[vshed(t, b)] = 3000.0;
      I->u[vshed(t, b)] = load(b, t);
    }
  }
  std::sort(I->integers, I->integers + I->n_int);
  return I;
}

}  // extern "C"
