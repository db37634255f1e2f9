"""Pins the CPU oracle (oracle/prima_oracle.c) before anything is compared
against it: the reference's own golden numbers and known-answer tests, plus
what the mathematics fixes (closed forms, brute force, invariants)."""
import math

import numpy as np
import pytest

from oracle.oracle import JRandom, OracleLp, default_mixed_params, default_params, fdlibm_log
from tests.lp_fixtures import FIXTURES, LADDER, OPTIMAL, RANDOM_LADDER, fixture
from noaidi_amd import prima


def oracle_of(p: prima.LpProblem) -> OracleLp:
    return OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                             p.offset)


# --- java.util.Random and StrictMath.log restatements -----------------------

def test_java_random_known_answers():
    # Values every JDK prints for these seeds.
    assert JRandom(0).next_double() == 0.730967787376657
    assert JRandom(42).next_double() == 0.7275636800328681
    assert JRandom(0).next_gaussian() == 0.8025330637390305
    assert JRandom(42).next_gaussian() == 1.1419053154730547
    r = JRandom(42)
    assert [r.next_int(10) for _ in range(5)] == [0, 3, 8, 4, 0]


def test_fdlibm_log_within_one_ulp_of_libm():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 1, 20000), rng.uniform(0, 2, 20000),
                         np.exp(rng.uniform(-700, 700, 20000)), [1.0, 2.0, 0.5, 1e-300, 1e300]])
    diff = 0
    for x in xs:
        a, b = fdlibm_log(float(x)), math.log(float(x))
        assert abs(a - b) <= math.ulp(b) + 1e-300, x
        diff += a != b
    assert diff < len(xs) // 10  # fdlibm is within 1 ulp, not correctly rounded
    assert fdlibm_log(1.0) == 0.0 and math.isinf(fdlibm_log(0.0)) and math.isnan(fdlibm_log(-1.0))


# --- SparseMatrix (SparseMatrixSuite.scala) ----------------------------------

def test_triplets_are_canonicalised():
    t = [(1, 2, 5.0), (0, 1, 2.0), (0, 1, 3.0), (1, 0, 1.0), (0, 0, 4.0), (0, 0, -4.0)]
    r, c, v = zip(*t)
    o = OracleLp.from_triplets(3, 2, 0, r, c, v, np.zeros(3), np.zeros(2), np.zeros(3),
                               np.ones(3))
    d = o.export()
    assert o.nnz == 3 and list(d["col"]) == [1, 0, 2] and list(d["val"]) == [5.0, 1.0, 5.0]
    rp, ci, va = prima.csr_from_triplets(2, 3, r, c, v)
    assert list(rp) == list(d["row_ptr"]) and list(ci) == [1, 0, 2] and list(va) == [5.0, 1.0, 5.0]


def test_empty_rows_and_multiply():
    o = OracleLp.from_triplets(2, 3, 0, [0, 2], [0, 1], [1.0, 2.0], np.zeros(2), np.zeros(3),
                               np.zeros(2), np.ones(2))
    assert list(o.export()["row_ptr"]) == [0, 1, 1, 2]
    assert list(o.spmv([3.0, 4.0])) == [3.0, 0.0, 8.0]


def test_spectral_norm_bound():
    col = OracleLp.from_triplets(1, 100, 0, range(100), [0] * 100, [1.0] * 100, np.zeros(1),
                                 np.zeros(100), np.zeros(1), np.ones(1))
    assert abs(col.spectral_norm_bound() - 10.0) < 1e-12
    diag = OracleLp.from_triplets(3, 3, 0, [0, 1, 2], [0, 1, 2], [2.0, -7.0, 3.0], np.zeros(3),
                                  np.zeros(3), np.zeros(3), np.ones(3))
    assert abs(diag.spectral_norm_bound() - 7.0) < 1e-12


def test_non_finite_entries_rejected():
    with pytest.raises(ValueError):
        OracleLp.from_triplets(1, 1, 0, [0], [0], [np.inf], [0.0], [0.0], [0.0], [1.0])
    with pytest.raises(ValueError):
        prima.csr_from_triplets(1, 1, [0], [0], [np.nan])


def test_spmv_and_transpose_against_dense():
    rng = np.random.default_rng(1)
    for _ in range(20):
        m, n = rng.integers(1, 15, 2)
        D = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)
        r, c = np.nonzero(D)
        o = OracleLp.from_triplets(n, m, 0, r, c, D[r, c], np.zeros(n), np.zeros(m),
                                   np.zeros(n), np.ones(n))
        x, y = rng.standard_normal(n), rng.standard_normal(m)
        assert np.allclose(o.spmv(x), D @ x, atol=1e-12)
        assert np.allclose(o.spmv(y, transpose=True), D.T @ y, atol=1e-12)


# --- Scaling (ScalingSuite.scala) -------------------------------------------

def test_scaling_equilibrates_and_is_diagonal():
    o = OracleLp.random_feasible(2, 200, 40, 80, 0.10)
    d = o.export()
    rs, cs, sv = o.scaling()
    rows = np.repeat(np.arange(o.m), np.diff(d["row_ptr"]))
    assert np.allclose(sv, rs[rows] * d["val"] * cs[d["col"]], rtol=1e-13)
    A = np.zeros((o.m, o.n))
    A[rows, d["col"]] = np.abs(sv)
    rn, cn = A.max(axis=1), A.max(axis=0)
    assert np.all((rn > 0.05) & (rn < 20)) and np.all((cn[cn > 0] > 0.05) & (cn[cn > 0] < 20))
    rs0, cs0, sv0 = o.scaling(ruiz=0, pc=False)
    assert np.all(rs0 == 1.0) and np.all(cs0 == 1.0) and np.array_equal(sv0, d["val"])


def test_scaling_leaves_empty_rows_and_columns_alone():
    o = OracleLp.from_triplets(3, 2, 0, [0], [0], [4.0], np.zeros(3), np.zeros(2), np.zeros(3),
                               np.ones(3))
    rs, cs, sv = o.scaling()
    assert rs[1] == 1.0 and cs[1] == 1.0 and cs[2] == 1.0 and np.all(np.isfinite(sv))


# --- Kkt and certificates (CertificatesSuite.scala) ---------------------------

def _contradiction():
    b = prima.builder(1)
    b.bounds(0, -np.inf, np.inf).greaterThan([(0, 1.0)], 1.0).lessThan([(0, 1.0)], 0.0)
    return oracle_of(b.build()[0])


def _absorbed():
    b = prima.builder(1)
    b.bounds(0, 0.0, 5.0).lessThan([(0, 1.0)], -1.0).lessThan([(0, 1.0)], 100.0)
    return oracle_of(b.build()[0])


def test_farkas_direction_certifies_and_is_scale_invariant():
    c = _contradiction()
    assert abs(c.primal_infeasibility([1.0, 1.0])) < 1e-12
    assert abs(c.primal_infeasibility([1e9, 1e9]) - c.primal_infeasibility([1.0, 1.0])) < 1e-12
    assert c.primal_infeasibility([0.0, 0.0]) is None
    assert c.primal_infeasibility([0.0, 1.0]) is None
    assert c.primal_infeasibility([-1.0, 0.0]) is None


def test_cone_violation_counts_exactly():
    a = _absorbed()
    assert abs(a.primal_infeasibility([2.0, 0.0])) < 1e-12
    assert abs(a.primal_infeasibility([2.0, -1.0]) - 1.0 / math.sqrt(5.0)) < 1e-12


def test_unbounded_ray_and_bound_violations():
    b = prima.builder(1)
    b.objectiveCoefficient(0, -1.0).bounds(0, 0.0, np.inf).greaterThan([(0, 1.0)], 0.0)
    ray = oracle_of(b.build()[0])
    assert abs(ray.dual_infeasibility([1.0])) < 1e-12
    assert ray.dual_infeasibility([0.0]) is None and ray.dual_infeasibility([-1.0]) is None
    b = prima.builder(1)
    b.objectiveCoefficient(0, -1.0).bounds(0, 0.0, 10.0).greaterThan([(0, 1.0)], 0.0)
    assert oracle_of(b.build()[0]).dual_infeasibility([1.0]) > 1e-6


def test_classify_prefers_primal_infeasibility():
    b = prima.builder(2)
    b.objectiveCoefficient(1, -1.0).bounds(0, -np.inf, np.inf).bounds(1, 0.0, np.inf)
    b.greaterThan([(0, 1.0)], 1.0).lessThan([(0, 1.0)], 0.0)
    o = oracle_of(b.build()[0])
    assert o.classify([0.0, 1.0], [1.0, 1.0], 1e-8) == "PrimalInfeasible"


def test_kkt_closed_form_optimum_has_zero_error():
    p = prima.config0()
    o = oracle_of(p)
    e = o.kkt(p.x_star, p.y_star)
    assert e[0] < 1e-9 and e[1] == 0.0
    assert abs(e[2] - p.known_objective) < 1e-9 and abs(e[3] - p.known_objective) < 1e-9


# --- LpBuilder: host mirror == oracle restatement, bit for bit ---------------

@pytest.mark.parametrize("name", list(FIXTURES))
def test_builder_matches_oracle_fixture(name):
    p, *_ = fixture(name)
    d = OracleLp.fixture(name).export()
    assert (p.n, p.m, p.n_eq) == (d["n"], d["m"], d["n_eq"])
    for k in ("row_ptr", "col", "val", "c", "q", "l", "u"):
        assert np.array_equal(getattr(p, k), d[k]), k


# --- PDHG: the reference ladder (NOTES.md:43-58) -----------------------------

LADDER_PARAMS = dict(eps_abs=1e-9, eps_rel=1e-9, max_iterations=500_000)


@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixture_ladder_exact(name):
    obj, _, iters, restarts = LADDER[name]
    _, eobj, ex, status = fixture(name)
    s = OracleLp.fixture(name).solve(default_params(**LADDER_PARAMS))
    assert s.status == status
    assert (s.iterations, s.restarts) == (iters, restarts)
    if eobj is not None:
        assert abs(s.objective - eobj) <= 1e-6 * max(1, abs(eobj))
        assert np.allclose(s.primal, ex, atol=1e-5)


def test_random_ladder():
    for name, args in RANDOM_LADDER.items():
        obj, oj, iters, restarts = LADDER[name]
        o = OracleLp.random_feasible(*args)
        s = o.solve(default_params(**LADDER_PARAMS))
        assert s.status == "Optimal"
        # published to 6 dp; the reference's own gate against ojAlgo is 1e-8
        assert abs(s.objective - oj) / max(1, abs(oj)) <= 1e-8
        if name == "random-60x30":  # short enough to be trajectory-exact
            assert (s.iterations, s.restarts) == (iters, restarts)
        else:  # libm vs JVM last-ulp pow/exp/log: trajectories drift
            assert 0.5 * iters <= s.iterations <= 1.5 * iters


def test_ladder_nnz_matches_published_sizes():
    # NOTES.md:56-58: 439 / 2,464 / 9,529 nonzeros
    for name, nnz in (("random-60x30", 439), ("random-200x120", 2464), ("random-600x400", 9529)):
        assert OracleLp.random_feasible(*RANDOM_LADDER[name]).nnz == nnz


def test_float32_pass_matches_published_iterations():
    # NOTES.md:228-233: float32 reaches 1e-5 in 1,792 iterations
    o = OracleLp.random_feasible(*RANDOM_LADDER["random-600x400"])
    s = o.solve(default_params(eps_abs=1e-5, eps_rel=1e-5, max_iterations=500_000), precision=32)
    assert s.status == "Optimal" and s.iterations == 1792


def test_mixed_precision_dispatch_savings_match_published():
    # NOTES.md:212-214: economic-dispatch saves 75% at 1e-9 and 100% at 1e-6
    o = OracleLp.fixture("economic-dispatch")
    for eps, saved in ((1e-9, 0.75), (1e-6, 1.0)):
        p = default_params(eps_abs=eps, eps_rel=eps, max_iterations=500_000)
        cold = o.solve(p)
        m = o.mixed_solve(p, 32)
        assert m.status == "Optimal" and abs(m.objective - 3700) < 1e-3
        assert abs((1 - m.mixed.refinement_iterations / cold.iterations) - saved) < 1e-12


def test_mixed_precision_fp64_device_skips_reduced_pass():
    o = OracleLp.fixture("economic-dispatch")
    m = o.mixed_solve(default_params(), 64)
    assert m.mixed.reduced_iterations == 0 and m.mixed.refined == 0


def test_iteration_limit_is_exact():
    s = OracleLp.fixture("economic-dispatch").solve(
        default_params(eps_abs=1e-14, eps_rel=1e-14, max_iterations=3))
    assert s.status == "IterationLimit" and s.iterations == 3


def test_trivial_problem_terminates_at_zero_iterations():
    b = prima.builder(2)
    b.bounds(0, 0.0, 1.0).bounds(1, 0.0, 1.0)
    s = oracle_of(b.build()[0]).solve(default_params(**LADDER_PARAMS))
    assert s.status == "Optimal" and s.iterations == 0


def test_closed_form_config0():
    p = prima.config0()
    s = oracle_of(p).solve(default_params())
    assert s.status == "Optimal"
    assert abs(s.objective - p.known_objective) / max(1, abs(p.known_objective)) <= 1e-8
    assert abs(p.c @ p.x_star - p.q @ p.y_star) <= 1e-9 * max(1, abs(p.known_objective))


def test_small_dispatch_against_highs():
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "highs_dispatch.json")) as f:
        gold = [d for d in json.load(f) if d["args"].get("buses") == 20]
    assert len(gold) == 8
    for d in gold[:4]:
        p = prima.dispatch(**d["args"])
        assert (p.n, p.m, p.nnz) == (d["n"], d["m"], d["nnz"])
        s = oracle_of(p).solve(default_params(max_iterations=500_000))
        assert s.status == "Optimal"
        assert abs(s.objective - d["objective"]) / max(1, abs(d["objective"])) <= 1e-8
