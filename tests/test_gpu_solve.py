"""Parity of the on-device restarted PDHG (prima_solve / prima_mixed_solve /
prima_solve_batch) against the oracle, the reference's fixtures and ladder,
closed-form optima and HiGHS goldens. Gates: same status; objective within
1e-8 relative (|a-b|/max(1,|b|), ValidationLadder.scala:45-46) at eps 1e-8
unless stated; an infeasible instance is never Optimal."""
import json
import os

import numpy as np
import pytest

from noaidi_amd import prima
from oracle.oracle import OracleLp, default_params
from tests.lp_fixtures import FIXTURES, LADDER, OPTIMAL, RANDOM_LADDER, fixture

pytestmark = pytest.mark.gpu

P9 = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, max_iterations=200_000)
GOLD = os.path.join(os.path.dirname(__file__), "golden", "highs_dispatch.json")


def gap(a, b):
    return abs(a - b) / max(1.0, abs(b))


def oracle_of(p):
    return OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                             p.offset)


def gold(pred):
    with open(GOLD) as f:
        return [d for d in json.load(f) if pred(d["args"])]


def assert_kkt(p, s, eps=1e-8):
    """The returned point itself meets Kkt.convergence (Kkt.scala:130-143) on
    the original data: the oracle's Kkt.evaluate (Kkt.scala:68-125) recomputes
    the residuals and both objectives from the primal and dual vectors alone
    (fresh K x and K^T y in index order; the device's own evaluation differs
    only in summation order, hence the 1e-6 relative slack). The primal lies
    in its box, the inequality rows' duals in their cone, and the reported
    reduced costs are c - K^T y."""
    e = oracle_of(p).kkt(s.primal, s.dual)
    slack = 1.0 + 1e-6
    assert np.all(np.isfinite(e)), e
    assert e[0] <= slack * (eps + eps * np.linalg.norm(p.q)), ("primal residual", e)
    assert e[1] <= slack * (eps + eps * np.linalg.norm(p.c)), ("dual residual", e)
    assert abs(e[2] - e[3]) <= slack * (eps + eps * (abs(e[2]) + abs(e[3]))), ("gap", e)
    assert abs(e[2] - s.objective_value) <= 1e-9 * max(1.0, abs(e[2]))
    room = 1e-12 * (1.0 + np.abs(s.primal))
    assert np.all(s.primal >= p.l - room) and np.all(s.primal <= p.u + room)
    assert np.all(s.dual[p.n_eq:] >= 0.0)
    rc = p.c - p.to_scipy().T @ s.dual
    assert np.allclose(s.reduced_costs, rc, rtol=1e-9, atol=1e-9 * (1.0 + np.abs(p.c).max()))


@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixtures_fp64(hip64, name):
    # PdhgSuite.scala:12-35 at eps 1e-9, plus the ladder's counts
    p, eobj, ex, status = fixture(name)
    s = hip64.solve(p, P9)
    assert s.status == status, s
    if eobj is not None:
        assert abs(s.objective_value - eobj) <= 1e-6 * max(1, abs(eobj))
        assert np.allclose(s.primal, ex, atol=1e-5)
        assert abs(s.objective_value - s.dual_objective_value) <= 1e-5 * max(1, abs(eobj))
    _, _, iters, restarts = LADDER[name]
    if name != "infeasible":  # certificate timing depends on summation order
        assert s.iterations == iters and s.restarts == restarts


def test_solution_satisfies_constraints(hip64):
    # PdhgSuite.scala:49-67
    for name in OPTIMAL:
        p, *_ = fixture(name)
        s = hip64.solve(p, P9)
        kx = p.to_scipy() @ s.primal
        assert np.all(np.abs(kx[: p.n_eq] - p.q[: p.n_eq]) <= 1e-6)
        assert np.all(kx[p.n_eq:] >= p.q[p.n_eq:] - 1e-6)
        assert np.all(s.primal >= p.l - 1e-9) and np.all(s.primal <= p.u + 1e-9)


def test_congestion_prices(hip64):
    # PdhgSuite.scala:69-90
    p, *_ = fixture("economic-dispatch")
    s = hip64.solve(p, P9)
    a1, b1, a2, b2 = s.dual[:4]
    assert abs(a1 - b1) <= 1e-4 and abs(a2 - b2) > 1.0
    assert abs(a2 - 10.0) <= 1e-3 and abs(b2 - 20.0) <= 1e-3


def test_scaling_does_not_change_the_answer(hip64):
    unscaled = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, ruiz_iterations=0,
                                apply_pock_chambolle=False)
    for name in OPTIMAL:
        p, *_ = fixture(name)
        a, b = hip64.solve(p, P9), hip64.solve(p, unscaled)
        assert b.status == "Optimal"
        assert abs(a.objective_value - b.objective_value) <= 1e-5 * max(1, abs(a.objective_value))


def test_iteration_limit_exact(hip64):
    p, *_ = fixture("economic-dispatch")
    s = hip64.solve(p, prima.PdhgParams(eps_abs=1e-14, eps_rel=1e-14, max_iterations=3))
    assert s.status == "IterationLimit" and s.iterations == 3


def test_trivial_problem_zero_iterations(hip64):
    b = prima.builder(2)
    b.bounds(0, 0.0, 1.0).bounds(1, 0.0, 1.0)
    s = hip64.solve(b.build()[0], P9)
    assert s.status == "Optimal" and s.iterations == 0


def test_time_limit(hip64):
    # PdhgSuite.scala:124-144
    p = prima.LpProblem.from_dict(OracleLp.random_feasible(5, 200, 40, 80, 0.3).export())
    s = hip64.solve(p, prima.PdhgParams(eps_abs=1e-16, eps_rel=1e-16, max_iterations=50_000_000,
                                        time_limit_ms=50))
    assert s.status == "TimeLimit" and s.iterations < 50_000_000


def test_invalid_arguments_raise(hip64):
    p, *_ = fixture("economic-dispatch")
    with pytest.raises(ValueError):
        hip64.solve(p, prima.PdhgParams(restart_artificial=0.0))
    with pytest.raises(ValueError):
        hip64.solve(p, prima.PdhgParams(infeasibility_tolerance=-1e-9))
    bad = prima.LpProblem(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.u, p.l)
    with pytest.raises(ValueError):
        hip64.solve(bad)
    with pytest.raises(ValueError):
        hip64.solve(p, P9, prima.WarmStart(np.full(p.n, np.nan), np.zeros(p.m)))


def test_spectral_bound_column(hip64):
    # PdhgSuite.scala:155-174
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 10.0)
    for _ in range(100):
        b.greaterThan([(0, 1.0)], 3.0)
    s = hip64.solve(b.build()[0], P9)
    assert s.status == "Optimal" and abs(s.objective_value - 3.0) <= 1e-6


def test_random_ladder_against_oracle_and_published(hip64):
    for name, args in RANDOM_LADDER.items():
        _, oj, _, _ = LADDER[name]
        p = prima.LpProblem.from_dict(OracleLp.random_feasible(*args).export())
        s = hip64.solve(p, prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, max_iterations=500_000))
        assert s.status == "Optimal"
        assert gap(s.objective_value, oj) <= 1e-8, (name, s.objective_value)
        assert_kkt(p, s, 1e-9)
        assert s.restarts > 0


def test_config0_closed_form_fp64(hip64):
    for seed in range(3):
        p = prima.config0(seed)
        s = hip64.solve(p)
        assert s.status == "Optimal"
        assert gap(s.objective_value, p.known_objective) <= 1e-8
        assert_kkt(p, s)


def test_dispatch_against_highs_fp64(hip64):
    for d in gold(lambda a: a.get("buses") == 20)[:4]:
        p = prima.dispatch(**d["args"])
        s = hip64.solve(p, prima.PdhgParams(max_iterations=500_000))
        assert s.status == "Optimal"
        assert gap(s.objective_value, d["objective"]) <= 1e-8
        assert_kkt(p, s)


def test_mixed_precision_f32_device(hip32):
    # MixedPrecisionSuite.scala:32-60: known optima, full fp64 KKT after refinement
    for name in OPTIMAL:
        p, eobj, _, _ = fixture(name)
        s = hip32.mixed_solve(p, P9)
        assert s.status == "Optimal" and s.mixed["refined"]
        assert abs(s.objective_value - eobj) <= 1e-6 * max(1, abs(eobj))
    p = prima.config0(1)
    s = hip32.mixed_solve(p, prima.PdhgParams())
    assert s.status == "Optimal" and gap(s.objective_value, p.known_objective) <= 1e-8
    assert_kkt(p, s)


def test_iterative_refinement_path(hip32):
    # MixedParams(refinement="iterative"): float32 device passes on correction
    # LPs (shifted bounds, shifted dual cone on inequality rows) around an
    # FP64 outer loop, with the reference's contract kept: the answer is
    # verified in FP64 (or comes from the warm FP64 fallback), certificates
    # are re-checked. Optima of the reference's fixtures (several with
    # inequality rows), config 0's closed form, a dispatch LP against HiGHS,
    # and the two certificates.
    ir = prima.MixedParams(refinement="iterative")
    for name in OPTIMAL:
        p, eobj, _, _ = fixture(name)
        s = hip32.mixed_solve(p, P9, ir)
        assert s.status == "Optimal", (name, s.status, s.mixed)
        assert abs(s.objective_value - eobj) <= 1e-6 * max(1, abs(eobj)), name
    p = prima.config0(3)
    s = hip32.mixed_solve(p, prima.PdhgParams(), ir)
    assert s.status == "Optimal" and gap(s.objective_value, p.known_objective) <= 1e-8
    assert_kkt(p, s)
    d = gold(lambda a: a.get("buses") == 20)[0]
    p = prima.dispatch(**d["args"])
    s = hip32.mixed_solve(p, prima.PdhgParams(max_iterations=500_000), ir)
    assert s.status == "Optimal", (s.status, s.mixed)
    assert gap(s.objective_value, d["objective"]) <= 1e-8
    assert_kkt(p, s)
    for name, status in (("infeasible", "PrimalInfeasible"), ("unbounded", "DualInfeasible")):
        p, *_ = fixture(name)
        assert hip32.mixed_solve(p, P9, ir).status == status
    # strongly active bounds fixed in the correction LPs: same answers
    fixed = prima.MixedParams(refinement="iterative", fix_threshold=1.0)
    for name in OPTIMAL:
        p, eobj, _, _ = fixture(name)
        s = hip32.mixed_solve(p, P9, fixed)
        assert s.status == "Optimal", (name, s.status, s.mixed)
        assert abs(s.objective_value - eobj) <= 1e-6 * max(1, abs(eobj)), name
    with pytest.raises(ValueError):
        hip32.mixed_solve(p, P9, prima.MixedParams(refinement="iterative", fix_threshold=-1.0))


def test_mixed_precision_fp64_device_skips_reduced_pass(hip64):
    p, *_ = fixture("economic-dispatch")
    s = hip64.mixed_solve(p, P9)
    assert s.status == "Optimal" and s.mixed["reduced_iterations"] == 0 and not s.mixed["refined"]


def test_mixed_certificates_rechecked(hip32):
    # MixedPrecisionSuite.scala:144-172: device certificates hold in fp64
    for name, status in (("infeasible", "PrimalInfeasible"), ("unbounded", "DualInfeasible")):
        p, *_ = fixture(name)
        s = hip32.mixed_solve(p, P9)
        assert s.status == status
    p, *_ = fixture("infeasible")
    s = hip32.mixed_solve(p, P9, prima.MixedParams(always_refine=True))
    assert s.status == "PrimalInfeasible" and s.mixed["refined"]


def test_float32_pass_reaches_reduced_tolerance(hip32):
    p = prima.LpProblem.from_dict(OracleLp.random_feasible(*RANDOM_LADDER["random-600x400"]).export())
    s = hip32.solve(p, prima.PdhgParams(eps_abs=1e-5, eps_rel=1e-5, max_iterations=100_000))
    assert s.status == "Optimal"
    ref = oracle_of(p).solve(default_params(eps_abs=1e-9, eps_rel=1e-9, max_iterations=500_000))
    assert gap(s.objective_value, ref.objective) <= 1e-4


def test_batch_keeps_order_and_matches_single(hip64):
    probs = [prima.config0(s) for s in range(4)] + [fixture("infeasible")[0],
                                                   fixture("economic-dispatch")[0]]
    sols = hip64.solve_all(probs, prima.PdhgParams())
    for p, s in zip(probs[:4], sols[:4]):
        assert s.status == "Optimal" and gap(s.objective_value, p.known_objective) <= 1e-8
        single = hip64.solve(p)
        assert s.iterations == single.iterations  # same device code, same order
        assert s.objective_value == single.objective_value
    assert sols[4].status == "PrimalInfeasible"
    assert sols[5].status == "Optimal" and abs(sols[5].objective_value - 3700) <= 1e-4


def test_warm_start_from_solution(hip64):
    p = prima.config0(2)
    cold = hip64.solve(p, prima.PdhgParams(eps_abs=1e-6, eps_rel=1e-6))
    warm = hip64.solve(p, prima.PdhgParams(), prima.WarmStart.of(cold))
    assert warm.status == "Optimal" and gap(warm.objective_value, p.known_objective) <= 1e-8


def test_deterministic_across_runs(hip64):
    p = prima.config0(0)
    a, b = hip64.solve(p), hip64.solve(p)
    assert a.iterations == b.iterations and a.objective_value == b.objective_value
    assert np.array_equal(a.primal, b.primal) and np.array_equal(a.dual, b.dual)


@pytest.mark.parametrize("wg", [1, 2, 5, 16])
def test_multi_workgroup_groups_agree(hip64, wg):
    p = prima.dispatch(seed=7, buses=60, chords=30, snapshots=24, storage=15)
    d = gold(lambda a: a.get("buses") == 60)[0]
    s = hip64.solve(p, prima.PdhgParams(max_iterations=500_000, workgroups_per_lp=wg))
    assert s.status == "Optimal" and s.workgroups == wg
    assert gap(s.objective_value, d["objective"]) <= 1e-8
    assert_kkt(p, s)


def test_mixed_stall_fallback(hip32):
    # scenario 150 of the dispatch sweep at one workgroup: the float32 pass
    # reaches 1e-5 in 18,368 iterations, and the float64 pass warm-started
    # from it lets the primal weight collapse and stalls (no convergence in
    # 400k iterations; from the same start the oracle needs 69,184 and a cold
    # device pass 69,312). With stall_iterations the stalled pass is rerun cold.
    g = {d["args"]["scenario_seed"]: d["objective"] for d in gold(
        lambda a: a.get("seed") == 1 and a.get("perturb") == 0.2 and "buses" not in a)}
    p = prima.dispatch(seed=1, perturb=0.2, scenario_seed=1150)
    s = hip32.mixed_solve(p, prima.PdhgParams(max_iterations=400_000, workgroups_per_lp=1),
                          prima.MixedParams(stall_iterations=100_000))
    assert s.status == "Optimal", (s.status, s.mixed)
    assert s.mixed["fallback_iterations"] > 0
    if 1150 in g:
        assert gap(s.objective_value, g[1150]) <= 1e-8


def test_mixed_stall_fallback_in_a_batch(hip32, hip64):
    # the stalled warm pass is raced by a cold twin (started once the warm
    # pass has run a quarter of its cap with workgroups to spare); the twin
    # converges first and stands for the LP. At one workgroup per LP every
    # LP's arithmetic is its own, so the twin's result equals the LP's cold
    # float64 solve, and the batched result the LP solved alone, bit for bit.
    P = prima.PdhgParams(max_iterations=400_000, workgroups_per_lp=1)
    M = prima.MixedParams(stall_iterations=100_000)
    probs = [prima.dispatch(seed=1, perturb=0.2, scenario_seed=1000 + k) for k in (150, 3, 7, 11)]
    alone = hip32.mixed_solve(probs[0], P, M)
    batch = hip32.solve_all(probs, P, mixed=M)
    cold = hip64.solve(probs[0], prima.PdhgParams(max_iterations=300_000, workgroups_per_lp=1))
    b = batch[0]
    assert b.status == "Optimal" and b.mixed["fallback_iterations"] > 0
    assert 25_000 <= b.mixed["refinement_iterations"] <= 100_000
    assert b.mixed["fallback_iterations"] == alone.mixed["fallback_iterations"] == cold.iterations
    assert b.iterations == alone.iterations == cold.iterations
    assert np.array_equal(b.primal, alone.primal) and np.array_equal(b.dual, alone.dual)
    assert np.array_equal(b.primal, cold.primal) and np.array_equal(b.dual, cold.dual)
    for p, s in zip(probs, batch):
        assert s.status == "Optimal"
        assert_kkt(p, s)


def test_uneven_sweep_batch_against_highs(hip32):
    # instances of the uneven scenario sweep (buses log-uniform in [50, 2000],
    # snapshots in {24, 72, 168}): the 16 smallest and one of median size
    # (~130k variables) in one mixed-precision batch; every LP Optimal and
    # within 1e-8 of HiGHS
    with open(os.path.join(os.path.dirname(__file__), "golden", "highs_sweep.json")) as f:
        gold = json.load(f)[:17]
    probs = [prima.config3(d["args"]["k"]) for d in gold]
    for p, d in zip(probs, gold):
        assert (p.n, p.m, p.nnz) == (d["n"], d["m"], d["nnz"])
    # a cold float64 pass of the median instance can need more than 700k
    # iterations, depending on the trajectory (workgroups, twin timing)
    sols = hip32.solve_all(probs, prima.PdhgParams(max_iterations=2_000_000),
                           mixed=prima.MixedParams(stall_iterations=300_000))
    for p, d, s in zip(probs, gold, sols):
        assert s.status == "Optimal", (d["args"], s.status, s.mixed)
        assert gap(s.objective_value, d["objective"]) <= 1e-8, (d["args"], s.objective_value)
        assert_kkt(p, s)


@pytest.mark.timeout(900)
def test_uneven_sweep_horizons_against_highs(hip32):
    # the size-strata midpoints of the sweep with HiGHS goldens beyond the
    # smallest: 23k to 359k variables, 24, 72 and 168 snapshots, in one
    # mixed-precision batch; every LP Optimal, within 1e-8 of HiGHS, and its
    # returned point meets the KKT test on the original data
    with open(os.path.join(os.path.dirname(__file__), "golden", "highs_sweep.json")) as f:
        gold = [d for d in json.load(f)[20:] if d["status"] == 0 and d["n"] <= 400_000]
    probs = [prima.config3(d["args"]["k"]) for d in gold]
    snaps = {d["config3_args"]["snapshots"] for d in gold}
    assert snaps == {24, 72, 168}, snaps
    for p, d in zip(probs, gold):
        assert (p.n, p.m, p.nnz) == (d["n"], d["m"], d["nnz"])
    sols = hip32.solve_all(probs, prima.PdhgParams(max_iterations=4_000_000),
                           mixed=prima.MixedParams(reduced_tolerance=1e-6,
                                                   reduced_max_iterations=60_000,
                                                   stall_iterations=1_000_000))
    for p, d, s in zip(probs, gold, sols):
        assert s.status == "Optimal", (d["args"], s.status, s.mixed)
        assert gap(s.objective_value, d["objective"]) <= 1e-8, (d["args"], s.objective_value)
        assert_kkt(p, s)


@pytest.mark.timeout(300)
def test_random_2m_lp_evaluation_matches_oracle(hip64):
    # the 2M-variable random LP (1M rows, 20M nonzeros, lognormal row lengths
    # capped at 2048, uniform random columns) on the whole device for a fixed
    # number of iterations: the iteration limit is reported, and the residual
    # and objectives the device evaluated are those the oracle's Kkt.evaluate
    # recomputes from the returned vectors alone, which lie in their box
    p = prima.config2()
    s = hip64.solve(p, prima.PdhgParams(max_iterations=2_000))
    assert s.status == "IterationLimit" and s.iterations == 2_000
    assert s.workgroups >= 128
    e = oracle_of(p).kkt(s.primal, s.dual)
    assert np.all(np.isfinite(e)), e
    assert abs(e[2] - s.objective_value) <= 1e-9 * max(1.0, abs(e[2]))
    assert abs(e[3] - s.dual_objective_value) <= 1e-6 * max(1.0, abs(e[3]))
    assert abs(e[0] - s.primal_residual) <= 1e-6 * max(1.0, e[0])
    assert abs(e[1] - s.dual_residual) <= 1e-6 * max(1.0, e[1])
    room = 1e-12 * (1.0 + np.abs(s.primal))
    assert np.all(s.primal >= p.l - room) and np.all(s.primal <= p.u + room)


def test_solver_autotuner():
    # prima_tune_solver on the 200-bus dispatch LP (65k nonzeros, so group
    # sizes 32, 64 and 128, each with three row-block cuts): every candidate
    # timed, the fastest kept on the context; the next solve of that matrix
    # runs alone on the chosen group size and is Optimal within 1e-8 of HiGHS
    # with the KKT check. A context of its own keeps the choice out of the
    # other tests.
    d = gold(lambda a: a == {"seed": 1})[0]
    p = prima.dispatch(**d["args"])
    with prima.HipKernels(0, 64) as k:
        t = k.tune_solver(p, prima.PdhgParams(), iterations=512)
        us = [c[2] for c in t["candidates"]]
        assert len(us) == 9 and all(np.isfinite(us)) and min(us) == t["us_per_iteration"]
        assert {c[0] for c in t["candidates"]} == {32, 64, 128}
        g, b, _ = min(t["candidates"], key=lambda c: c[2])
        assert (t["workgroups"], t["blocks_per_workgroup"]) == (g, b)
        s = k.solve(p, prima.PdhgParams(max_iterations=500_000))
        assert s.workgroups == t["workgroups"]
        assert s.status == "Optimal", s.status
        assert gap(s.objective_value, d["objective"]) <= 1e-8
        assert_kkt(p, s)
        again = k.solve(p, prima.PdhgParams(max_iterations=500_000))
        assert again.iterations == s.iterations and again.objective_value == s.objective_value
        with pytest.raises(ValueError):
            k.tune_solver(p, prima.PdhgParams(), iterations=0)


def test_batch_beyond_capacity_admits_waiting_lps(hip64):
    # more LPs than co-resident workgroups: the extra LPs wait and take over
    # the workgroups of LPs that finish. With one workgroup per LP every LP's
    # arithmetic is independent of the others, so each result equals the one
    # from a small batch, bit for bit.
    probs = [prima.random_known(s, 120, 80, row_len=6) for s in range(300)]
    P = prima.PdhgParams(workgroups_per_lp=1)
    big = hip64.solve_all(probs, P)
    pick = [0, 1, 2, 297, 298, 299]
    small = hip64.solve_all([probs[i] for i in pick], P)
    for i, b in zip(pick, small):
        a = big[i]
        assert a.status == b.status
        assert a.iterations == b.iterations and a.objective_value == b.objective_value
    optimal = 0
    for p, s in zip(probs, big):
        assert s.status in ("Optimal", "IterationLimit")
        if s.status == "Optimal":
            optimal += 1
            assert gap(s.objective_value, p.known_objective) <= 1e-7
    assert optimal >= 285
