"""Presolve and postsolve against PresolveSuite (PresolveSuite.scala:12-286):
solving with presolve must give the same status and objective as solving
without, the restored point must satisfy the original constraints, and the
restored duals must be a valid, optimal dual of the original problem. The CPU
tests solve with the oracle; the GPU test solves with the device."""
import math

import numpy as np
import pytest

from noaidi_amd import prima
from noaidi_amd.presolve import kkt_converged, presolve
from oracle.oracle import OracleLp, default_params
from tests.lp_fixtures import FIXTURES, fixture

INF = math.inf


def oracle_solver(p):
    if p.n == 0 and p.m == 0:
        # nothing left to iterate on: the empty LP at its offset
        return prima.LpSolution("Optimal", np.zeros(0), np.zeros(0), np.zeros(0), p.offset,
                                p.offset, 0, 0, 0.0, 0.0, 0.0, float("nan"), float("nan"))
    o = OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                          p.offset)
    s = o.solve(default_params(eps_abs=1e-9, eps_rel=1e-9, max_iterations=100_000))
    return prima.LpSolution(s.status, s.primal, s.dual, s.reduced_costs, s.objective,
                            s.dual_objective, s.iterations, s.restarts, s.solve_time_ms,
                            s.primal_residual, s.dual_residual, s.final_step_size,
                            s.final_primal_weight)


def agrees(problem, solver, hint):
    direct = solver(problem)
    res = presolve(problem)
    if res.proven_status is not None:
        assert res.proven_status == direct.status, hint
        with pytest.raises(RuntimeError):
            res.reduced
        with pytest.raises(RuntimeError):
            res.restore(direct)
        return
    restored = res.restore(solver(res.reduced))
    assert restored.status == direct.status, hint
    if direct.status != "Optimal":
        return
    assert abs(restored.objective_value - direct.objective_value) <= 1e-6 * max(
        1.0, abs(direct.objective_value)), hint
    A = problem.to_scipy()
    kx = A @ restored.primal
    assert np.all(np.abs(kx[: problem.n_eq] - problem.q[: problem.n_eq]) <= 1e-6), hint
    assert np.all(kx[problem.n_eq:] >= problem.q[problem.n_eq:] - 1e-6), hint
    assert np.all(restored.primal >= problem.l - 1e-6) and np.all(
        restored.primal <= problem.u + 1e-6), hint
    assert np.all(restored.dual[problem.n_eq:] >= -1e-6), hint
    kty = A.T @ restored.dual
    assert np.all(np.abs(restored.reduced_costs - (problem.c - kty)) <= 1e-6), hint
    kkt = (restored.primal_residual, restored.dual_residual, restored.objective_value,
           restored.dual_objective_value)
    assert kkt_converged(problem, kkt, 1e-6, 1e-6), (hint, kkt)


def random_feasible(seed):
    # LpFixtures.randomFeasible(seed, 40, 8, 14, 0.25) through the oracle's
    # java.util.Random restatement
    d = OracleLp.random_feasible(seed, 40, 8, 14, 0.25).export()
    return prima.LpProblem(d["n"], d["m"], d["n_eq"], d["row_ptr"], d["col"], d["val"], d["c"],
                           d["q"], d["l"], d["u"], d["offset"])


@pytest.mark.parametrize("name", list(FIXTURES))
def test_fixtures(name):
    agrees(fixture(name)[0], oracle_solver, name)


@pytest.mark.parametrize("seed", range(1, 9))
def test_random_sparse(seed):
    agrees(random_feasible(seed), oracle_solver, f"random {seed}")


def test_unbounded_survives():
    res = presolve(fixture("unbounded")[0])
    assert res.unbounded_if_feasible and res.proven_status is None
    assert res.restore(oracle_solver(res.reduced)).status == "DualInfeasible"


def test_improving_direction_on_infeasible_problem():
    b = prima.builder(3)
    b.objectiveCoefficient(0, -1.0)
    for i in range(3):
        b.bounds(i, 0.0, INF)
    b.greaterThan([(1, 1.0), (2, 1.0)], 10.0).lessThan([(1, 1.0), (2, 1.0)], 1.0)
    p = b.build()[0]
    assert oracle_solver(p).status == "PrimalInfeasible"
    res = presolve(p)
    assert res.unbounded_if_feasible and res.proven_status is None
    assert res.restore(oracle_solver(res.reduced)).status == "PrimalInfeasible"


def test_slack_singleton_priced_at_zero():
    p = fixture("range-row")[0]
    res = presolve(p)
    r = res.restore(oracle_solver(res.reduced))
    assert abs(r.objective_value - 2.0) <= 1e-6
    assert np.all(r.dual[p.n_eq:] >= -1e-9)
    assert abs(r.objective_value - r.dual_objective_value) <= 1e-5


def test_fixed_column_cascade():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 1.0).objectiveCoefficient(1, 5.0)
    b.bounds(0, 0.0, 10.0).bounds(1, 3.0, 3.0).greaterThan([(0, 1.0), (1, 1.0)], 4.0)
    res = presolve(b.build()[0])
    assert res.stats.fixed_columns == 1 and res.reduced.n == 0
    r = res.restore(oracle_solver(res.reduced))
    assert r.status == "Optimal" and r.primal[1] == 3.0 and abs(r.objective_value - 16.0) <= 1e-6


def test_singleton_row_dual_recovered():
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 100.0).greaterThan([(0, 2.0)], 6.0)
    res = presolve(b.build()[0])
    assert res.stats.singleton_rows == 1 and res.reduced.m == 0
    r = res.restore(oracle_solver(res.reduced))
    assert abs(r.primal[0] - 3.0) <= 1e-6 and abs(r.objective_value - 3.0) <= 1e-6
    assert abs(r.dual[0] - 0.5) <= 1e-5


def test_empty_rows():
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 2.0, 2.0).greaterThan([(0, 1.0)], 1.0)
    res = presolve(b.build()[0])
    assert not res.proven_infeasible
    r = res.restore(oracle_solver(res.reduced))
    assert abs(r.objective_value - 2.0) <= 1e-6 and r.dual[0] == 0.0
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 0.0).greaterThan([(0, 1.0)], 5.0)
    assert presolve(b.build()[0]).proven_infeasible
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 1.0).greaterThan([(0, 1.0)], 5.0)
    assert presolve(b.build()[0]).proven_infeasible


def test_empty_column_set_by_cost():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 1.0).objectiveCoefficient(1, -2.0)
    b.bounds(0, 0.0, 10.0).bounds(1, 0.0, 4.0).greaterThan([(0, 1.0)], 3.0)
    res = presolve(b.build()[0])
    assert res.stats.empty_columns >= 1
    r = res.restore(oracle_solver(res.reduced))
    assert r.primal[1] == 4.0 and abs(r.objective_value - (3.0 - 8.0)) <= 1e-6


def test_cascade_across_passes():
    b = prima.builder(3)
    for i in range(3):
        b.objectiveCoefficient(i, 1.0)
    b.bounds(0, 0.0, 10.0).bounds(1, 0.0, 10.0).bounds(2, 5.0, 5.0)
    b.equalityConstraint([(2, 1.0)], 5.0)
    b.equalityConstraint([(0, 1.0), (2, 1.0)], 9.0)
    b.equalityConstraint([(0, 1.0), (1, 1.0)], 6.0)
    res = presolve(b.build()[0])
    assert res.reduced.n == 0
    r = res.restore(oracle_solver(res.reduced))
    assert r.primal[2] == 5.0 and abs(r.primal[0] - 4.0) <= 1e-6
    assert abs(r.primal[1] - 2.0) <= 1e-6 and abs(r.objective_value - 11.0) <= 1e-6


def test_nothing_to_reduce():
    p = fixture("economic-dispatch")[0]
    res = presolve(p)
    assert res.stats.is_empty
    assert (res.reduced.n, res.reduced.m) == (p.n, p.m)


@pytest.mark.gpu
def test_presolve_then_device_solve(hip64):
    P = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9)

    def device(p):
        return hip64.solve(p, P)

    for name in FIXTURES:
        agrees(fixture(name)[0], device, name)
    for seed in range(1, 9):
        agrees(random_feasible(seed), device, f"random {seed}")
    # an LP that presolve empties completely is Optimal at its offset with no
    # device launch (Pdhg.scala:209-212)
    empty = prima.LpProblem(0, 0, 0, [0], [], [], [], [], [], [], 3.5)
    s = hip64.solve(empty, P)
    assert s.status == "Optimal" and s.objective_value == 3.5 and s.iterations == 0
