"""The oracle's branch-and-bound (a restatement of BranchAndBound.scala over
the oracle's FP64 PDHG) against BranchAndBoundSuite's hand-worked optima and
against brute-force enumeration of every integer point."""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests import milp_fixtures as mf


def params(**kw):
    return orc.default_params(eps_abs=1e-9, eps_rel=1e-9, **kw)


def olp(p):
    return orc.OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                                 p.offset)


def test_integer_optimum_differs_from_relaxation():
    p = mf.knapsackish()
    rel = olp(p).solve(params())
    assert rel.status == "Optimal" and abs(rel.objective + 21.0) < 1e-6
    assert abs(rel.primal[1] - 1.5) < 1e-5
    r = orc.bnb_solve(olp(p), [0, 1], params())
    assert r.status == "Optimal"
    assert abs(r.objective + 20.0) < 1e-6
    assert r.primal[0] == 4.0 and r.primal[1] == 0.0
    assert all(v == round(v) for v in r.primal)
    assert r.best_bound <= r.objective + 1e-9 and abs(r.objective - r.best_bound) <= 1e-6
    # the objective is c'x at the reported x
    assert abs(r.objective - (p.offset + float(np.dot(p.c, r.primal)))) <= 1e-12


def test_one_integer_column_keeps_the_other_fractional():
    r = orc.bnb_solve(olp(mf.knapsackish()), [0], params())
    assert r.status == "Optimal" and abs(r.objective + 21.0) < 1e-6
    assert r.primal[0] == 3.0 and abs(r.primal[1] - 1.5) < 1e-5


def test_empty_integer_set_short_circuits():
    r = orc.bnb_solve(olp(mf.knapsackish()), [], params())
    assert r.status == "Optimal" and abs(r.objective + 21.0) < 1e-6 and r.nodes_explored == 1


def test_no_integer_point_is_infeasible():
    p = mf.half_integer()
    assert olp(p).solve(params()).status == "Optimal"
    assert orc.bnb_solve(olp(p), [0], params()).status == "Infeasible"
    assert mf.brute_force(p, [0])[1] is None


def test_binaries_match_brute_force():
    p = mf.best_two_of_three()
    r = orc.bnb_solve(olp(p), [0, 1, 2], params())
    best, x = mf.brute_force(p, [0, 1, 2])
    assert r.status == "Optimal" and abs(r.objective - best) < 1e-6 and best == -7.0
    assert list(r.primal) == list(x) == [1.0, 0.0, 1.0]


def test_node_limit_downgrades():
    r = orc.bnb_solve(olp(mf.knapsackish()), [0, 1], params(), max_nodes=1)
    assert r.status not in ("Optimal", "Infeasible", "Unbounded")


def test_warm_start_does_not_change_the_answer():
    cold = orc.bnb_solve(olp(mf.knapsackish()), [0, 1], params(), warm_start=False)
    warm = orc.bnb_solve(olp(mf.knapsackish()), [0, 1], params(), warm_start=True)
    assert warm.status == cold.status and abs(warm.objective - cold.objective) <= 1e-9
    assert list(warm.primal) == list(cold.primal)


def test_out_of_range_column_rejected():
    with pytest.raises(ValueError):
        orc.bnb_solve(olp(mf.knapsackish()), [0, 7], params())


def test_predicates():
    # BranchAndBoundSuite.scala:130-158
    assert not orc.skips_before_solving(math.nan, 10.0, True)
    assert not orc.skips_before_solving(-math.inf, 10.0, True)
    assert orc.skips_before_solving(11.0, 10.0, True)
    assert not orc.skips_before_solving(9.0, 10.0, True)
    assert not orc.skips_before_solving(11.0, math.inf, False)
    assert orc.safe_bound_of(True, 5.0, 1.0) == 4.0
    assert orc.safe_bound_of(True, math.nan, 1.0) == -math.inf
    assert orc.safe_bound_of(True, 5.0, math.nan) == -math.inf
    assert orc.safe_bound_of(False, 5.0, 1.0) == -math.inf


def test_java_round():
    assert orc.java_round(2.5) == 3.0 and orc.java_round(-2.5) == -2.0
    assert orc.java_round(0.49999999999999994) == 0.0 and orc.java_round(math.nan) == 0.0
    assert orc.java_round(-0.5) == 0.0 and orc.java_round(1e300) == 1e300


def test_margin_costs_nodes_and_matches_enumeration():
    p = mf.knapsack16()
    ints = list(range(16))
    best, _ = mf.brute_force(p, ints)
    runs = {f: orc.bnb_solve(olp(p), ints, params(), pruning_safety_factor=f)
            for f in (0.0, 10.0, 1e9)}
    for r in runs.values():
        assert r.status == "Optimal" and abs(r.objective - best) <= 1e-6, r
    assert abs(runs[10.0].objective - runs[0.0].objective) <= 1e-9
    assert runs[10.0].nodes_explored >= runs[0.0].nodes_explored
    assert runs[1e9].nodes_explored > runs[10.0].nodes_explored
    assert runs[10.0].nodes_explored > 50


def test_starved_relaxations_are_counted():
    r = orc.bnb_solve(olp(mf.knapsackish()), [0, 1],
                      orc.default_params(eps_abs=1e-14, eps_rel=1e-14, max_iterations=2))
    assert r.status not in ("Optimal", "Infeasible", "Unbounded") and r.unproven_nodes > 0


def test_unbounded_relaxation():
    assert orc.bnb_solve(olp(mf.unbounded_integer()), [0], params()).status == "Unbounded"


def test_no_answer_means_no_point():
    r = orc.bnb_solve(olp(mf.outside_the_box()), [], params())
    assert r.status == "Infeasible" and len(r.primal) == 0


@pytest.mark.parametrize("case", range(3))
def test_unit_commitment_matches_highs_milp(case):
    # independent pin: HiGHS's MILP optimum of small multi-bus UC instances
    from noaidi_amd import prima
    g = mf.uc_milp_goldens()[case]
    p = prima.unit_commitment(**g["args"])
    r = orc.bnb_solve(olp(p), list(p.integers), orc.default_params(), max_nodes=2000)
    assert r.status == "Optimal"
    assert abs(r.objective - g["milp_objective"]) / abs(g["milp_objective"]) <= 1e-8
    assert all(r.primal[j] in (0.0, 1.0) for j in p.integers)


@pytest.mark.parametrize("name", list(mf.milp_ladder()))
def test_milp_ladder_matches_published_run(name):
    # NOTES.md:512-522 of prima-core: Optimal, objective (6 dp) and node count
    p, ints, obj, nodes = mf.milp_ladder()[name]
    r = orc.bnb_solve(olp(p), ints, orc.default_params(eps_abs=1e-9, eps_rel=1e-9,
                                                       max_iterations=200_000),
                      max_nodes=mf.LADDER_MAX_NODES)
    assert r.status == "Optimal" and r.unproven_nodes == 0
    assert abs(r.objective - obj) <= 1e-6  # published to 6 dp
    assert r.nodes_explored == nodes
