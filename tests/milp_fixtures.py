"""The integer programs of BranchAndBoundSuite.scala, written with the host
LpBuilder, and a brute-force enumerator that fixes their optima independently
of any LP solver."""
import itertools

import numpy as np

from noaidi_amd import prima

INF = np.inf
# BranchAndBoundSuite.scala:13 (the node relaxations run at 1e-9)
LP = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9)
KNAPSACK16_VALUES = [41.0, 50, 49, 59, 45, 47, 12, 33, 21, 44, 51, 19, 28, 37, 40, 25]
KNAPSACK16_WEIGHTS = [31.0, 27, 12, 34, 22, 19, 8, 16, 11, 24, 29, 9, 14, 20, 23, 13]


def _build(b):
    r = b.build()
    return r[0] if isinstance(r, tuple) else r


def knapsackish() -> prima.LpProblem:
    """maximise 5a + 4b s.t. 6a + 4b <= 24, a + 2b <= 6 (BranchAndBoundSuite.scala:16-24)."""
    b = prima.builder(2)
    b.objectiveCoefficient(0, -5.0).objectiveCoefficient(1, -4.0)
    b.bounds(0, 0.0, 10.0).bounds(1, 0.0, 10.0)
    b.lessThan([(0, 6.0), (1, 4.0)], 24.0).lessThan([(0, 1.0), (1, 2.0)], 6.0)
    return _build(b)


def half_integer() -> prima.LpProblem:
    """2x = 1, x in [0, 1] (BranchAndBoundSuite.scala:67-72)."""
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 1.0).equalityConstraint([(0, 2.0)], 1.0)
    return _build(b)


def best_two_of_three() -> prima.LpProblem:
    """min -(3x + 2y + 4z) s.t. x + y + z <= 2, binaries (BranchAndBoundSuite.scala:80-85)."""
    b = prima.builder(3)
    for i, c in enumerate([-3.0, -2.0, -4.0]):
        b.objectiveCoefficient(i, c)
        b.bounds(i, 0.0, 1.0)
    b.lessThan([(0, 1.0), (1, 1.0), (2, 1.0)], 2.0)
    return _build(b)


def knapsack16() -> prima.LpProblem:
    """The sixteen-item knapsack, capacity 120 (BranchAndBoundSuite.scala:160-166)."""
    b = prima.builder(16)
    for i, v in enumerate(KNAPSACK16_VALUES):
        b.objectiveCoefficient(i, -v)
        b.bounds(i, 0.0, 1.0)
    b.lessThan([(i, w) for i, w in enumerate(KNAPSACK16_WEIGHTS)], 120.0)
    return _build(b)


def unbounded_integer() -> prima.LpProblem:
    """min -x, x >= 0 free above, no rows (BranchAndBoundSuite.scala:263-266)."""
    b = prima.builder(1)
    b.objectiveCoefficient(0, -1.0).bounds(0, 0.0, INF)
    return _build(b)


def outside_the_box() -> prima.LpProblem:
    """x = 5 with x in [0, 1] (BranchAndBoundSuite.scala:306-310)."""
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 1.0).equalityConstraint([(0, 1.0)], 5.0)
    return _build(b)


def brute_force(p: prima.LpProblem, integers, box=None):
    """Minimum of c'x + offset over all integer points inside the bounds (all
    columns integer, finite bounds): the textbook answer with no LP solver in
    it. Rows are checked exactly; returns (objective, x) or (inf, None)."""
    assert sorted(integers) == list(range(p.n))
    A = p.to_scipy().toarray()
    ranges = [range(int(np.ceil(p.l[j])), int(np.floor(p.u[j])) + 1) for j in range(p.n)]
    X = np.array(list(itertools.product(*ranges)), dtype=np.float64)
    ax = X @ A.T
    ok = np.ones(len(X), bool)
    if p.n_eq:
        ok &= np.all(np.abs(ax[:, : p.n_eq] - p.q[: p.n_eq]) <= 1e-12, axis=1)
    if p.m > p.n_eq:
        ok &= np.all(ax[:, p.n_eq:] >= p.q[p.n_eq:] - 1e-12, axis=1)
    if not ok.any():
        return np.inf, None
    obj = X[ok] @ p.c + p.offset
    k = int(np.argmin(obj))
    return float(obj[k]), X[ok][k]


GOLDEN_DIR = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def uc_milp_goldens():
    """Small unit-commitment MILPs with their HiGHS MILP optimum."""
    import json
    import os
    with open(os.path.join(GOLDEN_DIR, "highs_uc_milp.json")) as f:
        return json.load(f)


def uc_frontier():
    """The 100-bus, 24-snapshot UC relaxation and its depth-9 frontier: node k
    fixes branch column i to bit i of k. Returns (root, golden, node(k))."""
    import json
    import os
    with open(os.path.join(GOLDEN_DIR, "highs_uc.json")) as f:
        g = json.load(f)
    root = prima.unit_commitment(**g["args"])
    cols = g["branch_columns"]

    def node(k: int) -> prima.LpProblem:
        l, u = root.l.copy(), root.u.copy()
        for i, c in enumerate(cols):
            l[c] = u[c] = float((k >> i) & 1)
        return prima.LpProblem(root.n, root.m, root.n_eq, root.row_ptr, root.col, root.val,
                               root.c, root.q, l, u, root.offset)

    return root, g, node


# ---------------------------------------------------------------------------
# MilpLadder (MilpLadder.scala:14-131) and its published run
# (modules/prima-core/NOTES.md:512-522): status Optimal everywhere, objective
# and B&B node count per instance, at the ladder's params (MilpLadder.scala:93-94).

LADDER_LP = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, max_iterations=200_000)
LADDER_MAX_NODES = 50_000


def _knapsack(values, weights, capacity):
    b = prima.builder(len(values))
    for i, v in enumerate(values):
        b.objectiveCoefficient(i, -float(v))
        b.bounds(i, 0.0, 1.0)
    b.lessThan([(i, float(w)) for i, w in enumerate(weights)], float(capacity))
    return _build(b), list(range(len(values)))


def _set_cover(sets, elements):
    b = prima.builder(len(sets))
    for j in range(len(sets)):
        b.objectiveCoefficient(j, 1.0)
        b.bounds(j, 0.0, 1.0)
    for e in range(elements):
        covering = [(j, 1.0) for j, s in enumerate(sets) if e in s]
        assert covering
        b.greaterThan(covering, 1.0)
    return _build(b), list(range(len(sets)))


def random_mixed(seed: int, n: int, m: int):
    """MilpLadder.randomMixed (MilpLadder.scala:71-84), drawn from the
    java.util.Random restatement in the oracle in the reference's order."""
    from oracle.oracle import JRandom
    rng = JRandom(seed)
    b = prima.builder(n)
    for j in range(n):
        b.objectiveCoefficient(j, -(1.0 + rng.next_int(9)))
        b.bounds(j, 0.0, 1.0 + rng.next_int(3))
    for _ in range(m):
        keep = [j for j in range(n) if rng.next_double() < 0.6]
        terms = [(j, 1.0 + float(rng.next_int(5))) for j in keep]
        if terms:
            b.lessThan(terms, 3.0 + rng.next_int(10))
    return _build(b), [j for j in range(n) if j % 2 == 0]


def milp_ladder():
    """name -> (problem, integer columns, published objective, published nodes)."""
    k5 = [{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 0}, {0, 2}, {1, 3}, {2, 4}, {3, 0}, {4, 1}]
    out = {
        "fractional-pair": (knapsackish(), [0, 1], -20.0, 5),
        "knapsack-8": _knapsack([10, 13, 18, 31, 7, 15, 22, 9], [3, 4, 5, 9, 2, 5, 7, 3], 17) +
        (-59.0, 23),
        "knapsack-16": (knapsack16(), list(range(16)), -271.0, 305),
        "set-cover-triangle": _set_cover([{0, 1}, {1, 2}, {0, 2}], 3) + (2.0, 5),
        "set-cover-k5-edges": _set_cover(k5, 5) + (3.0, 59),
        "random-mixed-1": random_mixed(1, 8, 4) + (-27.8, 5),
        "random-mixed-9": random_mixed(9, 10, 5) + (-31.4, 7),
        "random-mixed-3": random_mixed(3, 12, 6) + (-14.0, 9),
        "random-mixed-4": random_mixed(4, 14, 7) + (-33.512820, 5),
    }
    return out
