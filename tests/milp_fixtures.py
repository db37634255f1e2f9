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
