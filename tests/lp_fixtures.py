"""The reference's LpFixtures (LpFixtures.scala:25-165) written with the
host-side LpBuilder: hand-derived optima, primal points and statuses."""
import numpy as np

from noaidi_amd import prima

INF = np.inf


def _tiny_lower_bound():
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, 0.0, 10.0).greaterThan([(0, 1.0)], 3.0)
    return b, 3.0, [3.0], "Optimal"


def _product_mix():
    b = prima.builder(2)
    b.objectiveCoefficient(0, -3.0).objectiveCoefficient(1, -5.0)
    b.bounds(0, 0.0, INF).bounds(1, 0.0, INF)
    b.lessThan([(0, 1.0)], 4.0).lessThan([(1, 2.0)], 12.0).lessThan([(0, 3.0), (1, 2.0)], 18.0)
    return b, -36.0, [2.0, 6.0], "Optimal"


def _equality_split():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 2.0).objectiveCoefficient(1, 3.0)
    b.bounds(0, 0.0, 6.0).bounds(1, 0.0, 10.0)
    b.equalityConstraint([(0, 1.0), (1, 1.0)], 10.0)
    return b, 24.0, [6.0, 4.0], "Optimal"


def _degenerate():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 1.0).objectiveCoefficient(1, 1.0)
    b.bounds(0, 0.0, 5.0).bounds(1, 0.0, 5.0)
    b.greaterThan([(0, 1.0), (1, 1.0)], 2.0)
    b.equalityConstraint([(0, 1.0), (1, -1.0)], 0.0)
    return b, 2.0, [1.0, 1.0], "Optimal"


def _free_sign():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 1.0).objectiveCoefficient(1, -1.0)
    b.bounds(0, -1.0, 1.0).bounds(1, -1.0, 1.0)
    b.equalityConstraint([(0, 1.0), (1, 1.0)], 0.0)
    return b, -2.0, [-1.0, 1.0], "Optimal"


def _range_row():
    b = prima.builder(2)
    b.objectiveCoefficient(0, 1.0).bounds(0, -10.0, 10.0).bounds(1, 0.0, 0.0)
    b.constraint([(0, 1.0), (1, 1.0)], 2.0, 5.0)
    return b, 2.0, [2.0, 0.0], "Optimal"


def _infeasible():
    b = prima.builder(1)
    b.objectiveCoefficient(0, 1.0).bounds(0, -INF, INF)
    b.greaterThan([(0, 1.0)], 1.0).lessThan([(0, 1.0)], 0.0)
    return b, None, None, "PrimalInfeasible"


def _unbounded():
    b = prima.builder(1)
    b.objectiveCoefficient(0, -1.0).bounds(0, 0.0, INF).greaterThan([(0, 1.0)], 0.0)
    return b, None, None, "DualInfeasible"


def _economic_dispatch():
    loads_a, loads_b = [100.0, 80.0], [50.0, 70.0]
    costs, caps, line = [10.0, 30.0, 20.0], [120.0, 100.0, 90.0], 30.0
    b = prima.builder(8)
    for t in range(2):
        base = 4 * t
        for g in range(3):
            b.objectiveCoefficient(base + g, costs[g]).bounds(base + g, 0.0, caps[g])
        b.bounds(base + 3, -line, line)
        b.equalityConstraint([(base, 1.0), (base + 1, 1.0), (base + 3, -1.0)], loads_a[t])
        b.equalityConstraint([(base + 2, 1.0), (base + 3, 1.0)], loads_b[t])
    return b, 3700.0, [120.0, 0.0, 30.0, 20.0, 110.0, 0.0, 40.0, 30.0], "Optimal"


FIXTURES = {
    "tiny-lower-bound": _tiny_lower_bound,
    "product-mix": _product_mix,
    "equality-split": _equality_split,
    "degenerate": _degenerate,
    "free-sign": _free_sign,
    "range-row": _range_row,
    "economic-dispatch": _economic_dispatch,
    "infeasible": _infeasible,
    "unbounded": _unbounded,
}
OPTIMAL = [k for k in FIXTURES if k not in ("infeasible", "unbounded")]


def fixture(name):
    """(LpProblem, expected objective, expected primal, expected status)."""
    b, obj, x, status = FIXTURES[name]()
    return b.build()[0], obj, x, status


# The reference's published ladder (modules/prima-core/NOTES.md:43-58),
# eps = 1e-9: objective (6 dp), ojAlgo objective, iterations, restarts.
LADDER = {
    "tiny-lower-bound": (3.0, 3.0, 64, 0),
    "product-mix": (-36.0, -36.0, 128, 1),
    "equality-split": (24.0, 24.0, 64, 0),
    "degenerate": (2.0, 2.0, 64, 0),
    "free-sign": (-2.0, -2.0, 2, 0),
    "range-row": (2.0, 2.0, 64, 0),
    "economic-dispatch": (3700.0, 3700.0, 256, 3),
    "infeasible": (None, None, 1088, 5),
    "unbounded": (None, None, 64, 0),
    "random-60x30": (-773.369199, -773.369199, 1280, 9),
    "random-200x120": (-1796.551031, -1796.551031, 6464, 18),
    "random-600x400": (-5285.522448, -5285.522445, 35392, 17),
}
RANDOM_LADDER = {"random-60x30": (1, 60, 10, 20, 0.25), "random-200x120": (2, 200, 40, 80, 0.10),
                 "random-600x400": (3, 600, 120, 280, 0.04)}
