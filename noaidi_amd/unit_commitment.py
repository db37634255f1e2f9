"""Unit commitment over a PyPSA network, solved by the device branch-and-bound
(network-lopf's UnitCommitment.scala:72-425).

Per committable generator g and snapshot t: a binary status u, output p and
continuous start-up / shut-down indicators su, sd (driven to 0/1 by the rows
whenever u is integral, and priced non-negatively, so they need no branching):

    p_min_pu p_nom u <= p <= p_max_pu p_nom u
    su - sd = u[t] - u[t-1]            (u[-1] from up_time_before)
    u[t] = initial status for the first (min_up - up_before) or
           (min_down - down_before) snapshots
    sum su over the last min_up snapshots   <= u[t]
    sum sd over the last min_down snapshots <= 1 - u[t]

Bus balance holds per bus with no flows, so a network with any branch is
refused, as are global constraints, storage, ramp limits, stand-by and
quadratic costs and extendable generators (leaving any of them out would make
the schedule look cheaper than it is).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import numpy as np

from . import prima
from .lopf import expand_standard_types
from .network import Network, dangling_bus_references


class UnsupportedNetwork(RuntimeError):
    """UnitCommitment.UnsupportedNetwork (UnitCommitment.scala:74)."""


HANDLED = {"Generator", "Load", "Bus", "Carrier", "SubNetwork", "LineType", "TransformerType",
           "Shape"}


def _committable(g, e) -> bool:
    return (g.spec.attribute("committable") is not None and "committable" in g.static
            and g.bool("committable", e))


def _reject(net: Network):
    """UnitCommitment.reject (UnitCommitment.scala:327-424)."""
    gc = net.table("GlobalConstraint")
    if gc is not None and gc.size > 0:
        raise UnsupportedNetwork(
            f"network has {gc.size} global constraint(s): emission and primary-energy caps have "
            "no rows in this model, and leaving one out would make the schedule look cheaper "
            "than it is")
    unhandled = [t for t in net.tables.values() if t.size > 0 and t.name not in HANDLED]
    if unhandled:
        raise UnsupportedNetwork(
            "unit commitment is formulated without branch flows, so it models a single bus; this "
            "network carries " + ", ".join(f"{t.name} ({t.size})" for t in unhandled))
    dangling = dangling_bus_references(net)
    if dangling:
        c, e, port, bus = dangling[0]
        raise UnsupportedNetwork(f"{c} '{e}': its {port} names bus '{bus}', which the network "
                                 "does not define")
    g = net.table("Generator")
    if g is None:
        return
    for attribute in ("stand_by_cost", "marginal_cost_quadratic"):
        if attribute not in g.static and attribute not in g.series:
            continue
        for e in g.ids:
            checks = [("statically", g.float(attribute, e))] + [
                (f"at snapshot {t}", g.value_at(attribute, e, t)) for t in range(len(net.snapshots))]
            for where, v in checks:
                if math.isfinite(v) and v != 0.0:
                    raise UnsupportedNetwork(
                        f"generator '{e}' has {attribute} = {v} {where}; it enters PyPSA's "
                        "objective and not this one, so ignoring it would price the schedule "
                        "below the truth")
    for attribute in ("ramp_limit_up", "ramp_limit_down", "ramp_limit_start_up",
                      "ramp_limit_shut_down"):
        if attribute in g.static:
            for e in g.ids:
                v = g.float(attribute, e)
                if math.isfinite(v):
                    raise UnsupportedNetwork(f"generator '{e}': {attribute} = {v} asks for a "
                                             "ramp constraint, which this model does not build")
    if "p_nom_extendable" in g.static:
        for e in g.ids:
            if g.bool("p_nom_extendable", e):
                raise UnsupportedNetwork(f"generator '{e}': p_nom_extendable is set, but the "
                                         "commitment model needs a fixed capacity")


@dataclass
class UcModel:
    problem: prima.LpProblem
    integers: list
    columns: Dict[Tuple[str, str, int], int]
    committable: Dict[Tuple[str, int], bool]


def build(network: Network) -> UcModel:
    """The MILP of UnitCommitment.solve (UnitCommitment.scala:109-289)."""
    net = expand_standard_types(network)
    _reject(net)
    T = len(net.snapshots)
    if T == 0:
        raise UnsupportedNetwork("the network defines no snapshot to schedule")
    g = net.require("Generator")
    buses = net.require("Bus").ids
    units = []
    for e in g.ids:
        if not _committable(g, e):
            continue
        up_before, down_before = g.int("up_time_before", e), g.int("down_time_before", e)
        min_up, min_down = g.int("min_up_time", e), g.int("min_down_time", e)
        up = up_before > 0
        forced = max(min_up - up_before, 0) if up else max(min_down - down_before, 0)
        units.append(dict(id=e, start=g.float("start_up_cost", e), stop=g.float("shut_down_cost", e),
                          min_up=min_up, min_down=min_down, up=up, forced=forced))
    limits = {}
    for t in range(T):
        for e in g.ids:
            p_nom = g.float("p_nom", e)
            limits[(e, t)] = (p_nom * g.value_at("p_min_pu", e, t),
                              p_nom * g.value_at("p_max_pu", e, t), _committable(g, e))
    columns: Dict[Tuple[str, str, int], int] = {}
    lo, hi, cost = [], [], []

    def declare(kind, e, t, l, h, c):
        columns[(kind, e, t)] = len(lo)
        lo.append(l)
        hi.append(h)
        cost.append(c)

    for t in range(T):
        w = net.weighting("objective", t)
        for e in g.ids:
            a, b, com = limits[(e, t)]
            c = g.value_at("marginal_cost", e, t) * w
            if com:
                declare("p", e, t, 0.0, max(b, 0.0), c)
            else:
                declare("p", e, t, min(a, b), max(a, b), c)
        for u in units:
            declare("u", u["id"], t, 0.0, 1.0, 0.0)
            declare("su", u["id"], t, 0.0, 1.0, u["start"])
            declare("sd", u["id"], t, 0.0, 1.0, u["stop"])
    bld = prima.builder(len(lo))
    for i in range(len(lo)):
        bld.bounds(i, lo[i], hi[i])
        bld.objectiveCoefficient(i, cost[i])
    loads = net.table("Load")
    gens_at: Dict[str, list] = {}
    for e in g.ids:
        gens_at.setdefault(g.string("bus", e), []).append(e)
    for t in range(T):
        for bus in buses:
            terms = [(columns[("p", e, t)], 1.0) for e in gens_at.get(bus, ())]
            demand = 0.0
            if loads is not None:
                for e in loads.ids:
                    if loads.string("bus", e) == bus:
                        demand += loads.value_at("p_set", e, t)
            if terms or demand != 0.0:
                bld.equalityConstraint(terms, demand)
    for t in range(T):
        for u in units:
            e = u["id"]
            p, on = columns[("p", e, t)], columns[("u", e, t)]
            su, sd = columns[("su", e, t)], columns[("sd", e, t)]
            a, b, _ = limits[(e, t)]
            bld.lessThan([(p, 1.0), (on, -b)], 0.0)
            bld.greaterThan([(p, 1.0), (on, -a)], 0.0)
            if t > 0:
                bld.equalityConstraint([(su, 1.0), (sd, -1.0), (on, -1.0),
                                        (columns[("u", e, t - 1)], 1.0)], 0.0)
            else:
                bld.equalityConstraint([(su, 1.0), (sd, -1.0), (on, -1.0)],
                                       -(1.0 if u["up"] else 0.0))
            if t < u["forced"]:
                bld.equalityConstraint([(on, 1.0)], 1.0 if u["up"] else 0.0)
            if u["min_up"] > 1:
                window = [(columns[("su", e, i)], 1.0)
                          for i in range(max(0, t - u["min_up"] + 1), t + 1)]
                bld.lessThan(window + [(on, -1.0)], 0.0)
            if u["min_down"] > 1:
                window = [(columns[("sd", e, i)], 1.0)
                          for i in range(max(0, t - u["min_down"] + 1), t + 1)]
                bld.lessThan(window + [(on, 1.0)], 1.0)
    problem, _ = bld.build()
    integers = sorted(i for (kind, _, _), i in columns.items() if kind == "u")
    return UcModel(problem, integers, columns,
                   {(e, t): limits[(e, t)][2] for t in range(T) for e in g.ids})


class UcResult:
    """UcResult (UnitCommitment.scala:8-39)."""

    def __init__(self, model: UcModel, solution: prima.MilpSolution):
        self.model = model
        self.solution = solution

    @property
    def status(self) -> str:
        return self.solution.status

    @property
    def objective(self) -> float:
        return self.solution.objective_value

    @property
    def best_bound(self) -> float:
        return self.solution.best_bound

    @property
    def nodes_explored(self) -> int:
        return self.solution.nodes_explored

    def output(self, generator: str, t: int) -> float:
        return float(self.solution.primal[self.model.columns[("p", generator, t)]])

    def committed(self, generator: str, t: int) -> bool:
        """A non-committable generator is always available: True."""
        if not self.model.committable[(generator, t)]:
            return True
        return bool(self.solution.primal[self.model.columns[("u", generator, t)]] > 0.5)


def solve(network: Network, kernels, params: Optional[prima.BnbParams] = None,
          mixed: Optional[prima.MixedParams] = None) -> UcResult:
    """UnitCommitment.solve with the node relaxations on the device."""
    m = build(network)
    return UcResult(m, kernels.milp_solve(m.problem, m.integers, params or prima.BnbParams(),
                                          mixed))
