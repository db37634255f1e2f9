"""Linear optimal power flow: a PyPSA network turned into one LP over all its
snapshots, solved on the device (network-lopf's Lopf.scala:69-561 with
Expansion.scala, Storage.scala, Stores.scala, Losses.scala and Cycles.scala,
and network-model's StandardTypes.scala).

Columns, snapshot by snapshot: generator output, passive branch flow `p0`,
controllable branch (link) flow `p0`, four columns per storage unit
(dispatch, store, state of charge, spill) and two per store (energy, signed
power); one capacity column per extendable entity comes first. Rows: bus
balance per (snapshot, bus), storage and store energy balances, state of
charge set points, Kirchhoff's voltage law per (snapshot, independent cycle
of the passive branches), then the capacity coupling rows and the
`primary_energy` global constraints. Sign conventions are PyPSA's: a branch
takes `p0` out of `bus0` and delivers `efficiency * p0` at every other port;
load is consumption on the right-hand side.

Features the reference refuses are refused here with the same reasons
(unmodelled components, dangling buses, set points other than
`state_of_charge_set`, quadratic costs, multi-period flags, annuitised or
modular capacity, committable-and-extendable generators, standing losses
outside [0, 1), zero-impedance cycles).
"""
from __future__ import annotations

import math
import os
from collections import deque
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import prima
from .network import (DATA_DIR, Network, Table, branch_ports, dangling_bus_references,
                      port_efficiency, read_table_csv, role, sub_networks)

DISPATCH = "StorageUnit-p_dispatch"
STORE = "StorageUnit-p_store"
SOC = "StorageUnit-state_of_charge"
SPILL = "StorageUnit-spill"
STORE_E = "Store-e"
STORE_P = "Store-p"
NO_SNAPSHOT = -1

#: PyPSA's nominal_attrs (Expansion.scala:51-58).
NOMINAL = {"Generator": "p_nom", "Line": "s_nom", "Transformer": "s_nom", "Link": "p_nom",
           "StorageUnit": "p_nom", "Store": "e_nom"}


class UnsupportedNetwork(RuntimeError):
    """Lopf.UnsupportedNetwork (Lopf.scala:67)."""


class UnknownType(RuntimeError):
    """StandardTypes.UnknownType (StandardTypes.scala:42)."""


def capacity_key(component: str) -> str:
    return f"{component}-{NOMINAL[component]}"


# --------------------------------------------------------------------------
# standard types (StandardTypes.scala:53-301)


def _typed(net: Network, component: str) -> Optional[Table]:
    t = net.table(component)
    if t is None or t.kinds.get("type") != "str":
        return None
    return t if any(v.strip() for v in t.static["type"]) else None


def _library(net: Network, component: str) -> Dict[str, Table]:
    """The shipped library overridden row by row by the network's own types
    (StandardTypes.merged, :262-267)."""
    spec = net.schema[component]
    shipped = read_table_csv(os.path.join(DATA_DIR, f"{spec.list_name}.csv"), spec)
    out = {name: shipped for name in shipped.ids}
    own = net.table(component)
    if own is not None:
        out.update({name: own for name in own.ids})
    return out


def _unknown(component, entity, name, known):
    return UnknownType(f"{component} '{entity}': type '{name}' is not defined by the network "
                       f"and is not one of the {len(known)} standard types shipped here")


def expand_standard_types(net: Network) -> Network:
    """StandardTypes.expand (:53-62): impedance of typed lines and
    transformers from PyPSA's type library. Lines: r, x divided by
    num_parallel, b = 2*pi*1e-9*f_nom*c*length multiplied by it.
    Transformers: r = vscr/100, x = sqrt((vsc/100)^2 - r^2),
    g = pfe/(1000 s_nom), b = -sqrt((i0/100)^2 - g^2), the series terms
    divided and the shunt terms multiplied by num_parallel, s_nom from the
    type, tap_ratio = 1 + (tap_position - tap_neutral) * tap_step/100."""
    lines, trafos = _typed(net, "Line"), _typed(net, "Transformer")
    out = net
    if lines is not None:
        lib = _library(net, "LineType")
        r, x, b = [], [], []
        for e in lines.ids:
            name = lines.string("type", e).strip()
            if not name:
                r.append(lines.float("r", e)), x.append(lines.float("x", e))
                b.append(lines.float("b", e))
                continue
            if name not in lib:
                raise _unknown("Line", e, name, lib)
            t = lib[name]
            length, par = lines.float("length", e), lines.float("num_parallel", e)
            if not par > 0.0:
                raise UnknownType(f"Line '{e}': num_parallel = {par} cannot divide the "
                                  "impedance of its line type")
            r.append(t.float("r_per_length", name) * length / par)
            x.append(t.float("x_per_length", name) * length / par)
            b.append(2.0 * math.pi * 1e-9 * t.float("f_nom", name) * t.float("c_per_length", name)
                     * length * par)
        out = out.with_table(lines.with_floats({"r": r, "x": x, "b": b}))
    if trafos is not None:
        lib = _library(out, "TransformerType")
        keys = ("r", "x", "g", "b", "s_nom", "tap_ratio", "phase_shift")
        cols = {k: [] for k in keys}
        sides = []
        for e in trafos.ids:
            name = trafos.string("type", e).strip()
            if not name:
                for k in keys:
                    cols[k].append(trafos.float(k, e))
                sides.append(trafos.int("tap_side", e))
                continue
            if name not in lib:
                raise _unknown("Transformer", e, name, lib)
            t = lib[name]
            par = trafos.float("num_parallel", e)
            if not par > 0.0:
                raise UnknownType(f"Transformer '{e}': num_parallel {par} must be positive to "
                                  "scale its standard type's impedance")
            s_nom = t.float("s_nom", name)
            if not s_nom > 0.0:
                raise UnknownType(f"transformer type '{name}' has s_nom = {s_nom}")
            r = t.float("vscr", name) / 100.0
            vsc = t.float("vsc", name) / 100.0
            x = math.sqrt(max(0.0, vsc * vsc - r * r))
            g = t.float("pfe", name) / (1000.0 * s_nom)
            i0 = t.float("i0", name) / 100.0
            bb = -math.sqrt(max(0.0, i0 * i0 - g * g))
            if "tap_position" in trafos.static:
                pos = float(trafos.static["tap_position"][trafos.row(e)])
            else:
                pos = float(trafos.spec.attribute("tap_position").default)
            for k, v in zip(keys, (r / par, x / par, g * par, bb * par, s_nom,
                                   1.0 + (pos - t.int("tap_neutral", name))
                                   * (t.float("tap_step", name) / 100.0),
                                   t.float("phase_shift", name))):
                cols[k].append(v)
            sides.append(t.int("tap_side", name))
        nt = trafos.with_floats(cols)
        nt.static["tap_side"] = sides
        nt.kinds["tap_side"] = "int"
        out = out.with_table(nt)
    return out


# --------------------------------------------------------------------------
# Kirchhoff cycles (Cycles.scala:27-196)


def cycle_basis(net: Network) -> List[Tuple[str, List[Tuple[str, str, float]]]]:
    """A fundamental cycle basis of the passive branches: a breadth-first
    spanning tree per island, and for each branch outside it the cycle that
    branch closes, traversed bus0 -> bus1 along the branch and back through
    the tree. A branch carries +1 where the traversal runs bus0 -> bus1 along
    it and -1 the other way. That gives edges - buses + islands independent
    cycles, the circuit rank; any basis spans the same constraint set.
    Returns [(carrier of the island, [(component, id, orientation)])]."""
    edges = []
    for table in net.tables.values():
        if role(table.spec) == "passive":
            for e in table.ids:
                edges.append((table.name, e, table.string("bus0", e), table.string("bus1", e)))
    if not edges:
        return []
    adj: Dict[str, List[int]] = {}
    for k, (_, _, b0, b1) in enumerate(edges):
        adj.setdefault(b0, []).append(k)
        adj.setdefault(b1, []).append(k)
    carrier_of = {b: carrier for carrier, buses in sub_networks(net) for b in buses}
    parent: Dict[str, int] = {}
    depth: Dict[str, int] = {}
    in_tree = [False] * len(edges)
    island_of: Dict[str, str] = {}
    buses = net.table("Bus")
    for root in (buses.ids if buses else []):
        if root in depth or root not in adj:
            continue
        depth[root] = 0
        island_of[root] = root
        queue = deque([root])
        while queue:
            bus = queue.popleft()
            for k in adj[bus]:
                _, _, b0, b1 = edges[k]
                nxt = b1 if bus == b0 else b0
                if nxt not in depth:
                    depth[nxt] = depth[bus] + 1
                    parent[nxt] = k
                    island_of[nxt] = root
                    in_tree[k] = True
                    queue.append(nxt)
    cycles = []
    for k, (comp, e, b0, b1) in enumerate(edges):
        if in_tree[k]:
            continue
        terms = [(comp, e, 1.0)]
        up, down = [], []
        a, b = b1, b0
        # from bus1 up towards the common ancestor: travel node -> parent
        # from bus0's side the travel is parent -> node, so the sense flips
        while depth[a] > depth[b]:
            pe = edges[parent[a]]
            up.append((pe[0], pe[1], 1.0 if pe[2] == a else -1.0))
            a = pe[3] if pe[2] == a else pe[2]
        while depth[b] > depth[a]:
            pe = edges[parent[b]]
            down.append((pe[0], pe[1], 1.0 if pe[3] == b else -1.0))
            b = pe[3] if pe[2] == b else pe[2]
        while a != b:
            pa = edges[parent[a]]
            up.append((pa[0], pa[1], 1.0 if pa[2] == a else -1.0))
            a = pa[3] if pa[2] == a else pa[2]
            pb = edges[parent[b]]
            down.append((pb[0], pb[1], 1.0 if pb[3] == b else -1.0))
            b = pb[3] if pb[2] == b else pb[2]
        cycles.append((carrier_of.get(island_of[b0], ""), terms + up + down))
    return cycles


def tap_ratio(table: Table, entity: str) -> float:
    """Branches.tapRatio (network-pf Branches.scala:113-117)."""
    if table.spec.attribute("tap_ratio") is None and "tap_ratio" not in table.static:
        return 1.0
    v = table.float("tap_ratio", entity)
    return v if math.isfinite(v) and v > 0.0 else 1.0


def impedance(net: Network, component: str, entity: str, carrier: str) -> float:
    """Cycles.impedance (:168-192): r on a DC island, x otherwise; a line
    referred to v_nom(bus0)^2, a transformer to s_nom / tap_ratio."""
    table = net.require(component)
    raw = table.float("r" if carrier == "DC" else "x", entity)
    if component == "Line":
        v = net.require("Bus").float("v_nom", table.string("bus0", entity))
        return raw / (v * v) if v > 0.0 else raw
    if component == "Transformer":
        s_nom = table.float("s_nom", entity)
        if not s_nom > 0.0:
            raise UnsupportedNetwork(f"Transformer '{entity}': s_nom = {s_nom} cannot serve as "
                                     "the base of its per-unit impedance")
        return raw * tap_ratio(table, entity) / s_nom
    raise UnsupportedNetwork(f"{component} '{entity}': no per-unit impedance base is defined "
                             "for this kind of passive branch")


# --------------------------------------------------------------------------
# expansion (Expansion.scala:60-201)


def is_extendable(table: Table, entity: str) -> bool:
    attr = NOMINAL.get(table.name)
    if attr is None:
        return False
    flag = f"{attr}_extendable"
    return flag in table.static and table.bool(flag, entity)


def extendables(table: Table) -> List[str]:
    if table.name not in NOMINAL:
        return []
    return [e for e in table.ids if is_extendable(table, e)]


def periodized_cost(table: Table, entity: str) -> float:
    """capital_cost + fom_cost, non-finite parts read as 0."""
    cap = table.float("capital_cost", entity)
    fom = table.float("fom_cost", entity) if "fom_cost" in table.static else 0.0
    return (cap if math.isfinite(cap) else 0.0) + (fom if math.isfinite(fom) else 0.0)


def objective_constant(net: Network) -> float:
    """The capital cost of existing extendable capacity, which PyPSA
    subtracts from its objective (Expansion.scala:141-146)."""
    total = 0.0
    for component in sorted(NOMINAL):
        t = net.table(component)
        if t is None:
            continue
        for e in extendables(t):
            total += periodized_cost(t, e) * t.float(NOMINAL[component], e)
    return total


def operational_bounds(table: Table, entity: str, t: int):
    """(variable, min_pu, max_pu) per column a capacity bounds (:76-110)."""
    name = table.name
    if name in ("Line", "Transformer"):
        m = table.value_at("s_max_pu", entity, t)
        return [(name, -m, m)]
    if name == "StorageUnit":
        return [(DISPATCH, -math.inf, table.value_at("p_max_pu", entity, t)),
                (STORE, -math.inf, -table.value_at("p_min_pu", entity, t)),
                (SOC, -math.inf, table.float("max_hours", entity))]
    if name == "Store":
        return [(STORE_E, table.value_at("e_min_pu", entity, t),
                 table.value_at("e_max_pu", entity, t))]
    return [(name, table.value_at("p_min_pu", entity, t), table.value_at("p_max_pu", entity, t))]


def _reject_expansion(net: Network):
    for component, attr in NOMINAL.items():
        t = net.table(component)
        if t is None:
            continue
        for e in extendables(t):
            if "overnight_cost" in t.static and math.isfinite(t.float("overnight_cost", e)):
                raise UnsupportedNetwork(f"{component} '{e}': overnight_cost is given, but PyPSA would "
                                         "turn it into an annuity first; give capital_cost instead")
            mod = f"{attr}_mod"
            if mod in t.static and t.float(mod, e) != 0.0:
                raise UnsupportedNetwork(f"{component} '{e}': {mod} makes the capacity a whole "
                                         "number of modules, which needs integer variables")
            lo, hi = t.float(f"{attr}_min", e), t.float(f"{attr}_max", e)
            if lo > hi:
                raise UnsupportedNetwork(f"{component} '{e}' has {attr}_min = {lo} above "
                                         f"{attr}_max = {hi}")
            if not lo >= 0.0:
                raise UnsupportedNetwork(f"{component} '{e}' has {attr}_min = {lo}, which is "
                                         "not a capacity")
    g = net.table("Generator")
    if g is not None and "committable" in g.static:
        for e in g.ids:
            if g.bool("committable", e) and is_extendable(g, e):
                raise UnsupportedNetwork(f"Generator '{e}': a committable unit cannot also have "
                                         "an extendable capacity here")


# --------------------------------------------------------------------------
# storage guards (Storage.scala:53-114, Stores.scala:47-89, Losses.scala)


def _pinned(table: Table, attribute: str) -> List[str]:
    s = table.series.get(attribute)
    out = []
    for e in table.ids:
        static = attribute in table.static and math.isfinite(table.float(attribute, e))
        if (s is not None and s.covers(e)) or static:
            out.append(e)
    return out


def _flagged(table: Table, attribute: str) -> List[str]:
    return [e for e in table.ids if attribute in table.static and table.bool(attribute, e)]


def _quadratic(table: Table) -> List[str]:
    a = "marginal_cost_quadratic"
    s = table.series.get(a)
    return [e for e in table.ids if (a in table.static and table.float(a, e) != 0.0)
            or (s is not None and s.covers(e))]


def _reject_standing_loss(table: Table, component: str):
    def check(e, v, where=""):
        if not (v >= 0.0 and v < 1.0):
            raise UnsupportedNetwork(f"{component} '{e}': standing_loss = {v}{where} must lie "
                                     "in [0, 1)")
    s = table.series.get("standing_loss")
    for e in table.ids:
        if s is None or not s.covers(e):
            check(e, table.float("standing_loss", e))
    if s is not None:
        for j, e in enumerate(s.entities):
            for t in range(s.values.shape[0]):
                check(e, float(s.values[t, j]), f" at snapshot {t}")


def _reject_storage(table: Table):
    for a in ("p_set", "p_dispatch_set", "p_store_set"):
        p = _pinned(table, a)
        if p:
            raise UnsupportedNetwork(f"StorageUnit '{p[0]}': {a} fixes its power in a snapshot, "
                                     "and no such equality is generated")
    for a in ("cyclic_state_of_charge_per_period", "state_of_charge_initial_per_period"):
        f = _flagged(table, a)
        if f:
            raise UnsupportedNetwork(f"StorageUnit '{f[0]}': {a} belongs to PyPSA's investment "
                                     "periods, and the LP here spans one period only")
    q = _quadratic(table)
    if q:
        raise UnsupportedNetwork(f"StorageUnit '{q[0]}': marginal_cost_quadratic is non-zero, "
                                 "and a quadratic cost does not fit a linear program")
    for e in table.ids:
        mh = table.float("max_hours", e)
        if not mh > 0.0:
            raise UnsupportedNetwork(f"StorageUnit '{e}' has max_hours = {mh}, so it can store "
                                     "nothing")
    _reject_standing_loss(table, "StorageUnit")


def _reject_stores(table: Table):
    for a in ("e_set", "p_set"):
        p = _pinned(table, a)
        if p:
            raise UnsupportedNetwork(f"Store '{p[0]}' sets {a}; that constraint is not built here")
    for a in ("e_cyclic_per_period", "e_initial_per_period"):
        f = _flagged(table, a)
        if f:
            raise UnsupportedNetwork(f"Store '{f[0]}' sets {a}, a multi-investment-period concept")
    q = _quadratic(table)
    if q:
        raise UnsupportedNetwork(f"Store '{q[0]}': marginal_cost_quadratic is non-zero, and a "
                                 "quadratic cost does not fit a linear program")
    _reject_standing_loss(table, "Store")


HANDLED = {"Generator", "Load", "Line", "Transformer", "Link", "Bus", "Carrier",
           "GlobalConstraint", "SubNetwork", "LineType", "TransformerType", "Shape",
           "StorageUnit", "Store"}


# --------------------------------------------------------------------------
# the model


@dataclass
class Model:
    """Lopf.Model (Lopf.scala:64): the LP, how its rows map back to the rows
    as written, and where each variable and balance row sits."""

    problem: prima.LpProblem
    translation: prima.RowTranslation
    columns: Dict[Tuple[str, str, int], int]
    balance_rows: Dict[Tuple[str, int], int]

    @property
    def num_variables(self) -> int:
        return self.problem.n

    def column(self, component: str, entity: str, snapshot: int) -> int:
        try:
            return self.columns[(component, entity, snapshot)]
        except KeyError:
            raise KeyError(f"no variable for {component} '{entity}' at snapshot {snapshot}") from None


def build(network: Network) -> Model:
    """Lopf.build (Lopf.scala:71-474)."""
    net = expand_standard_types(network)
    unhandled = [t for t in net.tables.values() if t.size > 0 and t.name not in HANDLED]
    if unhandled:
        raise UnsupportedNetwork("the LP builder has no model for these components: "
                                 + ", ".join(f"{t.name} ({t.size})" for t in unhandled))
    dangling = dangling_bus_references(net)
    if dangling:
        c, e, port, bus = dangling[0]
        raise UnsupportedNetwork(f"{c} '{e}': its {port} names bus '{bus}', which the network "
                                 "does not define")
    _reject_expansion(net)
    T = len(net.snapshots)
    if T == 0:
        raise UnsupportedNetwork("the network defines no snapshot to dispatch")
    bus_table = net.table("Bus")
    buses = bus_table.ids if bus_table else []
    if not buses:
        raise UnsupportedNetwork("the network defines no bus")

    def tables_with(r):
        return [t for t in net.tables.values() if role(t.spec) == r and t.size > 0]

    passive, controllable, attached = (tables_with("passive"), tables_with("controllable"),
                                       tables_with("attached"))
    generators = [t for t in attached if t.name == "Generator"]
    load_tables = [t for t in attached if t.name == "Load"]
    storage = next((t for t in attached if t.name == "StorageUnit"), None)
    stores = next((t for t in attached if t.name == "Store"), None)
    if storage is not None:
        _reject_storage(storage)
    if stores is not None:
        _reject_stores(stores)

    columns: Dict[Tuple[str, str, int], int] = {}
    lo: List[float] = []
    hi: List[float] = []
    cost: List[float] = []

    def declare(component, entity, t, l, h, c):
        columns[(component, entity, t)] = len(lo)
        lo.append(l)
        hi.append(h)
        cost.append(c)

    expandable = [(t, extendables(t)) for t in passive + controllable + attached]
    expandable = [(t, ids) for t, ids in expandable if ids]
    for t, ids in expandable:
        attr = NOMINAL[t.name]
        for e in ids:
            declare(capacity_key(t.name), e, NO_SNAPSHOT, t.float(f"{attr}_min", e),
                    t.float(f"{attr}_max", e), periodized_cost(t, e))

    inf = math.inf
    for s in range(T):
        w = net.weighting("objective", s)
        for g in generators:
            for e in g.ids:
                c = g.value_at("marginal_cost", e, s) * w
                if is_extendable(g, e):
                    declare(g.name, e, s, -inf, inf, c)
                else:
                    p_nom = g.float("p_nom", e)
                    a, b = p_nom * g.value_at("p_min_pu", e, s), p_nom * g.value_at("p_max_pu", e, s)
                    declare(g.name, e, s, min(a, b), max(a, b), c)
        for br in passive:
            for e in br.ids:
                if is_extendable(br, e):
                    declare(br.name, e, s, -inf, inf, 0.0)
                else:
                    lim = br.float("s_nom", e) * br.value_at("s_max_pu", e, s)
                    declare(br.name, e, s, -lim, lim, 0.0)
        for br in controllable:
            for e in br.ids:
                c = br.value_at("marginal_cost", e, s) * w
                if is_extendable(br, e):
                    declare(br.name, e, s, -inf, inf, c)
                else:
                    p_nom = br.float("p_nom", e)
                    a, b = (p_nom * br.value_at("p_min_pu", e, s),
                            p_nom * br.value_at("p_max_pu", e, s))
                    declare(br.name, e, s, min(a, b), max(a, b), c)
        if storage is not None:
            for e in storage.ids:
                ext = is_extendable(storage, e)
                p_nom = storage.float("p_nom", e)

                def upper(per_unit):
                    return inf if ext else p_nom * per_unit

                declare(DISPATCH, e, s, 0.0, upper(storage.value_at("p_max_pu", e, s)),
                        storage.value_at("marginal_cost", e, s) * w)
                declare(STORE, e, s, 0.0, upper(-storage.value_at("p_min_pu", e, s)), 0.0)
                declare(SOC, e, s, 0.0, upper(storage.float("max_hours", e)),
                        storage.value_at("marginal_cost_storage", e, s) * w)
                inflow = storage.value_at("inflow", e, s)
                declare(SPILL, e, s, 0.0, math.nan if math.isnan(inflow) else max(0.0, inflow),
                        storage.value_at("spill_cost", e, s) * w)
        if stores is not None:
            for e in stores.ids:
                if is_extendable(stores, e):
                    a, b = -inf, inf
                else:
                    e_nom = stores.float("e_nom", e)
                    a = e_nom * stores.value_at("e_min_pu", e, s)
                    b = e_nom * stores.value_at("e_max_pu", e, s)
                declare(STORE_E, e, s, min(a, b), max(a, b),
                        stores.value_at("marginal_cost_storage", e, s) * w)
                declare(STORE_P, e, s, -inf, inf, stores.value_at("marginal_cost", e, s) * w)

    b = prima.builder(len(lo))
    for i in range(len(lo)):
        b.bounds(i, lo[i], hi[i])
        b.objectiveCoefficient(i, cost[i])
    if expandable:
        b.objectiveOffset(-objective_constant(net))

    # who touches each bus, in the order the balance lists them
    gens_at: Dict[str, list] = {}
    for g in generators:
        for e in g.ids:
            gens_at.setdefault(g.string("bus", e), []).append((g.name, e))
    storage_at: Dict[str, list] = {}
    if storage is not None:
        for e in storage.ids:
            storage_at.setdefault(storage.string("bus", e), []).append(e)
    stores_at: Dict[str, list] = {}
    if stores is not None:
        for e in stores.ids:
            stores_at.setdefault(stores.string("bus", e), []).append(e)
    ports_at: Dict[str, list] = {}
    for br in passive + controllable:
        ports = branch_ports(br)
        for e in br.ids:
            for port in ports:
                ports_at.setdefault(br.string(port, e), []).append((br, e, port))
    loads_at: Dict[str, list] = {}
    for lt in load_tables:
        for e in lt.ids:
            loads_at.setdefault(lt.string("bus", e), []).append((lt, e))

    balance_rows: Dict[Tuple[str, int], int] = {}
    row = 0
    for s in range(T):
        for bus in buses:
            terms = [(columns[(name, e, s)], 1.0) for name, e in gens_at.get(bus, ())]
            for e in storage_at.get(bus, ()):
                terms.append((columns[(DISPATCH, e, s)], 1.0))
                terms.append((columns[(STORE, e, s)], -1.0))
            for e in stores_at.get(bus, ()):
                terms.append((columns[(STORE_P, e, s)], 1.0))
            for br, e, port in ports_at.get(bus, ()):
                c = columns[(br.name, e, s)]
                terms.append((c, -1.0 if port == "bus0" else port_efficiency(br, e, port, s)))
            demand = 0.0
            for lt, e in loads_at.get(bus, ()):
                demand += lt.value_at("p_set", e, s)
            b.equalityConstraint(terms, demand)
            balance_rows[(bus, s)] = row
            row += 1

    if storage is not None:
        for e in storage.ids:
            cyclic = "cyclic_state_of_charge" in storage.static and storage.bool(
                "cyclic_state_of_charge", e)
            initial = storage.float("state_of_charge_initial", e)
            for s in range(T):
                eh = net.weighting("stores", s)
                eff_d = storage.value_at("efficiency_dispatch", e, s)
                if not eff_d > 0.0:
                    raise UnsupportedNetwork(f"StorageUnit '{e}': efficiency_dispatch {eff_d} in snapshot "
                                             f"{s} is not positive, and the state-of-charge row "
                                             "divides by it")
                stand = math.pow(1.0 - storage.value_at("standing_loss", e, s), eh)
                terms = [(columns[(SOC, e, s)], 1.0),
                         (columns[(DISPATCH, e, s)], eh / eff_d),
                         (columns[(STORE, e, s)], -eh * storage.value_at("efficiency_store", e, s)),
                         (columns[(SPILL, e, s)], eh)]
                prev = s - 1 if s > 0 else (T - 1 if cyclic else None)
                if prev is not None:
                    terms.append((columns[(SOC, e, prev)], -stand))
                rhs = eh * storage.value_at("inflow", e, s) + (initial if prev is None else 0.0)
                b.equalityConstraint(terms, rhs)
                target = storage.value_at("state_of_charge_set", e, s)
                if math.isfinite(target):
                    b.equalityConstraint([(columns[(SOC, e, s)], 1.0)], target)

    cycles = cycle_basis(net)
    imp = {}
    for carrier, terms in cycles:
        for comp, e, _ in terms:
            imp[(comp, e)] = impedance(net, comp, e, carrier)
    for carrier, terms in cycles:
        if not any(imp[(comp, e)] != 0.0 for comp, e, _ in terms):
            raise UnsupportedNetwork(
                f"a cycle of the {carrier} sub-network has zero impedance on all of its branches ("
                + ", ".join(f"{c} '{e}'" for c, e, _ in terms)
                + "): its voltage-law row would be empty")
    for s in range(T):
        for carrier, terms in cycles:
            b.equalityConstraint([(columns[(comp, e, s)], o * imp[(comp, e)])
                                  for comp, e, o in terms], 0.0)

    if stores is not None:
        for e in stores.ids:
            cyclic = "e_cyclic" in stores.static and stores.bool("e_cyclic", e)
            initial = stores.float("e_initial", e)
            for s in range(T):
                eh = net.weighting("stores", s)
                stand = math.pow(1.0 - stores.value_at("standing_loss", e, s), eh)
                terms = [(columns[(STORE_E, e, s)], 1.0), (columns[(STORE_P, e, s)], eh)]
                prev = s - 1 if s > 0 else (T - 1 if cyclic else None)
                if prev is not None:
                    terms.append((columns[(STORE_E, e, prev)], -stand))
                b.equalityConstraint(terms, initial if prev is None else 0.0)

    for t, ids in expandable:
        key = capacity_key(t.name)
        for e in ids:
            cap = columns[(key, e, NO_SNAPSHOT)]
            for s in range(T):
                for var, min_pu, max_pu in operational_bounds(t, e, s):
                    col = columns[(var, e, s)]
                    if math.isfinite(max_pu):
                        b.lessThan([(col, 1.0), (cap, -max_pu)], 0.0)
                    if math.isfinite(min_pu):
                        b.greaterThan([(col, 1.0), (cap, -min_pu)], 0.0)

    gc = net.table("GlobalConstraint")
    if gc is not None:
        for cid in gc.ids:
            sense, constant = gc.string("sense", cid), gc.float("constant", cid)
            kind = gc.string("type", cid)
            if kind != "primary_energy":
                raise UnsupportedNetwork(f"global constraint '{cid}': type '{kind}' is not "
                                         "built (the builder knows 'primary_energy')")
            if sense != "<=":
                raise UnsupportedNetwork(f"global constraint '{cid}' has sense '{sense}'; only "
                                         "'<=' is implemented")
            attribute = gc.string("carrier_attribute", cid)
            terms = []
            for s in range(T):
                w = net.weighting("generators", s)
                for g in generators:
                    for e in g.ids:
                        intensity = _carrier_attribute(net, g.string("carrier", e), attribute)
                        eff = g.value_at("efficiency", e, s)
                        if intensity != 0.0 and eff != 0.0:
                            terms.append((columns[(g.name, e, s)], intensity / eff * w))
            if terms:
                b.lessThan(terms, constant)

    problem, translation = b.build()
    return Model(problem, translation, columns, balance_rows)


def _carrier_attribute(net: Network, carrier: str, attribute: str) -> float:
    """Lopf.carrierAttribute (Lopf.scala:551-557): 0 for an unknown carrier,
    an absent column or a non-finite value."""
    t = net.table("Carrier")
    if t is None or not t.has(carrier):
        return 0.0
    if t.spec.attribute(attribute) is None and attribute not in t.static:
        return 0.0
    v = t.float(attribute, carrier)
    return v if math.isfinite(v) else 0.0


# --------------------------------------------------------------------------
# results


class LopfResult:
    """LopfResult (Lopf.scala:563-705): a solution addressed by component
    name. `network` is the type-expanded network the model was built from."""

    def __init__(self, network: Network, model: Model, solution: prima.LpSolution):
        self.network = network
        self.model = model
        self.solution = solution

    @property
    def status(self) -> str:
        return self.solution.status

    @property
    def objective(self) -> float:
        return self.solution.objective_value

    @property
    def total_system_cost(self) -> float:
        return self.objective + objective_constant(self.network)

    def _primal(self, var, entity, s):
        return float(self.solution.primal[self.model.column(var, entity, s)])

    def dispatch(self, component: str, entity: str, s: int) -> float:
        if component == "StorageUnit":
            return self._primal(DISPATCH, entity, s) - self._primal(STORE, entity, s)
        if component == "Store":
            return self._primal(STORE_P, entity, s)
        return self._primal(component, entity, s)

    def energy(self, entity: str, s: int) -> float:
        return self._primal(STORE_E, entity, s)

    def state_of_charge(self, entity: str, s: int) -> float:
        return self._primal(SOC, entity, s)

    def charging(self, entity: str, s: int) -> float:
        return self._primal(STORE, entity, s)

    def discharging(self, entity: str, s: int) -> float:
        return self._primal(DISPATCH, entity, s)

    def spill(self, entity: str, s: int) -> float:
        return self._primal(SPILL, entity, s)

    def capacity(self, component: str, entity: str) -> float:
        t = self.network.require(component)
        if not is_extendable(t, entity):
            return t.float(NOMINAL[component], entity)
        return self._primal(capacity_key(component), entity, NO_SNAPSHOT)

    def marginal_price(self, bus: str, s: int) -> float:
        """The balance row's dual, as written, per unit of objective weight."""
        row = self.model.balance_rows[(bus, s)]
        duals = self.model.translation.original_duals(self.solution.dual)
        return float(duals[row]) / self.network.weighting("objective", s)


def solve(network: Network, kernels, params: Optional[prima.PdhgParams] = None) -> LopfResult:
    """Lopf.solve (Lopf.scala:484-488) on the device: `kernels` is an open
    HipKernels context (f64 for the plain solve)."""
    net = expand_standard_types(network)
    model = build(net)
    return LopfResult(net, model, kernels.solve(model.problem, params or prima.PdhgParams()))


def solve_all(networks: Sequence[Network], kernels,
              params: Optional[prima.PdhgParams] = None) -> List[LopfResult]:
    """Every network's LP in one device batch, results in input order."""
    nets = [expand_standard_types(n) for n in networks]
    models = [build(n) for n in nets]
    sols = kernels.solve_all([m.problem for m in models], params or prima.PdhgParams())
    return [LopfResult(n, m, s) for n, m, s in zip(nets, models, sols)]
