"""The PyPSA reader and LOPF builder against LopfSuite (network-lopf
LopfSuite.scala:115-1089) and the reference's golden PyPSA results
(reference/goldens/results/*.json, kept as tests/golden/pypsa/optimize.json).

Parity is pinned three ways: the LP the builder makes is solved by scipy's
HiGHS and must reach PyPSA's own objective (so the model, not a solver, is
what is checked); the oracle's PDHG solves the small networks to the same
objectives; and the device solves every network, batched, to PyPSA's
objective and dispatch within LopfSuite's tolerances. The networks are
PyPSA 1.2.4's CSV exports, copied as data by tests/golden/make_golden_lopf.py.
"""
import json
import math
import os
import shutil

import numpy as np
import pytest
from scipy.optimize import linprog

from noaidi_amd import lopf, network, prima

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "pypsa")
with open(os.path.join(GOLD, "optimize.json")) as _f:
    OPTIMIZE = json.load(_f)
with open(os.path.join(GOLD, "highs.json")) as _f:
    HIGHS = json.load(_f)

# Every network whose PyPSA objective is the LP's. ac-pf-pv fixes g_pv's
# output through a generator p_set series, which PyPSA builds as a fixing
# constraint and Lopf.build (Lopf.scala:71-474) does not model; scigrid-de has
# no PyPSA result in the reference.
PYPSA_LP = ["ac-dc-co2", "ac-dc-dispatch", "ac-dc-meshed", "lodf-mesh", "phase-shift",
            "sclopf-triangle", "standard-types", "storage-cycle", "storage-hvdc", "store-bank",
            "transformer-levels"]
SMALL = ["ac-dc-dispatch", "ac-dc-co2", "lodf-mesh", "phase-shift", "sclopf-triangle",
         "standard-types", "storage-cycle", "store-bank", "transformer-levels", "ac-pf-pv"]
PARAMS = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, max_iterations=500_000)


def net(name):
    return network.read_folder(os.path.join(GOLD, name))


def frame(name, key, t, column):
    f = OPTIMIZE[name][key]
    v = f["values"][t][f["columns"].index(column)]
    return math.nan if isinstance(v, dict) else float(v)


def target(name):
    return OPTIMIZE[name]["objective"] if name in OPTIMIZE and name != "ac-pf-pv" else \
        HIGHS[name]["objective"]


def copy_of(name, tmp_path):
    d = tmp_path / name
    shutil.copytree(os.path.join(GOLD, name), d)
    return d


def set_column(text, column, value):
    """LopfSuite.setColumn: set (or add) one column of a static CSV."""
    lines = text.strip("\n").split("\n")
    head = lines[0].split(",")
    out = []
    if column in head:
        j = head.index(column)
        out.append(lines[0])
        for ln in lines[1:]:
            cells = ln.split(",")
            cells[j] = value(cells[0], cells[j]) if callable(value) else value
            out.append(",".join(cells))
    else:
        out.append(lines[0] + "," + column)
        for ln in lines[1:]:
            out.append(ln + "," + (value(ln.split(",")[0], "") if callable(value) else value))
    return "\n".join(out) + "\n"


def mutate(name, tmp_path, file, edit):
    d = copy_of(name, tmp_path)
    p = d / file
    before = p.read_text()
    after = edit(before)
    assert after != before
    p.write_text(after)
    return network.read_folder(str(d), name=name)


def highs(problem):
    """HiGHS on the LP, rows of tiny coefficients rescaled (its absolute
    feasibility tolerance would otherwise not hold Kirchhoff rows scaled by
    1/v_nom^2)."""
    A = problem.to_scipy().tocsr()
    scale = np.ones(problem.m)
    for r in range(problem.m):
        row = A.data[A.indptr[r]:A.indptr[r + 1]]
        if len(row) and np.abs(row).max() < 1e-3:
            scale[r] = 1.0 / np.abs(row).max()
    A = (A.T.multiply(scale).T).tocsr()
    q = problem.q * scale
    neq = problem.n_eq
    bounds = list(zip(np.where(np.isinf(problem.l), None, problem.l),
                      np.where(np.isinf(problem.u), None, problem.u)))
    res = linprog(problem.c, A_ub=-A[neq:] if problem.m > neq else None,
                  b_ub=-q[neq:] if problem.m > neq else None, A_eq=A[:neq] if neq else None,
                  b_eq=q[:neq] if neq else None, bounds=bounds, method="highs")
    assert res.status == 0, res.message
    return res.fun + problem.offset, res.x


# --------------------------------------------------------------------------
# reader (CsvReader.scala, ComponentTable.scala)


def test_reader_weightings_labels_and_series():
    n = net("storage-cycle")
    assert n.snapshots[:2] == ["0", "1"]
    assert n.weighting("objective", 0) == 2.0 and n.weighting("stores", 3) == 3.0
    assert n.weighting("generators", 1) == 1.0 and n.weighting("missing", 0) == 1.0
    d = net("ac-dc-dispatch")
    g = d.require("Generator")
    s = g.series["p_max_pu"]
    assert s.entities == ["Manchester Wind", "Frankfurt Wind", "Norway Wind"]
    assert g.value_at("p_max_pu", "Manchester Wind", 0) == 0.930019875
    # not covered by the series: static column absent, so the schema default
    assert g.value_at("p_max_pu", "Manchester Gas", 0) == 1.0
    assert d.require("Line").float("s_max_pu", d.require("Line").ids[0]) == 1.0
    assert d.snapshots[0] == "2015-01-01 00:00:00"


def test_reader_keeps_bus_names_and_custom_columns():
    h = net("storage-hvdc")
    assert h.require("Bus").ids[:3] == ["0", "1", "2"]
    assert h.require("Line").string("bus0", "0") == "0"
    assert h.require("StorageUnit").string("bus", "Storage 3") == "3"
    soc = h.require("StorageUnit").series["state_of_charge_set"]
    assert math.isnan(soc.get("Storage 3", 0)) and soc.get("Storage 4", 2) == 25.0
    d = net("ac-dc-dispatch")
    buses = d.require("Bus")
    assert buses.kinds["country"] == "str"  # a custom column, typed from its values
    with pytest.raises(KeyError):
        d.require("Generator").float("no_such_attribute", "Manchester Wind")


def test_reader_rejects_short_series(tmp_path):
    d = copy_of("ac-dc-dispatch", tmp_path)
    p = d / "generators-p_max_pu.csv"
    p.write_text("\n".join(p.read_text().split("\n")[:-3]) + "\n")
    with pytest.raises(network.MalformedNetwork):
        network.read_folder(str(d))


# --------------------------------------------------------------------------
# builder shape


def test_scigrid_de_builds_at_scale():
    # LopfSuite.scala:522-523: 60,552 variables and 23,688 rows
    m = lopf.build(net("scigrid-de"))
    assert (m.problem.n, m.problem.m) == (60_552, 23_688)


@pytest.mark.parametrize("name", PYPSA_LP + ["ac-pf-pv", "scigrid-de"])
def test_cycle_basis_has_the_circuit_rank(name):
    n = lopf.expand_standard_types(net(name))
    edges, touched = 0, set()
    for t in n.tables.values():
        if network.role(t.spec) == "passive":
            for e in t.ids:
                edges += 1
                touched.update((t.string("bus0", e), t.string("bus1", e)))
    islands = sum(1 for _, buses in network.sub_networks(n) if len(buses) > 1)
    cycles = lopf.cycle_basis(n)
    assert len(cycles) == edges - len(touched) + islands
    if name == "ac-dc-dispatch":
        assert len(cycles) == 2  # LopfSuite.scala:209-220


def test_store_power_is_unbounded():
    n = net("store-bank")
    m = lopf.build(n)
    for e in n.require("Store").ids:
        for t in range(len(n.snapshots)):
            c = m.column(lopf.STORE_P, e, t)
            assert m.problem.l[c] == -math.inf and m.problem.u[c] == math.inf


# --------------------------------------------------------------------------
# the model against PyPSA, through HiGHS


@pytest.mark.parametrize("name", PYPSA_LP)
def test_model_reaches_pypsa_objective(name):
    obj, _ = highs(lopf.build(net(name)).problem)
    t = OPTIMIZE[name]["objective"]
    assert abs(obj - t) <= 1e-6 * max(1.0, abs(t)), (obj, t)


def test_model_without_generator_set_points():
    # PyPSA pins g_pv at its p_set series (60, 80) for 7650; the builder, like
    # Lopf.build, leaves it free and the cheap generator covers all 265 MW
    obj, _ = highs(lopf.build(net("ac-pf-pv")).problem)
    assert abs(obj - 2650.0) <= 1e-9 and OPTIMIZE["ac-pf-pv"]["objective"] == 7650.0


def test_co2_cap_is_what_binds(tmp_path):
    # LopfSuite.scala:657-675: without the cap ac-dc-co2 costs 2819.52
    n = net("ac-dc-co2")
    uncapped = mutate("ac-dc-co2", tmp_path, "global_constraints.csv",
                      lambda s: set_column(s, "constant", "1e12"))
    capped, _ = highs(lopf.build(n).problem)
    free, _ = highs(lopf.build(uncapped).problem)
    assert abs(free - 2819.52) <= 0.01 and capped > free + 300
    assert OPTIMIZE["ac-dc-co2"]["global_constraint_mu"]["co2_limit"] != 0.0


def test_state_of_charge_set_point_binds(tmp_path):
    # LopfSuite.scala:557-594, on the LP itself
    n = mutate("storage-cycle", tmp_path, "storage_units.csv",
               lambda s: set_column(s, "state_of_charge_set",
                                    lambda e, _: "90.0" if e == "battery" else ""))
    m = lopf.build(n)
    obj, x = highs(m.problem)
    for t in range(len(n.snapshots)):
        assert abs(x[m.column(lopf.SOC, "battery", t)] - 90.0) <= 1e-6
    assert max(x[m.column(lopf.SOC, "cyclic", t)] for t in range(len(n.snapshots))) > 1.0
    assert obj > OPTIMIZE["storage-cycle"]["objective"] + 1.0


def test_discharge_only_extendable_storage_has_no_nan_bound(tmp_path):
    n = mutate("storage-hvdc", tmp_path, "storage_units.csv",
               lambda s: set_column(s, "p_min_pu", "0.0"))
    m = lopf.build(n)
    for e in n.require("StorageUnit").ids:
        for t in range(len(n.snapshots)):
            assert not math.isnan(m.problem.u[m.column(lopf.STORE, e, t)])


def test_extendable_store_energy_is_free_below(tmp_path):
    n = mutate("store-bank", tmp_path, "stores.csv",
               lambda s: set_column(s, "e_min_pu", lambda e, cur: "-0.5" if e == "grow" else cur))
    m = lopf.build(n)
    for t in range(len(n.snapshots)):
        assert m.problem.l[m.column(lopf.STORE_E, "grow", t)] == -math.inf


# --------------------------------------------------------------------------
# refusals (LopfSuite.scala:222-260, :417-450, :596-611, :746-779, :877-1040)


def _refused(n, phrase):
    with pytest.raises(lopf.UnsupportedNetwork) as e:
        lopf.build(n)
    assert phrase in str(e.value), str(e.value)


def test_dangling_bus_is_refused(tmp_path):
    n = mutate("ac-dc-dispatch", tmp_path, "loads.csv",
               lambda s: s.replace(",Manchester", ",Mancester", 1))
    _refused(n, "Mancester")


def test_unmodelled_component_is_refused(tmp_path):
    d = copy_of("ac-dc-dispatch", tmp_path)
    (d / "shunt_impedances.csv").write_text("name,bus,g,b\nshunt,Manchester,0.1,0.0\n")
    _refused(network.read_folder(str(d)), "ShuntImpedance")


@pytest.mark.parametrize("attribute,value,phrase", [
    ("e_set", "50.0", "e_set"), ("p_set", "10.0", "p_set"),
    ("e_cyclic_per_period", "True", "per_period"),
    ("marginal_cost_quadratic", "0.5", "quadratic"), ("standing_loss", "1.5", "standing_loss")])
def test_store_features_refused(tmp_path, attribute, value, phrase):
    _refused(mutate("store-bank", tmp_path, "stores.csv",
                    lambda s: set_column(s, attribute, value)), phrase)


@pytest.mark.parametrize("attribute", ["p_set", "p_dispatch_set", "p_store_set"])
def test_storage_set_points_refused(tmp_path, attribute):
    _refused(mutate("storage-cycle", tmp_path, "storage_units.csv",
                    lambda s: set_column(s, attribute, "10.0")), attribute)


def test_per_snapshot_standing_loss_refused(tmp_path):
    d = copy_of("store-bank", tmp_path)
    (d / "stores-standing_loss.csv").write_text(
        "name,tank\n" + "\n".join(f"{t},1.5" for t in range(6)) + "\n")
    with pytest.raises(lopf.UnsupportedNetwork) as e:
        lopf.build(network.read_folder(str(d)))
    assert "standing_loss" in str(e.value) and "snapshot" in str(e.value)


def test_static_standing_loss_overridden_by_series_is_accepted(tmp_path):
    d = copy_of("storage-cycle", tmp_path)
    (d / "storage_units-standing_loss.csv").write_text(
        "name,battery\n" + "\n".join(f"{t},0.02" for t in range(6)) + "\n")
    p = d / "storage_units.csv"
    p.write_text(set_column(p.read_text(), "standing_loss",
                            lambda e, cur: "1.5" if e == "battery" else cur))
    lopf.build(network.read_folder(str(d)))
    p.write_text(set_column(p.read_text(), "standing_loss",
                            lambda e, cur: "1.5" if e == "hydro" else cur))
    _refused(network.read_folder(str(d)), "hydro")


def test_capacity_floor_checks(tmp_path):
    n = mutate("ac-dc-meshed", tmp_path, "generators.csv",
               lambda s: set_column(set_column(s, "p_nom_min", "500.0"), "p_nom_max", "100.0"))
    _refused(n, "p_nom_min")
    n = mutate("ac-dc-meshed", tmp_path / "nan", "generators.csv",
               lambda s: set_column(s, "p_nom_min", "nan"))
    _refused(n, "p_nom_min")


def test_annuitised_and_modular_capacity_refused(tmp_path):
    _refused(mutate("ac-dc-meshed", tmp_path, "generators.csv",
                    lambda s: set_column(s, "overnight_cost", "1000.0")), "overnight_cost")
    _refused(mutate("ac-dc-meshed", tmp_path / "mod", "generators.csv",
                    lambda s: set_column(s, "p_nom_mod", "50.0")), "p_nom_mod")


def test_committable_and_extendable_refused(tmp_path):
    _refused(mutate("ac-dc-meshed", tmp_path, "generators.csv",
                    lambda s: set_column(s, "committable", "True")), "committable")


def test_zero_impedance_cycle_refused(tmp_path):
    _refused(mutate("lodf-mesh", tmp_path, "lines.csv",
                    lambda s: set_column(s, "x", "0.0")), "zero impedance")


def test_unknown_line_type_refused(tmp_path):
    n = mutate("standard-types", tmp_path, "lines.csv",
               lambda s: s.replace("Al/St 240/40 4-bundle 380.0", "Al/St 240/41 4-bundle 380.0"))
    with pytest.raises(lopf.UnknownType):
        lopf.build(n)


# --------------------------------------------------------------------------
# the oracle's PDHG on the small networks


@pytest.mark.parametrize("name", SMALL)
def test_oracle_solves_to_pypsa(name):
    from oracle.oracle import OracleLp, default_params
    p = lopf.build(net(name)).problem
    o = OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u, p.offset)
    s = o.solve(default_params(eps_abs=1e-9, eps_rel=1e-9, max_iterations=500_000))
    assert s.status == "Optimal"
    t = target(name)
    assert abs(s.objective - t) <= 1e-6 * max(1.0, abs(t)), (s.objective, t)


# --------------------------------------------------------------------------
# the device


@pytest.fixture(scope="module")
def device_results(hip64):
    names = PYPSA_LP + ["ac-pf-pv"]
    results = lopf.solve_all([net(n) for n in names], hip64, PARAMS)
    return dict(zip(names, results))


@pytest.mark.gpu
@pytest.mark.parametrize("name", PYPSA_LP + ["ac-pf-pv"])
def test_device_objective(device_results, name):
    r = device_results[name]
    assert r.status == "Optimal", r.solution
    t = target(name)
    assert abs(r.objective - t) <= 1e-6 * max(1.0, abs(t)), (r.objective, t)


@pytest.mark.gpu
def test_device_dispatch_and_flows(device_results):
    # LopfSuite.scala:132-207: 1e-3 MW, absolute
    r = device_results["ac-dc-dispatch"]
    n = r.network
    loads = n.require("Load")
    for t in range(len(n.snapshots)):
        for g in n.require("Generator").ids:
            assert abs(r.dispatch("Generator", g, t) - frame("ac-dc-dispatch", "generator_p", t, g)) \
                <= 1e-3
        for ln in n.require("Line").ids:
            flow = r.dispatch("Line", ln, t)
            assert abs(flow - frame("ac-dc-dispatch", "line_p0", t, ln)) <= 1e-3
            assert abs(flow) <= n.require("Line").float("s_nom", ln) + 1e-6
        gen = sum(r.dispatch("Generator", g, t) for g in n.require("Generator").ids)
        dem = sum(loads.value_at("p_set", e, t) for e in loads.ids)
        assert abs(gen - dem) <= 1e-3


@pytest.mark.gpu
def test_device_duals(device_results):
    # LopfSuite.scala:613-655 and :707-744: strong duality, finite prices, and
    # an interior zero-emission generator pins its bus price
    r = device_results["ac-dc-dispatch"]
    s = r.solution
    assert abs(s.objective_value - s.dual_objective_value) <= 1e-6 * max(1.0, abs(r.objective))
    n = r.network
    gens, carriers = n.require("Generator"), n.require("Carrier")
    checked = 0
    for t in range(len(n.snapshots)):
        for bus in n.require("Bus").ids:
            assert math.isfinite(r.marginal_price(bus, t))
        for g in gens.ids:
            p_nom = gens.float("p_nom", g)
            lo, hi = p_nom * gens.value_at("p_min_pu", g, t), p_nom * gens.value_at("p_max_pu", g, t)
            p = r.dispatch("Generator", g, t)
            clean = carriers.float("co2_emissions", gens.string("carrier", g)) == 0.0
            if clean and lo + 1e-2 < p < hi - 1e-2:
                checked += 1
                assert abs(r.marginal_price(gens.string("bus", g), t)
                           - gens.value_at("marginal_cost", g, t)) <= 1e-3
    assert checked >= 2


@pytest.mark.gpu
def test_device_co2_cap_reached(device_results):
    r = device_results["ac-dc-co2"]
    n = r.network
    gens, carriers = n.require("Generator"), n.require("Carrier")
    emitted = sum(r.dispatch("Generator", g, t)
                  * carriers.float("co2_emissions", gens.string("carrier", g))
                  / gens.value_at("efficiency", g, t) * n.weighting("generators", t)
                  for t in range(len(n.snapshots)) for g in gens.ids)
    cap = n.require("GlobalConstraint").float("constant", "co2_limit")
    assert emitted <= cap + 1e-3 and abs(emitted - cap) <= 1e-2


@pytest.mark.gpu
def test_device_storage_frames(device_results):
    # LopfSuite.scala:452-486: every storage frame at 1e-4
    r = device_results["storage-cycle"]
    n = r.network
    reads = {"storage_state_of_charge": r.state_of_charge, "storage_p_dispatch": r.discharging,
             "storage_p_store": r.charging, "storage_spill": r.spill,
             "storage_p": lambda e, t: r.dispatch("StorageUnit", e, t)}
    for key, read in reads.items():
        for t in range(len(n.snapshots)):
            for e in n.require("StorageUnit").ids:
                assert abs(read(e, t) - frame("storage-cycle", key, t, e)) <= 1e-4, (key, e, t)


@pytest.mark.gpu
def test_device_stores(device_results):
    # LopfSuite.scala:324-364
    r = device_results["store-bank"]
    n = r.network
    for t in range(len(n.snapshots)):
        for e in n.require("Store").ids:
            assert abs(r.energy(e, t) - frame("store-bank", "store_e", t, e)) <= 1e-4
            assert abs(r.dispatch("Store", e, t) - frame("store-bank", "store_p", t, e)) <= 1e-4
    for e in n.require("Store").ids:
        assert abs(r.capacity("Store", e) - OPTIMIZE["store-bank"]["nominal_opt"]["Store"][e]) <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ac-dc-meshed", "storage-hvdc"])
def test_device_expansion(device_results, name):
    # LopfSuite.scala:261-321: objective, constant, total cost, capacities
    r = device_results[name]
    opt = OPTIMIZE[name]
    const = opt["objective_constant"]
    assert abs((r.total_system_cost - r.objective) - const) <= 1e-6 * abs(const)
    assert abs(r.total_system_cost - opt["total_system_cost"]) <= 1e-6 * abs(opt["total_system_cost"])
    compared, moved = 0, False
    for component, chosen in opt["nominal_opt"].items():
        table = r.network.table(component)
        if table is None:
            continue
        for e, v in chosen.items():
            assert abs(r.capacity(component, e) - v) <= 1e-4 * max(1.0, abs(v)), (component, e)
            compared += 1
            moved |= abs(v - table.float(lopf.NOMINAL[component], e)) > 1.0
    assert compared > 10 and moved


@pytest.mark.gpu
@pytest.mark.parametrize("name,component", [("transformer-levels", "Transformer"),
                                            ("phase-shift", "Transformer"),
                                            ("standard-types", "Line"),
                                            ("standard-types", "Transformer")])
def test_device_branch_flows(device_results, name, component):
    r = device_results[name]
    key = "line_p0" if component == "Line" else "transformer_p0"
    for t in range(len(r.network.snapshots)):
        for e in r.network.require(component).ids:
            assert abs(r.dispatch(component, e, t) - frame(name, key, t, e)) <= 1e-3


@pytest.mark.gpu
def test_device_scigrid_de(hip64):
    # LopfSuite.scala:519-555: 1e-4 tolerances, objective within 2e-4 of the
    # optimum (HiGHS on the same LP; the reference holds no PyPSA result)
    r = lopf.solve(net("scigrid-de"), hip64,
                   prima.PdhgParams(eps_abs=1e-4, eps_rel=1e-4, max_iterations=100_000))
    assert r.status == "Optimal", r.solution
    t = HIGHS["scigrid-de"]["objective"]
    assert abs(r.objective - t) <= 2e-4 * abs(t), (r.objective, t)
