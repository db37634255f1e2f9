"""Unit commitment on PyPSA networks against UnitCommitmentSuite
(network-lopf UnitCommitmentSuite.scala:43-345) and PyPSA's golden
(reference/goldens/results/unit-commitment.json, objective 17000).

The MILP the builder makes is pinned by HiGHS's MILP solve reaching PyPSA's
objective; the oracle's branch-and-bound and the device branch-and-bound then
solve it to the same objective, schedule and dispatch."""
import json
import math
import os
import shutil

import numpy as np
import pytest
from scipy.optimize import Bounds, LinearConstraint, milp

from noaidi_amd import network, prima
from noaidi_amd import unit_commitment as uc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "pypsa")
with open(os.path.join(GOLD, "optimize.json")) as _f:
    GOLDEN = json.load(_f)["unit-commitment"]
# UnitCommitmentSuite.scala:30-33
LP = prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9, max_iterations=200_000)
BNB = prima.BnbParams(lp=LP, max_nodes=20_000)


def fixture():
    return network.read_folder(os.path.join(GOLD, "unit-commitment"))


def residual_case(tmp_path):
    """UnitCommitmentSuite.scala:229-277: an expensive unit held on for the
    two snapshots its min_up_time still owes, beside ample cheap capacity."""
    d = tmp_path / "uc-residual"
    d.mkdir()
    (d / "buses.csv").write_text("name,v_nom,carrier\nbus,110.0,AC\n")
    (d / "generators.csv").write_text(
        "name,bus,p_nom,p_min_pu,marginal_cost,committable,min_up_time,min_down_time\n"
        "unit,bus,100.0,0.5,500.0,True,3,1\ncheap,bus,500.0,0.0,10.0,False,0,0\n")
    (d / "loads.csv").write_text("name,bus,p_set\nd,bus,200.0\n")
    (d / "snapshots.csv").write_text(",snapshot\n0,0\n1,1\n2,2\n3,3\n")
    return network.read_folder(str(d))


def frame(key, t, column):
    f = GOLDEN[key]
    return float(f["values"][t][f["columns"].index(column)])


def highs_milp(m):
    p = m.problem
    A = p.to_scipy()
    integ = np.zeros(p.n)
    integ[m.integers] = 1
    cons = [LinearConstraint(A[:p.n_eq], p.q[:p.n_eq], p.q[:p.n_eq])]
    if p.m > p.n_eq:
        cons.append(LinearConstraint(A[p.n_eq:], p.q[p.n_eq:], np.inf))
    r = milp(p.c, constraints=cons, integrality=integ, bounds=Bounds(p.l, p.u))
    assert r.status == 0, r.message
    return r.fun + p.offset


def oracle_solve(m):
    from oracle.oracle import OracleLp, bnb_solve, default_params
    p = m.problem
    o = OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u, p.offset)
    s = bnb_solve(o, m.integers, default_params(eps_abs=1e-9, eps_rel=1e-9, max_iterations=200_000),
                  max_nodes=20_000)
    return uc.UcResult(m, prima.MilpSolution(s.status, s.primal, s.objective, s.best_bound,
                                             s.nodes_explored, s.unproven_nodes, 0.0))


def check_against_golden(r, n):
    # UnitCommitmentSuite.scala:43-130
    assert r.status == "Optimal"
    assert abs(r.objective - GOLDEN["objective"]) <= 1e-6 * 17000.0
    gens, loads = n.require("Generator"), n.require("Load")
    off = 0
    for t in range(len(n.snapshots)):
        for g in gens.ids:
            if "committable" in gens.static and gens.bool("committable", g):
                theirs = frame("generator_status", t, g) > 0.5
                assert r.committed(g, t) == theirs, (t, g)
                off += not theirs
            assert abs(r.output(g, t) - frame("generator_p", t, g)) <= 1e-4, (t, g)
            p = r.output(g, t)
            if r.committed(g, t):
                floor = gens.float("p_nom", g) * gens.value_at("p_min_pu", g, t)
                assert p >= floor - 1e-4
            else:
                assert abs(p) <= 1e-4
        supplied = sum(r.output(g, t) for g in gens.ids)
        demanded = sum(loads.value_at("p_set", e, t) for e in loads.ids)
        assert abs(supplied - demanded) <= 1e-4
    assert off >= 3
    # minimum down time (UnitCommitmentSuite.scala:132-157)
    for g in gens.ids:
        md = gens.int("min_down_time", g) if "min_down_time" in gens.static else 0
        if md > 1:
            start = -1
            for t in range(len(n.snapshots)):
                on = r.committed(g, t)
                if not on and start < 0:
                    start = t
                if on and start >= 0:
                    assert t - start >= md
                    start = -1


def check_residual(r, n):
    gens = n.require("Generator")
    forced = max(gens.int("min_up_time", "unit") - gens.int("up_time_before", "unit"), 0)
    assert forced == 2 and r.status == "Optimal"
    for t in range(forced):
        assert r.committed("unit", t) and abs(r.output("unit", t) - 50.0) <= 1e-4
    for t in range(forced, len(n.snapshots)):
        assert not r.committed("unit", t)
    assert abs(r.objective - 57000.0) <= 1e-3


def test_model_reaches_pypsa_objective(tmp_path):
    m = uc.build(fixture())
    assert (m.problem.n, m.problem.m, len(m.integers)) == (72, 82, 16)
    assert abs(highs_milp(m) - GOLDEN["objective"]) <= 1e-9
    assert abs(highs_milp(uc.build(residual_case(tmp_path))) - 57000.0) <= 1e-6


def test_oracle_branch_and_bound():
    n = fixture()
    check_against_golden(oracle_solve(uc.build(n)), n)


def test_oracle_holds_the_residual(tmp_path):
    n = residual_case(tmp_path)
    check_residual(oracle_solve(uc.build(n)), n)


def copy_fixture(tmp_path):
    d = tmp_path / "uc"
    shutil.copytree(os.path.join(GOLD, "unit-commitment"), d)
    return d


def refused(n, phrase, absent=None):
    with pytest.raises(uc.UnsupportedNetwork) as e:
        uc.build(n)
    assert phrase in str(e.value), str(e.value)
    if absent:
        assert absent not in str(e.value)


def test_ramp_limits_refused(tmp_path):
    d = copy_fixture(tmp_path)
    p = d / "generators.csv"
    lines = p.read_text().strip().split("\n")
    p.write_text("\n".join([lines[0] + ",ramp_limit_up"] + [ln + ",0.5" for ln in lines[1:]]) + "\n")
    refused(network.read_folder(str(d)), "ramp")


def test_transmission_refused(tmp_path):
    d = copy_fixture(tmp_path)
    (d / "buses.csv").write_text((d / "buses.csv").read_text().rstrip() + "\nfar,110.0\n")
    (d / "lines.csv").write_text("name,bus0,bus1,x,r,s_nom\nl,bus,far,1.0,0.0,500.0\n")
    refused(network.read_folder(str(d)), "Line")


def test_dangling_load_refused(tmp_path):
    d = copy_fixture(tmp_path)
    p = d / "loads.csv"
    rows = p.read_text().strip().split("\n")
    p.write_text("\n".join([rows[0]] + [r.replace(",bus", ",buss") for r in rows[1:]]) + "\n")
    refused(network.read_folder(str(d)), "buss")


def test_global_constraint_refused_with_its_own_reason(tmp_path):
    d = copy_fixture(tmp_path)
    (d / "global_constraints.csv").write_text(
        "name,type,carrier_attribute,sense,constant\nco2_limit,primary_energy,co2_emissions,<=,"
        "1000.0\n")
    refused(network.read_folder(str(d)), "global constraint", absent="branch flows")


def test_series_only_stand_by_cost_refused(tmp_path):
    d = copy_fixture(tmp_path)
    (d / "generators-stand_by_cost.csv").write_text(
        ",base\n0,0.0\n1,0.0\n2,25.0\n3,0.0\n4,0.0\n5,0.0\n6,0.0\n7,0.0\n")
    refused(network.read_folder(str(d)), "stand_by_cost")


def test_priced_attribute_refused_without_snapshots(tmp_path):
    d = tmp_path / "uc-empty"
    d.mkdir()
    (d / "buses.csv").write_text("name,v_nom,carrier\nbus,110.0,AC\n")
    (d / "generators.csv").write_text(
        "name,bus,p_nom,p_min_pu,marginal_cost,committable,stand_by_cost\n"
        "unit,bus,100.0,0.5,10.0,True,25.0\n")
    (d / "loads.csv").write_text("name,bus,p_set\nd,bus,50.0\n")
    (d / "snapshots.csv").write_text(",snapshot\n")
    n = network.read_folder(str(d))
    assert len(n.snapshots) == 0
    refused(n, "stand_by_cost")


@pytest.mark.gpu
@pytest.mark.parametrize("frontier", [1, 8])
def test_device_branch_and_bound(hip64, frontier):
    n = fixture()
    params = prima.BnbParams(lp=LP, max_nodes=20_000, frontier=frontier)
    check_against_golden(uc.solve(n, hip64, params), n)


@pytest.mark.gpu
def test_device_holds_the_residual(hip64, tmp_path):
    n = residual_case(tmp_path)
    check_residual(uc.solve(n, hip64, BNB), n)
