"""Write HiGHS goldens for the unit-commitment branch-and-bound frontier:
the root relaxation of the 100-bus, 24-snapshot UC, the nine most
fractional commitment binaries of its HiGHS optimum (at most one per
generator, so the frontier is not dominated by minimum up/down time
conflicts of one unit), and all 512 node LPs
that fix those nine binaries to every 0/1 pattern (node k fixes column
cols[i] to bit i of k), each with its HiGHS status and objective.

It also writes the HiGHS MILP optimum (scipy.optimize.milp, zero gap) of
small unit-commitment instances that branch-and-bound closes in a few
hundred nodes.

  python tests/golden/make_golden_uc.py
"""
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

ARGS = dict(seed=1, buses=100, chords=50, snapshots=24, gens_per_bus=3, min_up=3, min_down=3)
DEPTH = 9


def problem():
    from noaidi_amd import prima
    return prima.unit_commitment(**ARGS)


def highs(p, l, u):
    from scipy.optimize import linprog
    A = p.to_scipy()
    ne = p.n_eq
    r = linprog(p.c, A_eq=A[:ne], b_eq=p.q[:ne], A_ub=-A[ne:], b_ub=-p.q[ne:],
                bounds=list(zip(l, u)), method="highs",
                options=dict(primal_feasibility_tolerance=1e-10, dual_feasibility_tolerance=1e-10))
    return r


def node(args):
    k, cols = args
    p = problem()
    l, u = p.l.copy(), p.u.copy()
    for i, c in enumerate(cols):
        l[c] = u[c] = float((k >> i) & 1)
    t = time.time()
    r = highs(p, l, u)
    return dict(node=k, status=int(r.status), objective=float(r.fun) if r.status == 0 else None,
                seconds=time.time() - t)


SMALL = [dict(seed=2, buses=6, chords=2, snapshots=6, gens_per_bus=1, min_up=3, min_down=3),
         dict(seed=5, buses=5, chords=2, snapshots=6, gens_per_bus=1, min_up=2, min_down=2),
         dict(seed=6, buses=4, chords=1, snapshots=8, gens_per_bus=1, min_up=3, min_down=3)]


def small_milp(args):
    from scipy.optimize import Bounds, LinearConstraint, milp
    from noaidi_amd import prima
    p = prima.unit_commitment(**args)
    integ = np.zeros(p.n)
    integ[p.integers] = 1
    lo, hi = p.q.copy(), p.q.copy()
    hi[p.n_eq:] = np.inf
    r = milp(p.c, constraints=LinearConstraint(p.to_scipy(), lo, hi), integrality=integ,
             bounds=Bounds(p.l, p.u), options=dict(mip_rel_gap=0))
    assert r.status == 0
    rel = highs(p, p.l, p.u)
    return dict(args=args, n=p.n, m=p.m, n_int=len(p.integers), milp_objective=float(r.fun),
                relaxation_objective=float(rel.fun))


def main():
    with open(os.path.join(HERE, "highs_uc_milp.json"), "w") as f:
        json.dump([small_milp(a) for a in SMALL], f, indent=1)
    p = problem()
    r = highs(p, p.l, p.u)
    assert r.status == 0
    ints = np.asarray(p.integers)
    xu = r.x[ints]
    frac = np.abs(xu - np.round(xu))
    per_t = p.n // ARGS["snapshots"]
    gens = ARGS["buses"] * ARGS["gens_per_bus"]
    cols, units = [], set()
    for i in sorted(range(len(ints)), key=lambda i: (-frac[i], ints[i])):
        unit = (int(ints[i]) % per_t) - gens  # u columns follow the G output columns
        if unit not in units:
            units.add(unit)
            cols.append(int(ints[i]))
        if len(cols) == DEPTH:
            break
    with ProcessPoolExecutor(8) as ex:
        nodes = list(ex.map(node, [(k, cols) for k in range(1 << DEPTH)]))
    out = dict(args=ARGS, n=p.n, m=p.m, n_eq=p.n_eq, nnz=p.nnz, root_objective=float(r.fun),
               fractional=int((frac > 1e-6).sum()), branch_columns=cols,
               branch_root_values=[float(r.x[c]) for c in cols], nodes=nodes)
    with open(os.path.join(HERE, "highs_uc.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(len(nodes), "nodes written;", sum(n["status"] == 0 for n in nodes), "optimal")


if __name__ == "__main__":
    main()
