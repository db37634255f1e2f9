"""Write golden objectives for the seeded dispatch instances with HiGHS
(scipy.optimize.linprog, method="highs"), an LP oracle independent of the
reference and of this backend. Run in a container that has scipy; the JSON
it writes is the fixture the tests and the benchmark read.

  python tests/golden/make_golden.py          (dispatch and sweep goldens)
  python tests/golden/make_golden.py sweep    (only the missing sweep goldens)
"""
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def highs(spec):
    from scipy.optimize import linprog
    from noaidi_amd import prima
    kind, kw = spec
    if kind == "dispatch":
        p = prima.dispatch(**kw)
    elif kind == "sweep":
        p = prima.config3(kw["k"])
    else:
        p = prima.random_known(**kw)
    A = p.to_scipy()
    ne = p.n_eq
    t = time.time()
    method = "highs"
    # large sweep instances: the interior-point solver first (the simplex
    # takes more than ten minutes beyond ~0.5M variables)
    order = ("highs-ipm", "highs") if p.n > 100_000 else ("highs", "highs-ipm")
    for method in order:
        # a few instances end in HiGHS's "numerical difficulties" (status 4)
        # with the simplex at these tolerances; the interior-point solver then
        # settles them
        r = linprog(p.c, A_eq=A[:ne] if ne else None, b_eq=p.q[:ne] if ne else None,
                    A_ub=-A[ne:] if p.m > ne else None, b_ub=-p.q[ne:] if p.m > ne else None,
                    bounds=list(zip(p.l, p.u)), method=method,
                    options=dict(primal_feasibility_tolerance=1e-10,
                                 dual_feasibility_tolerance=1e-10))
        if r.status != 4:
            break
    return dict(kind=kind, args=kw, n=p.n, m=p.m, n_eq=ne, nnz=p.nnz, status=int(r.status),
                objective=float(r.fun) if r.status == 0 else None, method=method,
                seconds=time.time() - t)


def sweep_pick():
    """Sweep instances with goldens: the 16 smallest, the 4 around the median
    size, and the midpoints of 16 equal-count size strata (every snapshot
    count, up to 2.4M variables)."""
    from noaidi_amd import prima
    args = [prima.config3_args(k) for k in range(256)]
    size = sorted(range(256), key=lambda k: (args[k]["buses"] * args[k]["snapshots"], k))
    strata = [size[int(i * 255 / 15)] for i in range(16)]
    pick = size[:16] + size[126:130]
    return pick + [k for k in strata if k not in pick]


def write_sweep():
    from noaidi_amd import prima
    pick = sweep_pick()
    path = os.path.join(HERE, "highs_sweep.json")
    done = {}
    if os.path.exists(path):
        with open(path) as f:
            done = {d["args"]["k"]: d for d in json.load(f)}
    todo = [k for k in pick if k not in done]
    with ProcessPoolExecutor(4) as ex:
        for d in ex.map(highs, [("sweep", dict(k=k)) for k in todo]):
            d["config3_args"] = prima.config3_args(d["args"]["k"])
            done[d["args"]["k"]] = d
            print(d["args"]["k"], d["n"], d["status"], d["method"], round(d["seconds"]), flush=True)
    with open(path, "w") as f:
        json.dump([done[k] for k in pick if k in done], f, indent=1)
    print(len(done), "sweep instances written")


def main():
    if sys.argv[1:] == ["sweep"]:
        write_sweep()
        return
    specs = [("dispatch", dict(seed=1))]  # the 200-bus, 24-snapshot dispatch LP
    specs += [("dispatch", dict(seed=1, perturb=0.2, scenario_seed=1000 + k)) for k in range(256)]
    specs += [("dispatch", dict(seed=s, buses=20, chords=10, snapshots=6, storage=5))
              for s in range(8)]
    specs += [("dispatch", dict(seed=7, buses=60, chords=30, snapshots=24, storage=15))]
    with ProcessPoolExecutor(8) as ex:
        out = list(ex.map(highs, specs))
    with open(os.path.join(HERE, "highs_dispatch.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(out), "instances written")
    write_sweep()


if __name__ == "__main__":
    main()
