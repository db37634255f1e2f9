"""Write the PyPSA LOPF fixtures under tests/golden/pypsa/ (run in the build
container, where the reference tree is present; the output travels, the
reference does not).

- `<network>/*.csv`: PyPSA 1.2.4's own CSV exports of the reference's golden
  networks (reference/goldens/networks/), copied as data.
- `optimize.json`: the `optimize` section of each reference/goldens/results/
  <network>.json (PyPSA + HiGHS objectives, chosen capacities, dispatch frames).
- `highs.json`: the objective scipy's HiGHS reaches on the LP that
  `noaidi_amd.lopf.build` makes from each network, with rows of tiny
  coefficients rescaled first (HiGHS's absolute feasibility tolerance would
  otherwise leave Kirchhoff rows scaled by 1/v_nom^2 unenforced). scigrid-de
  has no PyPSA result in the reference, so this is its only objective.
"""
import json
import os
import shutil
import sys

import numpy as np
from scipy.optimize import linprog

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from noaidi_amd import lopf, network  # noqa: E402

SRC = "/root/reference/reference/goldens"
OUT = os.path.join(HERE, "pypsa")
NETWORKS = ["ac-dc-co2", "ac-dc-dispatch", "ac-dc-meshed", "ac-pf-pv", "lodf-mesh", "phase-shift",
            "scigrid-de", "sclopf-triangle", "standard-types", "storage-cycle", "storage-hvdc",
            "store-bank", "transformer-levels", "unit-commitment"]


def highs_objective(p):
    A = p.to_scipy().tocsr().astype(np.float64)
    q = p.q.copy()
    scale = np.ones(p.m)
    for r in range(p.m):
        row = A.data[A.indptr[r]:A.indptr[r + 1]]
        if len(row):
            mx = np.abs(row).max()
            if mx < 1e-3:
                scale[r] = 1.0 / mx
    A = (A.T.multiply(scale).T).tocsr()
    q = q * scale
    neq = p.n_eq
    bounds = list(zip(np.where(np.isinf(p.l), None, p.l), np.where(np.isinf(p.u), None, p.u)))
    res = linprog(p.c, A_ub=-A[neq:] if p.m > neq else None, b_ub=-q[neq:] if p.m > neq else None,
                  A_eq=A[:neq] if neq else None, b_eq=q[:neq] if neq else None, bounds=bounds,
                  method="highs")
    assert res.status == 0, res.message
    return float(res.fun + p.offset)


def main():
    os.makedirs(OUT, exist_ok=True)
    optimize, highs = {}, {}
    for name in NETWORKS:
        src = os.path.join(SRC, "networks", name)
        dst = os.path.join(OUT, name)
        if os.path.isdir(dst):
            shutil.rmtree(dst)
        os.makedirs(dst)
        for f in sorted(os.listdir(src)):
            if f.endswith(".csv"):
                shutil.copy(os.path.join(src, f), os.path.join(dst, f))
        res = os.path.join(SRC, "results", f"{name}.json")
        if os.path.exists(res):
            with open(res) as f:
                optimize[name] = json.load(f)["optimize"]
        if name == "unit-commitment":
            continue  # a MILP: committable generators are not an LP
        net = network.read_folder(dst)
        m = lopf.build(net)
        highs[name] = {"objective": highs_objective(m.problem), "n": m.problem.n,
                       "m": m.problem.m, "n_eq": m.problem.n_eq, "nnz": int(len(m.problem.val))}
        print(name, highs[name], optimize.get(name, {}).get("objective"), flush=True)
    with open(os.path.join(OUT, "optimize.json"), "w") as f:
        json.dump(optimize, f, indent=0, sort_keys=True)
    with open(os.path.join(OUT, "highs.json"), "w") as f:
        json.dump(highs, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
