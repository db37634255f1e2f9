"""Benchmark: LPs/sec to 1e-8 relative KKT on one MI355X.

Workload: the 200-bus, 24-snapshot PyPSA-shaped dispatch LP (34,800
variables, 14,400 rows, 64,750 nonzeros), as a batch of demand scenarios
(per-bus demand x U[0.8, 1.2], scenario seeds 1000..1000+B-1). One step
solves the whole batch with mixed precision (float32 device pass, then a
float64 device pass warm-started from it) inside shared persistent launches.
Plain FP64 PDHG solving the first --unbatched scenarios one at a time is timed
beside it: the metric's denominator. LPs that finish early hand their
workgroups to the ones still running (elastic batching).

Every LP must come back Optimal with an objective within 1e-8 (relative,
|a-b|/max(1,|b|)) of the HiGHS golden in tests/golden/highs_dispatch.json;
otherwise the run fails.

  python bench.py [--steps K] [--warmup W] [--batch B] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from noaidi_amd import prima  # noqa: E402

PEAK_HBM_GBS = 8000.0
GOLDEN = os.path.join(ROOT, "tests", "golden", "highs_dispatch.json")


def scenario(k: int) -> prima.LpProblem:
    return prima.dispatch(seed=1, perturb=0.2, scenario_seed=1000 + k)


def golden_objectives() -> dict:
    with open(GOLDEN) as f:
        data = json.load(f)
    out = {}
    for d in data:
        a = d["args"]
        if a.get("seed") == 1 and a.get("perturb") == 0.2 and "buses" not in a:
            out[a["scenario_seed"] - 1000] = d["objective"]
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_gap(a, b):
    return abs(a - b) / max(1.0, abs(b))


def check(sols, idx, gold, what):
    worst = 0.0
    for k, s in zip(idx, sols):
        if s.status != "Optimal":
            raise SystemExit(f"{what}: scenario {k} ended {s.status}")
        if k in gold:
            g = rel_gap(s.objective_value, gold[k])
            worst = max(worst, g)
            if g > 1e-8:
                raise SystemExit(f"{what}: scenario {k} objective {s.objective_value!r} vs "
                                 f"HiGHS {gold[k]!r} (gap {g:.2e})")
    return worst


# Iterations the oracle needs to bring scenario 0 to 1e-8 (deterministic:
# same restatement, same libm). The baseline times a bounded prefix of that
# run and scales it to one full LP.
ORACLE_ITERATIONS_SCENARIO0 = 776_640
CPU_SAMPLE_ITERATIONS = 100_000


def cpu_baseline(problem: prima.LpProblem):
    """The oracle (C restatement of Prima's FP64 CPU path, single thread) on
    scenario 0: time per iteration over a bounded prefix, scaled to LPs/s."""
    from oracle.oracle import OracleLp, default_params
    o = OracleLp.from_csr(problem.n, problem.m, problem.n_eq, problem.row_ptr, problem.col,
                          problem.val, problem.c, problem.q, problem.l, problem.u,
                          problem.offset)
    t = time.perf_counter()
    s = o.solve(default_params(max_iterations=CPU_SAMPLE_ITERATIONS))
    dt = time.perf_counter() - t
    per_lp = dt / s.iterations * ORACLE_ITERATIONS_SCENARIO0
    return dict(value=1.0 / per_lp, unit="LPs/s", cores=1, kind="port",
                sample=f"scenario 0 on the C restatement of Prima's ScalaKernels+Pdhg path: first "
                       f"{s.iterations} of the {ORACLE_ITERATIONS_SCENARIO0} iterations it needs "
                       f"to reach 1e-8 ({dt:.1f} s), scaled to one LP ({per_lp:.0f} s)",
                cpu=platform.processor() or platform.machine())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--wg", type=int, default=0, help="workgroups per LP (0: tuner)")
    ap.add_argument("--unbatched", type=int, default=4,
                    help="scenarios solved one at a time for the FP64 denominator")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build solves independent LPs per GPU "
              "and has no multi-GPU path to measure", file=sys.stderr)
        sys.exit(2)

    B = args.batch
    batch = [scenario(k) for k in range(B)]
    gold = golden_objectives()
    params = prima.PdhgParams(max_iterations=1_000_000, workgroups_per_lp=args.wg)
    mixed = prima.MixedParams()
    k32 = prima.HipKernels(0, 32)
    k64 = prima.HipKernels(0, 64)

    for _ in range(args.warmup):
        t = time.perf_counter()
        check(k32.solve_all(batch, params, mixed=mixed), range(B), gold, "warmup")
        log(f"warmup step: {time.perf_counter() - t:.2f} s")

    k32.reset_stats()
    t0 = time.perf_counter()
    worst = 0.0
    iters = []
    for _ in range(args.steps):
        sols = k32.solve_all(batch, params, mixed=mixed)
        worst = max(worst, check(sols, range(B), gold, "mixed batched"))
        iters += [(s.mixed["reduced_iterations"], s.mixed["refinement_iterations"]) for s in sols]
        log(f"step done: {time.perf_counter() - t0:.2f} s cumulative, worst gap {worst:.2e}")
    elapsed = time.perf_counter() - t0
    st = k32.stats()
    value = B * args.steps / elapsed

    # plain FP64 PDHG, one LP at a time (the denominator), on the first
    # --unbatched scenarios of the same batch
    U = min(args.unbatched, B)
    k64.reset_stats()
    t = time.perf_counter()
    s64 = [k64.solve(batch[k], params) for k in range(U)]
    t64 = time.perf_counter() - t
    check(s64, range(U), gold, "fp64 unbatched")
    fp64_lps = U / t64
    st64 = k64.stats()
    log(f"mixed batched {value:.3f} LPs/s; fp64 unbatched {fp64_lps:.3f} LPs/s "
        f"({[s.iterations for s in s64]} iterations, {t64:.2f} s); kernel stats {st}")

    achieved = st["algorithmic_bytes"] / (st["kernel_ms"] * 1e-3) / 1e9 if st["kernel_ms"] else 0
    launches = max(1, st["launches"])
    line = {
        "metric": "LPs/sec at 1e-8 rel. KKT: mixed-prec batched vs plain FP64 PDHG, 1 MI355X",
        "value": value,
        "unit": "LPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 iterations + f64 refinement",
        "data": "synthetic",
        "config": {
            "workload": "dispatch LP, 200 buses x 24 snapshots (34,800 vars, 14,400 rows, "
                        "64,750 nnz), batch of demand scenarios",
            "batch": B,
            "n": batch[0].n, "m": batch[0].m, "nnz": batch[0].nnz,
            "tolerance": 1e-8,
        },
        "fp64_unbatched_lps": fp64_lps,
        "fp64_unbatched_sample": U,
        "speedup_vs_fp64_unbatched": value / fp64_lps,
        "fp64_unbatched_iterations": [s.iterations for s in s64],
        "mixed_iterations_mean": [float(np.mean([a for a, _ in iters])),
                                  float(np.mean([b for _, b in iters]))],
        "worst_rel_gap_vs_highs": worst,
        "roofline": {
            "bound": "hbm",
            "kernel": "pdhg_persistent",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": st["algorithmic_bytes"] / launches,
            "avg_launch_ms": st["kernel_ms"] / launches,
            "launches": st["launches"],
        },
        "fp64_roofline": {
            "achieved": (st64["algorithmic_bytes"] / (st64["kernel_ms"] * 1e-3) / 1e9
                         if st64["kernel_ms"] else 0),
            "unit": "GB/s",
        },
    }
    if not args.no_cpu:
        log("timing the CPU baseline (oracle, one LP, one core)")
        line["cpu_baseline"] = cpu_baseline(batch[0])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
