"""Benchmark: LPs/sec to 1e-8 relative KKT on one MI355X.

Headline workload (the default, --workload scenarios): 256 demand scenarios
of the 200-bus, 24-snapshot PyPSA-shaped dispatch LP (the config-1 LP: 34,800
variables, 14,400 rows, 64,750 nonzeros; per-bus demand x U[0.8, 1.2] with
seed 1000 + k). The metric is a batched one; this batch finishes inside a
600 s run, while the uneven 256-LP sweep takes more than 19 minutes to 1e-8. One step solves all 256 in one batched
call with mixed precision: a float32 device pass, then a float64 device pass
warm-started from it (the reference's MixedPrecision contract), every LP in the same
persistent launches with its own step size, restarts and termination. Every
LP must come back Optimal and within 1e-8 (|a-b|/max(1,|b|)) of its HiGHS
objective (tests/golden/highs_dispatch.json), or the run fails.

--workload sweep runs the same measurement on the uneven scenario sweep
(prima.config3, k = 0..--batch-1: buses log-uniform in [50, 2000], snapshots
in {24, 72, 168}; 88.1M variables over 256 LPs). Its largest LPs (up to 2.4M
variables) keep the batch running for more than 19 minutes.

Timing follows a fixed harness command (--steps K --warmup W): warm-up is
capped at one small solve per context (the persistent kernel has no JIT to
warm), and steps are timed until the wall budget (--budget seconds) would be
exceeded, at least one; `steps` reports the steps actually timed, and a
repeated step must reproduce the first bit for bit. Then, in the same run:
the whole batch in plain FP64 batched; plain FP64 one LP at a time (the
metric's denominator) on a sample stratified by the FP64 iteration counts
(by size for the sweep); and the oracle (the C restatement of Prima's FP64
CPU path) on the host cores. One JSON line, printed and flushed at the end.

  python bench.py [--steps K] [--warmup W] [--budget S] [--no-cpu]
                  [--workload scenarios|sweep|sweep-fp64|uc-frontier|kernel|lopf|spmv|
                              random-lp|tune]

--workload uc-frontier: the 100-bus UC relaxation's depth-9 branch-and-bound
frontier (--batch nodes, shared K, warm-started from the root). --workload
kernel: the solver kernel alone for profiling (dispatch LPs for --iterations
iterations each), after a 2^26-double copy that calibrates hardware byte
counters. --workload lopf: PyPSA's scigrid-de. --workload spmv: every SpMV
schedule on the 2M-column random LP and on the sweep's 256 matrices side by
side. --workload random-lp: the 2M-variable random LP in explicit float32 and
float64 passes under time limits (--pass-seconds). --workload tune: the
solver autotuner's timings on the 200-bus dispatch LP, the sweep's largest
LP and the 2M-variable random LP (--iterations per candidate).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from noaidi_amd import prima  # noqa: E402

METRIC = "LPs/sec at 1e-8 rel. KKT: mixed-prec batched vs plain FP64 PDHG, 1 MI355X"
PEAK_HBM_GBS = 8000.0
GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")
# HBM bytes per algorithmic byte of pdhg_persistent<double> on the headline
# step, from rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by the counters'
# reading of a 512 MiB device copy (profiles/pdhg_persistent_scenarios256_pmc.md)
TRAFFIC_PER_ALGORITHMIC_BYTE = {32: None, 64: 1.143}
# host cores of the GPU box's share for one GPU
CPU_CORES = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_gap(a, b):
    return abs(a - b) / max(1.0, abs(b))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def generate(fn, ks):
    """Build instances on host threads (the generators are C++ and release
    the GIL)."""
    with ThreadPoolExecutor(min(CPU_CORES, os.cpu_count() or 1)) as ex:
        return list(ex.map(fn, ks))


def run_watched(k, fn, what, every=20.0):
    """Run fn() in a worker thread; print the library's progress counters
    while it runs."""
    out = {}

    def work():
        try:
            out["v"] = fn()
        except BaseException as e:  # re-raised in the caller
            out["e"] = e

    th = threading.Thread(target=work)
    th.start()
    t = time.perf_counter()
    while th.is_alive():
        th.join(every)
        if th.is_alive():
            pr = k.progress()
            log(f"{what}: {time.perf_counter() - t:.0f} s, {pr['launches']} launches, "
                f"{pr['accepted_iterations']} iterations, {pr['finished_lps']} passes finished")
    if "e" in out:
        raise out["e"]
    return out["v"]


def check(sols, labels, gold, what):
    worst = 0.0
    for k, s in zip(labels, sols):
        if s.status != "Optimal":
            raise SystemExit(f"{what}: LP {k} ended {s.status}")
        if k in gold:
            g = rel_gap(s.objective_value, gold[k])
            worst = max(worst, g)
            if g > 1e-8:
                raise SystemExit(f"{what}: LP {k} objective {s.objective_value!r} vs HiGHS "
                                 f"{gold[k]!r} (gap {g:.2e})")
    return worst


def dominant(stats):
    """The kernel instantiation (float32 or float64 iterations) that took the
    most device time, with its launches, time and algorithmic bytes."""
    ms = stats["kernel_ms_by_precision"]
    i = 1 if ms[1] >= ms[0] else 0
    return dict(precision=64 if i else 32, launches=stats["launches_by_precision"][i],
                kernel_ms=ms[i], algorithmic_bytes=stats["algorithmic_bytes_by_precision"][i])


def roofline(stats, copy=None):
    d = dominant(stats)
    launches = max(1, d["launches"])
    per_launch = d["algorithmic_bytes"] / launches
    avg_ms = d["kernel_ms"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    ratio = TRAFFIC_PER_ALGORITHMIC_BYTE.get(d["precision"])
    out = {"bound": "hbm",
           "kernel": f"pdhg_persistent<{'double' if d['precision'] == 64 else 'float'}>",
           "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": achieved / PEAK_HBM_GBS,
           "traffic": ratio * per_launch if ratio else None,
           "algorithmic_bytes_per_launch": per_launch, "avg_launch_ms": avg_ms,
           "launches": d["launches"],
           "all_launches": {"kernel_ms": stats["kernel_ms"], "launches": stats["launches"],
                            "achieved_GBs": (stats["algorithmic_bytes"] / (stats["kernel_ms"] * 1e-3)
                                             / 1e9 if stats["kernel_ms"] else 0.0)}}
    if copy:
        out["achievable_copy"] = copy
        out["frac_of_copy"] = achieved / copy["GBs"]
    return out


def copy_calibration(k):
    """A plain copy of 2^26 doubles through the seam's copy op (512 MiB read,
    512 MiB written): the reference dispatch that scales FETCH_SIZE and
    WRITE_SIZE to bytes in a counter run of the same command."""
    n = 1 << 26
    a = k.upload(np.arange(n, dtype=np.float64))
    b = k.allocate(n)
    k.copy(a, b)
    k.dot(b, b)  # synchronises
    del a, b
    return {"elements": n, "bytes_read": 8 * n, "bytes_written": 8 * n}


def copy_bandwidth(k, reps: int = 20) -> dict:
    """Achievable copy over 2^26 doubles (512 MiB read and 512 MiB written):
    the seam's copy op (a device-to-device hipMemcpyAsync) and its scale op
    (the library's own streaming kernel), each wall-clocked over back-to-back
    launches between two synchronising one-element dot products."""
    n = 1 << 26
    a = k.upload(np.ones(n))
    b = k.allocate(n)
    s = k.allocate(1)

    def rate(op):
        op()
        k.dot(s, s)
        t = time.perf_counter()
        for _ in range(reps):
            op()
        k.dot(s, s)
        return 16 * n / ((time.perf_counter() - t) / reps) / 1e9

    memcpy = rate(lambda: k.copy(a, b))
    scale = rate(lambda: k.scale(1.0, a, b))
    del a, b, s
    return {"GBs": max(memcpy, scale), "hipMemcpy_d2d_GBs": memcpy, "scale_kernel_GBs": scale}


# ---------------------------------------------------------------------------
# the uneven scenario sweep (the headline)

def sweep_goldens() -> dict:
    with open(os.path.join(GOLDEN_DIR, "highs_sweep.json")) as f:
        return {d["args"]["k"]: d["objective"] for d in json.load(f) if d["status"] == 0}


def stratified(sizes, count):
    """Indices at the midpoints of `count` equal-count size strata."""
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i], i))
    return [order[int((j + 0.5) * len(order) / count)] for j in range(count)]


def cpu_baseline(batch, pick, fp64_iterations, seconds=12.0):
    """The oracle (C restatement of Prima's FP64 CPU path: ScalaKernels +
    Pdhg + Scaling + Kkt, -O3, no fast-math), one LP per core as
    PrimaZio.solveAll runs one LP per available processor (PrimaZio.scala:
    61-66): each sampled LP on CPU_CORES / len(pick) cores at once, each copy
    for about `seconds` of iterations. A whole sweep LP takes the CPU minutes
    to hours, so the figure is extrapolated: seconds per iteration on a loaded
    host times the iterations the same LP took in the device's FP64 solve."""
    from oracle.oracle import OracleLp, default_params
    cores = max(1, min(CPU_CORES, os.cpu_count() or 1))
    per = max(1, cores // len(pick))
    jobs = []
    for i in pick:
        p = batch[i]
        for _ in range(per):
            jobs.append((i, OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c,
                                              p.q, p.l, p.u, p.offset)))
    # a short probe sizes each copy's iteration count to about `seconds`
    probe = {}
    for i in pick:
        o = next(o for j, o in jobs if j == i)
        t = time.perf_counter()
        o.solve(default_params(max_iterations=64, eps_abs=1e-30, eps_rel=1e-30))
        probe[i] = max((time.perf_counter() - t) / 64, 1e-7)

    def run(job):
        i, o = job
        its = max(64, int(seconds / probe[i]) // 64 * 64)
        t = time.perf_counter()
        s = o.solve(default_params(max_iterations=its, eps_abs=1e-30, eps_rel=1e-30))
        return i, (time.perf_counter() - t) / max(1, s.iterations)

    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(jobs)) as ex:  # the oracle's C solve releases the GIL
        runs = list(ex.map(run, jobs))
    wall = time.perf_counter() - t0
    sec_per_it = {i: float(np.mean([r for j, r in runs if j == i])) for i in pick}
    per_lp = [sec_per_it[i] * fp64_iterations[i] for i in pick]
    value = cores / float(np.mean(per_lp))
    return dict(value=value, unit="LPs/s", cores=cores, kind="port",
                sample=(f"the denominator's {len(pick)} sampled LPs ({', '.join(str(i) for i in pick)}), "
                        f"{per} copies each on {len(jobs)} threads at once for {wall:.0f} s; "
                        f"extrapolated: seconds per iteration x the device FP64 iteration count "
                        f"of the same LP (mean {np.mean(per_lp):.0f} s per LP per core), one LP "
                        f"per core"),
                seconds_per_lp=dict(zip([str(i) for i in pick], per_lp)), cpu=cpu_model())


def headline(args, batch, gold, workload, sample_by_size):
    """Mixed-precision batched steps until the budget, then plain FP64
    batched, then plain FP64 one LP at a time on a stratified sample (the
    metric's denominator), then the CPU baseline; one JSON line."""
    t_all = time.perf_counter()
    B = len(batch)
    params = prima.PdhgParams(max_iterations=args.max_iterations)
    mixed = prima.MixedParams(reduced_tolerance=args.reduced_tolerance,
                              reduced_max_iterations=args.reduced_max_iterations,
                              stall_iterations=args.stall_iterations)
    k32 = prima.HipKernels(0, 32)
    k64 = prima.HipKernels(0, 64)
    try:
        copy_calibration(k64)  # scales FETCH_SIZE / WRITE_SIZE in a counter run
        warm = 0
        for _ in range(min(args.warmup, 1)):  # no JIT: one small solve per context
            for k in (k32, k64):
                s = k.solve(prima.config0(), prima.PdhgParams(eps_abs=1e-4, eps_rel=1e-4))
                if s.status != "Optimal":
                    raise SystemExit(f"warm-up solve ended {s.status}")
            warm += 1
        k32.reset_stats()
        steps, elapsed, worst, sols, first = 0, 0.0, 0.0, None, None
        while steps < max(1, args.steps):
            t = time.perf_counter()
            sols = run_watched(k32, lambda: k32.solve_all(batch, params, mixed=mixed), "step")
            dt = time.perf_counter() - t
            worst = max(worst, check(sols, range(B), gold, "mixed batched"))
            if first is None:
                first = sols
            elif any(a.objective_value != b.objective_value or a.iterations != b.iterations
                     for a, b in zip(first, sols)):
                raise SystemExit("mixed batched: a repeated step gave different results")
            steps += 1
            elapsed += dt
            log(f"step {steps}: {dt:.1f} s, worst gap vs HiGHS {worst:.2e}")
            if elapsed + dt > args.budget:
                break
        st = k32.stats()
        value = B * steps / elapsed
        copy = copy_bandwidth(k64)
        its = [(s.mixed["reduced_iterations"], s.mixed["refinement_iterations"],
                s.mixed["fallback_iterations"]) for s in sols]
        heavy = sorted(range(B), key=lambda i: -sols[i].total_time_ms)[:6]
        log("slowest LPs (n; f32/f64/cold f64 iterations; seconds): " + "; ".join(
            f"#{i} {batch[i].n}; {its[i][0]}/{its[i][1]}/{its[i][2]}; "
            f"{sols[i].total_time_ms / 1e3:.0f}" for i in heavy))

        # plain FP64, the whole batch in one call
        fp64_batched, s64b = None, None
        if not args.no_fp64_batched:
            k64.reset_stats()
            t = time.perf_counter()
            s64b = run_watched(k64, lambda: k64.solve_all(batch, params), "fp64 batched")
            tb = time.perf_counter() - t
            fp64_batched = {"value": B / tb, "unit": "LPs/s", "seconds": tb,
                            "iterations_mean": float(np.mean([s.iterations for s in s64b])),
                            "iterations_max": int(max(s.iterations for s in s64b)),
                            "roofline": roofline(k64.stats())}
            try:  # a comparison leg: an LP off the gate is reported, not fatal
                check(s64b, range(B), gold, "fp64 batched")
            except SystemExit as e:
                fp64_batched["gate_failure"] = str(e)
            log(f"fp64 batched: {fp64_batched}")

        # plain FP64 one LP at a time (the denominator) on a stratified
        # sample: by size, or, for LPs of one size, by their FP64 iteration
        # counts from the batched run (--unbatched 0 skips this leg)
        fp64, pick, s64 = None, [], {}
        if args.unbatched > 0:
            key = ([p.n for p in batch] if sample_by_size or s64b is None
                   else [s.iterations for s in s64b])
            pick = stratified(key, args.unbatched)
            k64.reset_stats()
            times = {}
            for i in pick:
                t = time.perf_counter()
                s64[i] = k64.solve(batch[i], params)
                times[i] = time.perf_counter() - t
                log(f"fp64 unbatched: LP {i} {s64[i].iterations} iterations, {times[i]:.1f} s")
            t64 = sum(times.values())
            gate = None
            try:
                check([s64[i] for i in pick], pick, gold, "fp64 unbatched")
            except SystemExit as e:
                gate = str(e)
            fp64 = {"value": len(pick) / t64, "unit": "LPs/s", "seconds": t64,
                    "sample": (f"LPs {pick}: midpoints of {len(pick)} equal-count strata by "
                               f"{'size' if sample_by_size or s64b is None else 'FP64 iteration count'}"),
                    "iterations": [s64[i].iterations for i in pick],
                    "seconds_each": [round(times[i], 3) for i in pick]}
            if gate:
                fp64["gate_failure"] = gate
            if s64b is not None:
                # every LP's batched FP64 iteration count times the sample's
                # seconds per iteration alone: the whole batch one at a time
                rate = t64 / sum(s64[i].iterations for i in pick)
                fp64["whole_batch_estimate"] = B / (rate * sum(s.iterations for s in s64b))
            log(f"fp64 unbatched sample: {fp64}")
        line = {
            "metric": METRIC, "value": value, "unit": "LPs/s", "n_gpus": 1, "steps": steps,
            "warmup": warm, "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f32 iterations, f64 final pass",
            "data": "synthetic",
            "config": dict(workload, tolerance=1e-8,
                           mixed={"reduced_tolerance": mixed.reduced_tolerance,
                                  "reduced_max_iterations": mixed.reduced_max_iterations,
                                  "stall_iterations": mixed.stall_iterations},
                           max_iterations=params.max_iterations),
            "fp64_unbatched": fp64,
            "speedup_vs_fp64_unbatched": value / fp64["value"] if fp64 else None,
            "fp64_batched": fp64_batched,
            "speedup_vs_fp64_batched": value / fp64_batched["value"] if fp64_batched else None,
            "worst_rel_gap_vs_highs": worst,
            "goldens_checked": sum(1 for i in range(B) if i in gold),
            "iterations_mean": [float(np.mean([a for a, _, _ in its])),
                                float(np.mean([b for _, b, _ in its])),
                                float(np.mean([c for _, _, c in its]))],
            "cold_fallbacks": int(sum(1 for _, _, c in its if c > 0)),
            "device_lps": B * steps / (st["kernel_ms"] * 1e-3) if st["kernel_ms"] else None,
            "roofline": roofline(st, copy),
        }
        if not args.no_cpu and pick:
            log("timing the CPU baseline (oracle)")
            try:
                line["cpu_baseline"] = cpu_baseline(batch, pick,
                                                    {i: s64[i].iterations for i in pick})
            except Exception as e:  # the line is printed whatever the host leg does
                line["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
        line["wall_seconds"] = time.perf_counter() - t_all
        print(json.dumps(line), flush=True)
    finally:
        k32.close()
        k64.close()


def sweep(args):
    """The uneven scenario sweep (prima.config3, k = 0..--batch-1)."""
    t = time.perf_counter()
    batch = generate(prima.config3, range(args.batch))
    n_tot, m_tot, z_tot = (sum(p.n for p in batch), sum(p.m for p in batch),
                           sum(p.nnz for p in batch))
    log(f"generated {len(batch)} sweep LPs in {time.perf_counter() - t:.1f} s: {n_tot} vars, "
        f"{m_tot} rows, {z_tot} nnz")
    headline(args, batch, sweep_goldens(),
             {"workload": f"uneven scenario sweep: {len(batch)} dispatch LPs, buses "
                          f"log-uniform in [50, 2000], snapshots in {{24, 72, 168}} "
                          f"({n_tot} vars, {m_tot} rows, {z_tot} nnz)", "lps": len(batch)},
             sample_by_size=True)


def sweep_fp64(args):
    """The uneven scenario sweep in plain FP64 batched (no float32 pass): the
    whole batch in one call, every LP checked against its HiGHS golden where
    one exists; one JSON line with the batch's LPs/s."""
    batch = generate(prima.config3, range(args.batch))
    gold = sweep_goldens()
    params = prima.PdhgParams(max_iterations=args.max_iterations)
    with prima.HipKernels(0, 64) as k:
        t = time.perf_counter()
        sols = run_watched(k, lambda: k.solve_all(batch, params), "fp64 sweep")
        dt = time.perf_counter() - t
        st = k.stats()
    worst, off = 0.0, []
    for i, x in enumerate(sols):
        g = rel_gap(x.objective_value, gold[i]) if i in gold else 0.0
        if x.status != "Optimal" or g > 1e-8:
            off.append({"lp": i, "status": x.status, "rel_gap_vs_highs": g if i in gold else None})
        else:
            worst = max(worst, g)
    print(json.dumps({
        "metric": "LPs/sec at 1e-8 rel. KKT, plain FP64 batched, 1 MI355X", "value": len(batch) / dt,
        "unit": "LPs/s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": 1e3 * dt,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"uneven scenario sweep: {len(batch)} dispatch LPs "
                               f"({sum(p.n for p in batch)} vars)", "tolerance": 1e-8,
                   "max_iterations": params.max_iterations},
        "iterations_mean": float(np.mean([x.iterations for x in sols])),
        "iterations_max": int(max(x.iterations for x in sols)),
        "worst_rel_gap_vs_highs": worst, "goldens_checked": sum(1 for i in range(len(batch)) if i in gold),
        "off": off, "roofline": roofline(st)}), flush=True)


def scenarios(args):
    """Demand scenarios of the 200-bus, 24-snapshot dispatch LP (the config-1
    LP: 34,800 vars, 14,400 rows, 64,750 nnz), per-bus demand x U[0.8, 1.2]
    with seed 1000 + k; every objective checked against HiGHS
    (tests/golden/highs_dispatch.json)."""
    with open(os.path.join(GOLDEN_DIR, "highs_dispatch.json")) as f:
        gold = {d["args"]["scenario_seed"] - 1000: d["objective"] for d in json.load(f)
                if d["args"].get("seed") == 1 and d["args"].get("perturb") == 0.2
                and "buses" not in d["args"]}
    batch = generate(lambda k: prima.dispatch(seed=1, perturb=0.2, scenario_seed=1000 + k),
                     range(args.batch))
    p = batch[0]
    headline(args, batch, gold,
             {"workload": f"demand scenarios of the 200-bus, 24-snapshot dispatch LP "
                          f"({p.n} vars, {p.m} rows, {p.nnz} nnz each)", "lps": len(batch),
              "note": "the 200-bus LP as a batch of demand scenarios; the uneven 256-LP "
                      "sweep needs more than 19 minutes to 1e-8 on one MI355X and runs as "
                      "--workload sweep"},
             sample_by_size=False)


# ---------------------------------------------------------------------------
# other workloads

def uc_frontier(args):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import milp_fixtures as mf
    root, g, node = mf.uc_frontier()
    B = min(args.batch, 512)
    nodes = [node(k) for k in range(B)]
    P = prima.PdhgParams(max_iterations=args.max_iterations)
    mixed = prima.MixedParams(reduced_tolerance=args.reduced_tolerance,
                              reduced_max_iterations=args.reduced_max_iterations,
                              stall_iterations=args.stall_iterations)
    with prima.HipKernels(0, 64) as k64, prima.HipKernels(0, 32) as k32:
        t = time.perf_counter()
        r = k64.solve(root, P)
        root_s = time.perf_counter() - t
        if r.status != "Optimal" or rel_gap(r.objective_value, g["root_objective"]) > 1e-8:
            raise SystemExit(f"root relaxation: {r.status} {r.objective_value!r}")
        log(f"root relaxation: {r.iterations} iterations, {root_s:.2f} s")
        warm = [prima.WarmStart.of(r)] * B
        k32.reset_stats()
        t0 = time.perf_counter()
        sols = run_watched(k32, lambda: k32.solve_all(nodes, P, mixed=mixed, warms=warm),
                           "frontier")
        elapsed = time.perf_counter() - t0
        worst, bad = 0.0, []
        for k, s in enumerate(sols):
            gold = g["nodes"][k]["objective"]
            if s.status != "Optimal" or rel_gap(s.objective_value, gold) > 1e-8:
                bad.append({"node": k, "status": s.status, "objective": s.objective_value,
                            "highs": gold, "rel_gap": rel_gap(s.objective_value, gold),
                            "iterations": s.iterations})
            else:
                worst = max(worst, rel_gap(s.objective_value, gold))
        if bad:
            log(f"frontier nodes off: {bad}")
        st = k32.stats()
    good = B - len(bad)
    print(json.dumps({
        "metric": "node LPs/sec at 1e-8 rel. KKT, batched B&B frontier, 1 MI355X",
        "value": good / elapsed, "unit": "LPs/s", "all_nodes_per_s": B / elapsed,
        "nodes_within_1e-8_of_highs": good, "off_nodes": bad, "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * elapsed, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 iterations, f64 final pass", "data": "synthetic",
        "config": {"workload": "unit-commitment B&B frontier: 100 buses x 24 snapshots "
                               "(37,200 vars, 49,200 rows, 154,800 nnz), depth-9 frontier, "
                               "warm-started from the root", "batch": B, "tolerance": 1e-8},
        "root_iterations": r.iterations, "root_seconds": root_s,
        "iterations_mean": [float(np.mean([s.mixed["reduced_iterations"] for s in sols])),
                            float(np.mean([s.mixed["refinement_iterations"] for s in sols])),
                            float(np.mean([s.mixed["fallback_iterations"] for s in sols]))],
        "iterations_max": int(max(s.iterations for s in sols)),
        "worst_rel_gap_vs_highs": worst, "roofline": roofline(st)}), flush=True)


def kernel_profile(args):
    B = args.batch
    batch = generate(lambda k: prima.dispatch(seed=1, perturb=0.2, scenario_seed=1000 + k),
                     range(B))
    out = {"workload": f"{B} dispatch LPs (200 buses x 24 snapshots), {args.iterations} "
                       f"iterations each, tolerance 1e-12"}
    with prima.HipKernels(0, 64) as k64, prima.HipKernels(0, 32) as k32:
        out["copy"] = copy_calibration(k64)
        for prec, kk in ((64, k64), (32, k32)):
            P = prima.PdhgParams(eps_abs=1e-12, eps_rel=1e-12, max_iterations=args.iterations,
                                 workgroups_per_lp=args.wg)
            kk.reset_stats()
            t = time.perf_counter()
            sols = kk.solve_all(batch, P)
            wall = time.perf_counter() - t
            st = kk.stats()
            its = sum(s.iterations for s in sols)
            out[f"f{prec}"] = dict(roofline(st), wall_s=wall,
                                   lp_iterations_per_s=its / (st["kernel_ms"] * 1e-3),
                                   workgroups_per_lp=sols[0].workgroups)
            log(f"f{prec}: {out[f'f{prec}']}")
    print(json.dumps(out), flush=True)


def lopf_scigrid(args):
    """scigrid-de, PyPSA's German transmission example (60,552 variables,
    23,688 rows, 24 snapshots): read from its CSV export, built into one LP
    and solved on the device in float64 and in mixed precision, against
    HiGHS's objective on the same LP."""
    from noaidi_amd import lopf, network
    gold_dir = os.path.join(GOLDEN_DIR, "pypsa")
    with open(os.path.join(gold_dir, "highs.json")) as f:
        target = json.load(f)["scigrid-de"]["objective"]
    net = network.read_folder(os.path.join(gold_dir, "scigrid-de"))
    p = lopf.build(net).problem
    out = {}
    for label, prec, mixed in (("f64", 64, None),
                               ("mixed", 32, prima.MixedParams(reduced_tolerance=1e-6,
                                                               reduced_max_iterations=60_000))):
        with prima.HipKernels(0, prec) as k:
            P = prima.PdhgParams(max_iterations=args.max_iterations)
            t = time.perf_counter()
            s = run_watched(k, (lambda: k.mixed_solve(p, P, mixed)) if mixed else
                            (lambda: k.solve(p, P)), label)
            el = time.perf_counter() - t
        out[label] = {"status": s.status, "seconds": el, "iterations": s.iterations,
                      "rel_gap_vs_highs": rel_gap(s.objective_value, target)}
        log(f"scigrid-de {label}: {out[label]}")
    print(json.dumps({"metric": "PyPSA LOPF solve time, scigrid-de, 1 MI355X",
                      "value": out["mixed"]["seconds"], "unit": "s", "n_gpus": 1, "steps": 1,
                      "warmup": 0, "ms_per_step": 1e3 * out["mixed"]["seconds"],
                      "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
                      "dtype": "f32 iterations, f64 final pass",
                      "data": "PyPSA scigrid-de CSV export",
                      "config": {"workload": f"scigrid-de LOPF ({p.n} vars, {p.m} rows, "
                                             f"{p.nnz} nnz)", "tolerance": 1e-8},
                      "solves": out}), flush=True)


SCHEDULES = {0: "auto", 1: "vector-1", 2: "vector-2", 3: "vector-4", 4: "vector-8",
             5: "vector-16", 6: "vector-32", 7: "vector-64", 8: "row-blocks", 9: "panels",
             10: "row-blocks-small"}


def tune_report(args):
    """The solver autotuner (prima_tune_solver) on three matrix families: the
    200-bus dispatch LP, the largest LP of the uneven sweep and
    the 2M-variable random LP; --iterations accepted iterations per candidate
    in float64. Prints every candidate's device time per iteration."""
    cases = [("dispatch 200 buses x 24 snapshots", lambda: prima.config1()),
             ("sweep LP 49 (1,988 buses x 168 snapshots)", lambda: prima.config3(49)),
             ("random 2M x 1M", lambda: prima.config2())]
    out = {"iterations_per_candidate": args.iterations, "precision": 64, "cases": []}
    with prima.HipKernels(0, 64) as k:
        for name, make in cases:
            p = make()
            t = time.perf_counter()
            r = k.tune_solver(p, prima.PdhgParams(), iterations=args.iterations)
            rec = {"lp": name, "n": p.n, "m": p.m, "nnz": p.nnz,
                   "chosen": {"workgroups": r["workgroups"],
                              "blocks_per_workgroup": r["blocks_per_workgroup"],
                              "us_per_iteration": r["us_per_iteration"]},
                   "candidates": [{"workgroups": g, "blocks_per_workgroup": b,
                                   "us_per_iteration": u} for g, b, u in r["candidates"]],
                   "seconds": time.perf_counter() - t}
            log(json.dumps(rec))
            out["cases"].append(rec)
    print(json.dumps(out), flush=True)


def random_lp_passes(args):
    """The large random LP (2M variables, 1M rows, 20M nonzeros, lognormal
    row lengths capped at 2048, known optimum c'x* = b'y*) in explicit
    mixed-precision passes, each under a time limit: a float32 pass to 1e-6,
    a float64 pass to 1e-6 warm-started from it, a float64 pass to 1e-8
    warm-started from that. Each pass reports its status, iterations, time
    and the relative error of its objective against the known optimum."""
    t = time.perf_counter()
    p = prima.config2()
    log(f"random LP: n={p.n} m={p.m} nnz={p.nnz} known={p.known_objective!r} "
        f"(built in {time.perf_counter() - t:.1f} s)")
    lim = args.pass_seconds

    def rep(what, s, dt):
        return {"pass": what, "status": s.status, "iterations": s.iterations, "seconds": dt,
                "ms_per_iteration": 1e3 * dt / max(1, s.iterations),
                "objective": s.objective_value, "dual_objective": s.dual_objective_value,
                "rel_err_vs_known": rel_gap(s.objective_value, p.known_objective),
                "primal_residual": s.primal_residual, "dual_residual": s.dual_residual,
                "workgroups": s.workgroups}

    out = {"workload": f"random LP: {p.n} vars, {p.m} rows, {p.nnz} nnz",
           "known_objective": p.known_objective, "passes": []}
    with prima.HipKernels(0, 32) as k32, prima.HipKernels(0, 64) as k64:
        plan = [(k32, "float32 to 1e-6", 1e-6, lim[0]), (k64, "float64 to 1e-6, warm", 1e-6, lim[1]),
                (k64, "float64 to 1e-8, warm", 1e-8, lim[2])]
        warm = None
        for k, what, eps, sec in plan:
            P = prima.PdhgParams(eps_abs=eps, eps_rel=eps, max_iterations=4_000_000,
                                 time_limit_ms=int(sec * 1e3))
            t = time.perf_counter()
            s = run_watched(k, lambda: k.solve(p, P, warm=warm), what)
            r = rep(what, s, time.perf_counter() - t)
            log(json.dumps(r))
            out["passes"].append(r)
            warm = prima.WarmStart.of(s)
    print(json.dumps(out), flush=True)


def spmv_roofline(args, reps: int = 20):
    """The per-op SpMV at the metric's scales, every schedule timed by the
    autotuner (HIP events), algorithmic bytes per SpMV Z(s+4) + 4(R+1) +
    s*C + s*R: (1) A of the 2M-variable, 1M-row random LP (lognormal row
    lengths, sigma 1, mean about 20, capped at 2048; uniform random
    columns); (2) the constraint matrices of all 256 LPs of the uneven
    scenario sweep side by side (block diagonal: the rows and columns one
    batched iteration multiplies), and its transpose."""
    p = prima.random_known(2, 2_000_000, 1_000_000, row_len=0, mu=math.log(20.0) - 0.5,
                           sigma=1.0, max_len=2048)
    mats = [("random 1M x 2M", p.m, p.n, p.row_ptr, p.col, p.val, p.to_scipy)]
    if args.batch > 0:
        sw = generate(prima.config3, range(args.batch))
        ro = np.cumsum([0] + [q.m for q in sw])
        co = np.cumsum([0] + [q.n for q in sw])
        zo = np.cumsum([0] + [q.nnz for q in sw])
        ptr = np.concatenate([np.asarray(q.row_ptr[:-1], np.int64) + zo[j]
                              for j, q in enumerate(sw)] + [np.array([zo[-1]])]).astype(np.int32)
        col = np.concatenate([np.asarray(q.col, np.int64) + co[j]
                              for j, q in enumerate(sw)]).astype(np.int32)
        val = np.concatenate([np.asarray(q.val) for q in sw])
        del sw
        import scipy.sparse as sp
        S = sp.csr_matrix((val, col, ptr), shape=(int(ro[-1]), int(co[-1])))
        T = S.T.tocsr()
        mats.append((f"sweep block diagonal ({args.batch} LPs)", S.shape[0], S.shape[1],
                     S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data, lambda: S))
        mats.append((f"sweep block diagonal ({args.batch} LPs), transposed", T.shape[0],
                     T.shape[1], T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data,
                     lambda: T))
    out = {}
    with prima.HipKernels(0, 64) as kc:
        out["achievable_copy"] = copy_bandwidth(kc)
    for name, R, Cn, rp, ci, vv, sci in mats:
        Z = len(vv)
        rec = {"rows": R, "cols": Cn, "nnz": Z}
        xh = np.random.default_rng(0).standard_normal(Cn)
        ref = sci() @ xh
        mag = abs(sci()) @ np.abs(xh) + 1e-30
        for prec in (32, 64):
            with prima.HipKernels(0, prec) as k:
                A = k.uploadMatrix(R, Cn, rp, ci, vv)
                x, y = k.upload(xh), k.allocate(R)
                choice, best_ms = k.autotune(A, reps)
                sb = 4 if prec == 32 else 8
                bytes_ = Z * (sb + 4) + 4 * (R + 1) + sb * Cn + sb * R
                k.set_schedule(A, choice)
                k.spmv(A, x, y)
                got = k.download(y)
                err = float(np.max(np.abs(got - ref) / mag))
                rec[f"f{prec}"] = {"schedule": SCHEDULES[choice], "ms": best_ms,
                                   "achieved_GBs": bytes_ / (best_ms * 1e-3) / 1e9,
                                   "frac_of_8TBs": bytes_ / (best_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                   "algorithmic_bytes": bytes_, "max_rel_err_vs_f64_host": err}
                log(f"spmv {name} f{prec}: {rec[f'f{prec}']}")
                del A, x, y
        out[name] = rec
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--budget", type=float, default=150.0,
                    help="seconds of timed steps after which no further step starts")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--unbatched", type=int, default=8,
                    help="strata sampled for the FP64 one-at-a-time denominator")
    ap.add_argument("--wg", type=int, default=0, help="workgroups per LP (0: scheduler)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-batched", action="store_true",
                    help="skip solving the whole batch in plain FP64 batched")
    ap.add_argument("--max-iterations", type=int, default=4_000_000)
    ap.add_argument("--reduced-tolerance", type=float, default=1e-6)
    ap.add_argument("--reduced-max-iterations", type=int, default=60_000)
    ap.add_argument("--stall-iterations", type=int, default=1_000_000)
    ap.add_argument("--iterations", type=int, default=2048,
                    help="iterations per LP for --workload kernel")
    ap.add_argument("--pass-seconds", type=float, nargs=3, default=[240.0, 186.7, 373.3],
                    help="time limits of the three passes of --workload random-lp")
    ap.add_argument("--workload", default="scenarios",
                    choices=["sweep", "sweep-fp64", "scenarios", "uc-frontier", "kernel", "lopf",
                             "spmv", "random-lp", "tune"])
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on exactly one GPU: this build solves independent LPs per GPU "
              "and has no multi-GPU path to measure", file=sys.stderr)
        sys.exit(2)
    {"sweep": sweep, "sweep-fp64": sweep_fp64, "scenarios": scenarios, "uc-frontier": uc_frontier,
     "kernel": kernel_profile, "lopf": lopf_scigrid, "spmv": spmv_roofline,
     "random-lp": random_lp_passes, "tune": tune_report}[args.workload](args)


if __name__ == "__main__":
    main()
