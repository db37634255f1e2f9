"""The benchmark's contract on the CPU: it refuses more than one GPU with a
message, it offers every workload the measurements name, the uneven sweep
has HiGHS goldens across its horizons, and the FP64 denominator samples
size strata at their midpoints."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""


def test_every_measured_workload_is_offered():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for w in ("scenarios", "sweep", "sweep-fp64", "uc-frontier", "kernel", "lopf", "spmv", "random-lp",
              "tune"):
        assert w in r.stdout, w
    assert "--pass-seconds" in r.stdout and "--unbatched" in r.stdout


def test_sweep_goldens_cover_every_horizon():
    # the uneven sweep (--workload sweep): HiGHS goldens include instances
    # of every snapshot count, and the relative gap is |a-b| / max(1, |b|)
    import bench
    from noaidi_amd import prima
    gold = bench.sweep_goldens()
    assert len(gold) >= 20
    assert {prima.config3_args(k)["snapshots"] for k in gold} == {24, 72, 168}
    assert bench.rel_gap(1.0 + 1e-9, 1.0) <= 1e-8 < bench.rel_gap(1.0 + 2e-8, 1.0)
    assert bench.rel_gap(1e-3, 0.0) == 1e-3


def test_stratified_sample_takes_strata_midpoints():
    import bench
    sizes = [50, 10, 40, 20, 30, 80, 70, 60]
    pick = bench.stratified(sizes, 4)
    # sorted sizes 10 20 30 40 | 50 60 70 80: strata of two, midpoints the upper
    assert [sizes[i] for i in pick] == [20, 40, 60, 80]
    assert len(set(bench.stratified(list(range(256)), 4))) == 4


def test_roofline_arithmetic():
    # the dominant instantiation by device time; achieved = algorithmic bytes
    # per launch / average launch duration; frac against 8 TB/s; traffic is
    # the PMC factor times the bytes per launch (float64 only has one)
    import bench
    stats = {"kernel_ms_by_precision": [10.0, 40.0], "launches_by_precision": [2, 4],
             "algorithmic_bytes_by_precision": [1e9, 8e10], "kernel_ms": 50.0, "launches": 6,
             "algorithmic_bytes": 8.1e10}
    r = bench.roofline(stats)
    assert r["kernel"] == "pdhg_persistent<double>" and r["launches"] == 4
    assert abs(r["avg_launch_ms"] - 10.0) < 1e-12
    assert abs(r["algorithmic_bytes_per_launch"] - 2e10) < 1e-3
    assert abs(r["achieved"] - 2000.0) < 1e-9 and abs(r["frac"] - 0.25) < 1e-12
    assert abs(r["traffic"] - bench.TRAFFIC_PER_ALGORITHMIC_BYTE[64] * 2e10) < 1e-3
    f32 = dict(stats, kernel_ms_by_precision=[40.0, 10.0])
    r = bench.roofline(f32)
    assert r["kernel"] == "pdhg_persistent<float>" and r["traffic"] is None
