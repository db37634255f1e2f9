"""The benchmark's contract on the CPU: it refuses more than one GPU with a
message, and every scenario of the headline batch has a HiGHS golden to be
checked against."""
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


def test_every_headline_scenario_has_a_golden():
    import bench
    gold = bench.golden_objectives()
    assert sorted(gold) == list(range(256))
    assert bench.rel_gap(1.0 + 1e-9, 1.0) <= 1e-8 < bench.rel_gap(1.0 + 2e-8, 1.0)
    assert bench.rel_gap(1e-3, 0.0) == 1e-3  # |a-b| / max(1, |b|)
