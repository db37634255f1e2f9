"""Failure handling and synchronisation of the persistent solver on the device:
an iteration cap that falls on a checkpoint with a multi-workgroup group, a
group barrier that gives up (forced from the host), and cancellation of a
running solve from another thread (PrimaZioSuite.scala:40-58)."""
import threading
import time

import numpy as np
import pytest

from noaidi_amd import prima
from oracle.oracle import OracleLp

pytestmark = pytest.mark.gpu


def oracle_of(p):
    return OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                             p.offset)


@pytest.mark.parametrize("prec", [64, 32])
def test_cap_on_a_checkpoint_with_a_group(hip64, hip32, prec):
    # max_iterations a multiple of evaluationFrequency (64): the checkpoint
    # and the iteration-limit evaluation run back to back on the same
    # partial-sum slots (Pdhg.scala:236-244). Eight workgroups per LP.
    k = hip64 if prec == 64 else hip32
    p = prima.config1()
    P = prima.PdhgParams(eps_abs=1e-14, eps_rel=1e-14, max_iterations=640, workgroups_per_lp=8)
    a = k.solve(p, P)
    b = k.solve(p, P)
    assert a.status == "IterationLimit" and a.iterations == 640 and a.workgroups == 8
    # deterministic: every workgroup took the same decisions both times
    assert np.array_equal(a.primal, b.primal) and np.array_equal(a.dual, b.dual)
    assert a.objective_value == b.objective_value
    # the whole primal and dual were written at the returned point: the
    # oracle's Kkt.evaluate from the vectors alone gives the reported values
    e = oracle_of(p).kkt(a.primal, a.dual)
    assert abs(e[2] - a.objective_value) <= 1e-9 * max(1.0, abs(e[2]))
    assert abs(e[3] - a.dual_objective_value) <= 1e-6 * max(1.0, abs(e[3]))
    assert abs(e[0] - a.primal_residual) <= 1e-6 * max(1.0, e[0])
    # float32 iterates are clamped to float32-rounded scaled bounds
    room = (1e-12 if prec == 64 else 1e-6) * (1.0 + np.abs(a.primal))
    assert np.all(a.primal >= p.l - room) and np.all(a.primal <= p.u + room)
    assert np.count_nonzero(a.primal) > p.n // 4


def test_barrier_that_gives_up_is_a_device_error(hip64):
    # the host forces the last workgroup of every group to wait on a word
    # nothing writes, with a short limit: its wait gives up, the group's other
    # workgroups give up at their next barrier, every workgroup exits, and the
    # call fails with PRIMA_EDEVICE rather than returning a SolveStatus
    p = prima.config1()
    P = prima.PdhgParams(max_iterations=10_000, workgroups_per_lp=4)
    hip64._test_stall_barriers(True)
    try:
        with pytest.raises(prima.DeviceError, match="group barrier"):
            hip64.solve_all([p, p], P)
    finally:
        hip64._test_stall_barriers(False)
    # the context and the device are still usable
    s = hip64.solve(prima.config0(), prima.PdhgParams(workgroups_per_lp=4))
    assert s.status == "Optimal"
    assert abs(s.objective_value - prima.config0().known_objective) <= 1e-8 * max(
        1.0, abs(s.objective_value))


def test_cancel_stops_a_long_solve(hip64):
    # PrimaZioSuite.scala:40-58: tolerances that cannot be met, so the solve
    # would run for its whole budget; another thread cancels it
    p = prima.config1()
    stubborn = prima.PdhgParams(eps_abs=1e-16, eps_rel=1e-16, max_iterations=50_000_000)
    out = {}

    def run():
        out["s"] = hip64.solve(p, stubborn)

    th = threading.Thread(target=run)
    th.start()
    time.sleep(1.0)
    t = time.perf_counter()
    hip64.cancel()
    th.join(30.0)
    waited = time.perf_counter() - t
    assert not th.is_alive()
    s = out["s"]
    assert s.status == "Interrupted", s.status
    # one launch boundary (chunks of about 50 ms) plus the final evaluation
    assert waited < 2.0, waited
    assert 0 < s.iterations < 50_000_000
    e = oracle_of(p).kkt(s.primal, s.dual)
    assert abs(e[2] - s.objective_value) <= 1e-9 * max(1.0, abs(e[2]))
