"""The allocation-free solve (prima_solve_noalloc) against prima_solve: two
calls in a row through the C ABI, the second on a smaller LP, in both
precisions, must give the same results bit for bit (same device code, same
group size, same arithmetic); an LP larger than the workspace is refused."""
import numpy as np
import pytest

from noaidi_amd import prima

pytestmark = pytest.mark.gpu


def same(a, b):
    assert a.status == b.status and a.iterations == b.iterations and a.restarts == b.restarts
    assert a.objective_value == b.objective_value
    assert a.dual_objective_value == b.dual_objective_value
    assert np.array_equal(a.primal, b.primal) and np.array_equal(a.dual, b.dual)
    assert np.array_equal(a.reduced_costs, b.reduced_costs)
    assert a.workgroups == b.workgroups


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_solve_matches_solve(prec):
    with prima.HipKernels(0, prec) as k:
        tol = 1e-8 if prec == 64 else 1e-5
        P = prima.PdhgParams(eps_abs=tol, eps_rel=tol, max_iterations=400_000)
        # the 200-bus dispatch LP, then the smaller config-0 LP
        for p in (prima.config1(), prima.config0(4)):
            a = k.solve(p, P)
            b = k.solve_noalloc(p, P)
            same(a, b)
            assert b.status == "Optimal"
            if prec == 64 and np.isfinite(p.known_objective):
                assert abs(b.objective_value - p.known_objective) <= 1e-8 * max(
                    1.0, abs(p.known_objective))
        # a warm start goes through as well
        p = prima.config0(5)
        w = prima.WarmStart.of(k.solve(p, prima.PdhgParams(eps_abs=1e-4, eps_rel=1e-4)))
        same(k.solve(p, P, w), k.solve_noalloc(p, P, w))


def test_noalloc_refuses_an_lp_beyond_its_workspace():
    with prima.HipKernels(0, 64) as k:
        big = prima.random_known(9, 400_000, 200_000, row_len=12)
        with pytest.raises(ValueError, match="allocation-free solve holds at most"):
            k.solve_noalloc(big, prima.PdhgParams(max_iterations=64))
        # the context still works afterwards
        s = k.solve_noalloc(prima.config0(6), prima.PdhgParams())
        assert s.status == "Optimal"
