"""CPU-side checks of the C-ABI boundary: the built library loads and
exports every function include/prima_hip.h declares, the ctypes mirror binds
them, and the defaults match the reference's PdhgParams / MixedPrecision.Params.
No device calls are made here."""
import ctypes as C
import os
import re

import pytest

from noaidi_amd import prima

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "prima_hip.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(prima_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(prima.LIB_PATH)
    names = declared()
    assert len(names) >= 28
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(prima.EXPORTED) == names


def test_ctypes_bindings_load():
    L = prima.lib()
    for n in prima.EXPORTED:
        assert getattr(L, n).argtypes is not None


def test_defaults_match_reference():
    # PdhgParams.scala:9-59, Scaling.scala:11-24
    p = prima._Params()
    prima.lib().prima_default_params(C.byref(p))
    assert (p.eps_abs, p.eps_rel, p.max_iterations, p.evaluation_frequency) == (1e-8, 1e-8,
                                                                                 100000, 64)
    assert p.time_limit_ms == -1 and p.ruiz_iterations == 10 and p.apply_pock_chambolle == 1
    assert p.pock_chambolle_alpha == 1.0 and p.infeasibility_tolerance == 1e-8
    assert (p.restart_sufficient, p.restart_necessary, p.restart_artificial,
            p.primal_weight_smoothing) == (0.2, 0.8, 0.36, 0.5)
    # MixedPrecision.scala:37-66
    m = prima._Mixed()
    prima.lib().prima_default_mixed_params(C.byref(m))
    assert (m.reduced_tolerance, m.reduced_max_iterations, m.always_refine,
            m.reduced_time_share) == (1e-5, 20000, 0, 0.5)
    # the Python mirror's defaults agree with the library's
    q = prima.PdhgParams()._c()
    for f, _ in prima._Params._fields_:
        assert getattr(q, f) == getattr(p, f), f


def test_bnb_defaults_match_reference():
    # BnbParams (BranchAndBound.scala:57-108); frontier 1 is the sequential search
    b = prima._BnbParams()
    prima.lib().prima_default_bnb_params(C.byref(b))
    assert (b.integrality_tolerance, b.gap_abs, b.gap_rel, b.pruning_safety_factor) == (
        1e-6, 1e-6, 1e-6, 10.0)
    assert (b.max_nodes, b.warm_start, b.time_limit_ms, b.frontier) == (100000, 1, -1, 1)
    assert b.lp.eps_abs == 1e-8 and b.lp.max_iterations == 100000
    q = prima.BnbParams()._c()
    for f, _ in prima._BnbParams._fields_:
        if f != "lp":
            assert getattr(q, f) == getattr(b, f), f


def test_bnb_predicates_in_the_library():
    # BranchAndBoundSuite.scala:130-158, evaluated by the C++ driver's own helpers
    nan, inf = float("nan"), float("inf")
    assert not prima.skips_before_solving(nan, 10.0, True)
    assert not prima.skips_before_solving(-inf, 10.0, True)
    assert prima.skips_before_solving(11.0, 10.0, True)
    assert not prima.skips_before_solving(9.0, 10.0, True)
    assert not prima.skips_before_solving(11.0, inf, False)
    assert prima.safe_bound_of(True, 5.0, 1.0) == 4.0
    assert prima.safe_bound_of(True, nan, 1.0) == -inf
    assert prima.safe_bound_of(True, 5.0, nan) == -inf
    assert prima.safe_bound_of(False, 5.0, 1.0) == -inf


def test_struct_layouts_match_header(tmp_path):
    # compile the header and compare every field offset with the ctypes mirror
    structs = {"prima_lp": prima._Lp, "prima_params": prima._Params,
               "prima_result": prima._Result, "prima_warm": prima._Warm,
               "prima_mixed_params": prima._Mixed, "prima_mixed_info": prima._MixedInfo,
               "prima_stats": prima._Stats, "prima_capabilities": prima._Caps,
               "prima_bnb_params": prima._BnbParams, "prima_milp_result": prima._MilpResult,
               "prima_tune_result": prima._TuneResult}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "prima_hip.h"',
             "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                 check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_null_context_is_rejected_without_a_device():
    L = prima.lib()
    assert L.prima_close(None) == prima.PRIMA_EINVAL
    assert L.prima_cancel(None) == prima.PRIMA_EINVAL
    assert L.prima_vec_length(None) == -1
    assert L.prima_tune_solver(None, None, None, 1, None) == prima.PRIMA_EINVAL


def test_builder_validation():
    b = prima.builder(1)
    with pytest.raises(ValueError):
        b.bounds(0, 1.0, 0.0)
    with pytest.raises(ValueError):
        b.constraint([(0, 1.0)], float("-inf"), float("inf"))
    with pytest.raises(ValueError):
        b.objectiveCoefficient(3, 1.0)
