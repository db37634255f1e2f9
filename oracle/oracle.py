"""ctypes binding to the CPU oracle (oracle/_build/libprima_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, always as the checker or the baseline,
never as the measured product path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libprima_oracle.so")

STATUS_NAMES = [
    "Optimal",
    "PrimalInfeasible",
    "DualInfeasible",
    "IterationLimit",
    "TimeLimit",
    "NumericalError",
    "Interrupted",
]


class Params(C.Structure):
    _fields_ = [
        ("eps_abs", C.c_double),
        ("eps_rel", C.c_double),
        ("max_iterations", C.c_int32),
        ("evaluation_frequency", C.c_int32),
        ("time_limit_ms", C.c_int64),
        ("ruiz_iterations", C.c_int32),
        ("apply_pock_chambolle", C.c_int32),
        ("pock_chambolle_alpha", C.c_double),
        ("infeasibility_tolerance", C.c_double),
        ("restart_sufficient", C.c_double),
        ("restart_necessary", C.c_double),
        ("restart_artificial", C.c_double),
        ("primal_weight_smoothing", C.c_double),
    ]


class Result(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("iterations", C.c_int32),
        ("restarts", C.c_int32),
        ("pad", C.c_int32),
        ("objective", C.c_double),
        ("dual_objective", C.c_double),
        ("primal_residual", C.c_double),
        ("dual_residual", C.c_double),
        ("solve_time_ms", C.c_double),
        ("final_step_size", C.c_double),
        ("final_primal_weight", C.c_double),
    ]


class MixedParams(C.Structure):
    _fields_ = [
        ("reduced_tolerance", C.c_double),
        ("reduced_max_iterations", C.c_int32),
        ("always_refine", C.c_int32),
        ("reduced_time_share", C.c_double),
    ]


class MixedInfo(C.Structure):
    _fields_ = [
        ("reduced_iterations", C.c_int32),
        ("reduced_restarts", C.c_int32),
        ("refinement_iterations", C.c_int32),
        ("refined", C.c_int32),
        ("reduced_ms", C.c_double),
        ("refinement_ms", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        L.orc_lp_from_csr.restype = P
        L.orc_lp_from_csr.argtypes = [C.c_int32, C.c_int32, C.c_int32, ip, ip, dp, dp, dp, dp, dp,
                                      C.c_double]
        L.orc_lp_from_triplets.restype = P
        L.orc_lp_from_triplets.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, ip, ip, dp,
                                           dp, dp, dp, dp, C.c_double]
        L.orc_lp_free.argtypes = [P]
        L.orc_lp_dims.argtypes = [P, ip, ip, ip, C.POINTER(C.c_int64)]
        L.orc_lp_export.argtypes = [P, ip, ip, dp, dp, dp, dp, dp, dp]
        L.orc_fixture.restype = P
        L.orc_fixture.argtypes = [C.c_char_p]
        L.orc_random_feasible.restype = P
        L.orc_random_feasible.argtypes = [C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_double]
        L.orc_jrandom_init.argtypes = [P, C.c_int64]
        L.orc_jrandom_next_int.restype = C.c_int32
        L.orc_jrandom_next_int.argtypes = [P, C.c_int32]
        L.orc_jrandom_next_double.restype = C.c_double
        L.orc_jrandom_next_double.argtypes = [P]
        L.orc_jrandom_next_gaussian.restype = C.c_double
        L.orc_jrandom_next_gaussian.argtypes = [P]
        L.orc_fdlibm_log.restype = C.c_double
        L.orc_fdlibm_log.argtypes = [C.c_double]
        L.orc_spmv.argtypes = [P, C.c_int32, dp, dp]
        L.orc_spectral_norm_bound.restype = C.c_double
        L.orc_spectral_norm_bound.argtypes = [P]
        L.orc_scaling.argtypes = [P, C.c_int32, C.c_int32, C.c_double, dp, dp, dp]
        L.orc_kkt_evaluate.argtypes = [P, dp, dp, dp, dp, dp]
        L.orc_kkt_converged.restype = C.c_int32
        L.orc_kkt_converged.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_primal_infeasibility.restype = C.c_int32
        L.orc_primal_infeasibility.argtypes = [P, dp, dp, dp]
        L.orc_dual_infeasibility.restype = C.c_int32
        L.orc_dual_infeasibility.argtypes = [P, dp, dp, dp]
        L.orc_classify.restype = C.c_int32
        L.orc_classify.argtypes = [P, dp, dp, C.c_double]
        L.orc_default_params.argtypes = [C.POINTER(Params)]
        L.orc_default_mixed_params.argtypes = [C.POINTER(MixedParams)]
        L.orc_pdhg_solve.restype = C.c_int32
        L.orc_pdhg_solve.argtypes = [P, C.POINTER(Params), C.c_int32, dp, dp, C.c_double,
                                     C.c_double, C.POINTER(Result), dp, dp, dp]
        L.orc_mixed_solve.restype = C.c_int32
        L.orc_mixed_solve.argtypes = [P, C.POINTER(Params), C.c_int32, C.POINTER(MixedParams),
                                      C.POINTER(Result), C.POINTER(MixedInfo), dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def default_params(**kw) -> Params:
    p = Params()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def default_mixed_params(**kw) -> MixedParams:
    p = MixedParams()
    lib().orc_default_mixed_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@dataclass
class Solution:
    status: str
    iterations: int
    restarts: int
    objective: float
    dual_objective: float
    primal_residual: float
    dual_residual: float
    solve_time_ms: float
    final_step_size: float
    final_primal_weight: float
    primal: np.ndarray
    dual: np.ndarray
    reduced_costs: np.ndarray
    mixed: MixedInfo | None = None


class OracleLp:
    """An LpProblem held by the oracle (LpProblem.scala:62-104)."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("the oracle rejected the LP (LpProblem requirements)")
        self.h = handle
        n, m, ne = C.c_int32(), C.c_int32(), C.c_int32()
        nnz = C.c_int64()
        lib().orc_lp_dims(self.h, C.byref(n), C.byref(m), C.byref(ne), C.byref(nnz))
        self.n, self.m, self.n_eq, self.nnz = n.value, m.value, ne.value, nnz.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_lp_free(self.h)
            self.h = None

    @classmethod
    def fixture(cls, name: str) -> "OracleLp":
        return cls(lib().orc_fixture(name.encode()))

    @classmethod
    def random_feasible(cls, seed, n, n_eq, n_ineq, density=0.3) -> "OracleLp":
        return cls(lib().orc_random_feasible(seed, n, n_eq, n_ineq, density))

    @classmethod
    def from_csr(cls, n, m, n_eq, row_ptr, col, val, c, q, l, u, offset=0.0) -> "OracleLp":
        arrs = [np.ascontiguousarray(a, dtype=np.int32) for a in (row_ptr, col)]
        darr = [np.ascontiguousarray(a, dtype=np.float64) for a in (val, c, q, l, u)]
        return cls(lib().orc_lp_from_csr(n, m, n_eq, _ip(arrs[0]), _ip(arrs[1]), *map(_dp, darr),
                                         offset))

    @classmethod
    def from_triplets(cls, n, m, n_eq, rows, cols, vals, c, q, l, u, offset=0.0) -> "OracleLp":
        r = np.ascontiguousarray(rows, dtype=np.int32)
        cc = np.ascontiguousarray(cols, dtype=np.int32)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        darr = [np.ascontiguousarray(a, dtype=np.float64) for a in (c, q, l, u)]
        return cls(lib().orc_lp_from_triplets(n, m, n_eq, len(r), _ip(r), _ip(cc), _dp(v),
                                              *map(_dp, darr), offset))

    def export(self):
        rp = np.zeros(self.m + 1, np.int32)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.float64)
        c = np.zeros(max(self.n, 1))
        l = np.zeros(max(self.n, 1))
        u = np.zeros(max(self.n, 1))
        q = np.zeros(max(self.m, 1))
        off = C.c_double()
        lib().orc_lp_export(self.h, _ip(rp), _ip(col), _dp(val), _dp(c), _dp(q), _dp(l), _dp(u),
                            C.byref(off))
        return dict(n=self.n, m=self.m, n_eq=self.n_eq, row_ptr=rp, col=col[: self.nnz],
                    val=val[: self.nnz], c=c[: self.n], q=q[: self.m], l=l[: self.n],
                    u=u[: self.n], offset=off.value)

    def spmv(self, x, transpose=False):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(max(self.n if transpose else self.m, 1))
        lib().orc_spmv(self.h, 1 if transpose else 0, _dp(x), _dp(out))
        return out[: (self.n if transpose else self.m)]

    def spectral_norm_bound(self) -> float:
        return lib().orc_spectral_norm_bound(self.h)

    def scaling(self, ruiz=10, pc=True, alpha=1.0):
        rs = np.zeros(max(self.m, 1))
        cs = np.zeros(max(self.n, 1))
        vals = np.zeros(max(self.nnz, 1))
        lib().orc_scaling(self.h, ruiz, 1 if pc else 0, alpha, _dp(rs), _dp(cs), _dp(vals))
        return rs[: self.m], cs[: self.n], vals[: self.nnz]

    def kkt(self, x, y, kx=None, kty=None):
        out = np.zeros(4)
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        kx = None if kx is None else np.ascontiguousarray(kx, np.float64)
        kty = None if kty is None else np.ascontiguousarray(kty, np.float64)
        lib().orc_kkt_evaluate(self.h, _dp(x), _dp(y), _dp(kx), _dp(kty), _dp(out))
        return out

    def primal_infeasibility(self, y):
        v = C.c_double()
        y = np.ascontiguousarray(y, np.float64)
        ok = lib().orc_primal_infeasibility(self.h, _dp(y), None, C.byref(v))
        return v.value if ok else None

    def dual_infeasibility(self, d):
        v = C.c_double()
        d = np.ascontiguousarray(d, np.float64)
        ok = lib().orc_dual_infeasibility(self.h, _dp(d), None, C.byref(v))
        return v.value if ok else None

    def classify(self, x, y, tol):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        s = lib().orc_classify(self.h, _dp(x), _dp(y), tol)
        return None if s < 0 else STATUS_NAMES[s]

    def solve(self, params: Params | None = None, precision: int = 64, warm_x=None, warm_y=None,
              warm_step=float("nan"), warm_weight=float("nan")) -> Solution:
        params = params or default_params()
        res = Result()
        x = np.zeros(max(self.n, 1))
        y = np.zeros(max(self.m, 1))
        rc = np.zeros(max(self.n, 1))
        wx = None if warm_x is None else np.ascontiguousarray(warm_x, np.float64)
        wy = None if warm_y is None else np.ascontiguousarray(warm_y, np.float64)
        r = lib().orc_pdhg_solve(self.h, C.byref(params), precision, _dp(wx), _dp(wy), warm_step,
                                 warm_weight, C.byref(res), _dp(x), _dp(y), _dp(rc))
        if r != 0:
            raise ValueError("oracle rejected the solve arguments")
        return _solution(res, x[: self.n], y[: self.m], rc[: self.n])

    def mixed_solve(self, params: Params | None = None, device_precision: int = 32,
                    mixed: MixedParams | None = None) -> Solution:
        params = params or default_params()
        mixed = mixed or default_mixed_params()
        res = Result()
        info = MixedInfo()
        x = np.zeros(max(self.n, 1))
        y = np.zeros(max(self.m, 1))
        rc = np.zeros(max(self.n, 1))
        r = lib().orc_mixed_solve(self.h, C.byref(params), device_precision, C.byref(mixed),
                                  C.byref(res), C.byref(info), _dp(x), _dp(y), _dp(rc))
        if r != 0:
            raise ValueError("oracle rejected the mixed-precision arguments")
        s = _solution(res, x[: self.n], y[: self.m], rc[: self.n])
        s.mixed = info
        return s


def _solution(res: Result, x, y, rc) -> Solution:
    return Solution(STATUS_NAMES[res.status], res.iterations, res.restarts, res.objective,
                    res.dual_objective, res.primal_residual, res.dual_residual, res.solve_time_ms,
                    res.final_step_size, res.final_primal_weight, x.copy(), y.copy(), rc.copy())


class JRandom:
    """java.util.Random restatement, for known-answer tests."""

    def __init__(self, seed: int):
        self._buf = C.create_string_buffer(32)
        lib().orc_jrandom_init(self._buf, seed)

    def next_int(self, bound):
        return lib().orc_jrandom_next_int(self._buf, bound)

    def next_double(self):
        return lib().orc_jrandom_next_double(self._buf)

    def next_gaussian(self):
        return lib().orc_jrandom_next_gaussian(self._buf)


def fdlibm_log(x: float) -> float:
    return lib().orc_fdlibm_log(x)


# ---------------------------------------------------------------------------
# Branch-and-bound (BranchAndBound.scala:168-469), sequential depth-first, with
# every node relaxation solved by the oracle's FP64 PDHG.

MILP_STATUS_NAMES = ["Optimal", "Feasible", "Infeasible", "Unbounded", "NoSolutionFound"]
_CONCLUSIVE = {"Optimal", "PrimalInfeasible", "DualInfeasible"}


@dataclass
class MilpSolution:
    status: str
    primal: np.ndarray
    objective: float
    best_bound: float
    nodes_explored: int
    unproven_nodes: int


def java_round(v: float) -> float:
    """Math.round(v).toDouble: nearest whole number, ties toward +inf, NaN -> 0."""
    if v != v:
        return 0.0
    if not np.isfinite(v) or abs(v) >= 2.0 ** 52:
        return v
    f = float(np.floor(v))
    return f + 1.0 if v - f >= 0.5 else f


def skips_before_solving(safe_bound: float, incumbent: float, have: bool) -> bool:
    """BranchAndBound.scala:407-413."""
    return have and safe_bound >= incumbent


def safe_bound_of(conclusive: bool, bound: float, slack: float) -> float:
    """BranchAndBound.scala:422-427."""
    if not conclusive:
        return float("-inf")
    c = bound - slack
    return c if np.isfinite(c) else float("-inf")


def most_fractional(x, integers, tol):
    """BranchAndBound.scala:441-453: the furthest from whole, first on ties."""
    best, worst = -1, tol
    for j in integers:
        g = abs(x[j] - java_round(x[j]))
        if g > worst:
            worst, best = g, j
    return best


def _node_lp(data, lower, upper) -> OracleLp:
    return OracleLp.from_csr(data["n"], data["m"], data["n_eq"], data["row_ptr"], data["col"],
                             data["val"], data["c"], data["q"], lower, upper, data["offset"])


def _primal_objective(data, x) -> float:
    """LpProblem.primalObjective (LpProblem.scala:42-50): offset first, then in order."""
    s = data["offset"]
    for ci, xi in zip(data["c"], x):
        s += ci * xi
    return s


def bnb_solve(lp: OracleLp, integers, params: Params | None = None,
              integrality_tolerance=1e-6, gap_abs=1e-6, gap_rel=1e-6,
              pruning_safety_factor=10.0, max_nodes=100_000, warm_start=True) -> MilpSolution:
    """BranchAndBound.solveWith (BranchAndBound.scala:168-392), no time limit."""
    params = params or default_params()
    data = lp.export()
    n = data["n"]
    for j in integers:
        if not (0 <= j < n):
            raise ValueError(f"integer column {j} is out of range [0, {n})")
    if not integers:  # :181-207
        root = lp.solve(params)
        st = {"Optimal": "Optimal", "PrimalInfeasible": "Infeasible",
              "DualInfeasible": "Unbounded"}.get(root.status, "NoSolutionFound")
        ok = st == "Optimal"
        return MilpSolution(st, root.primal if ok else np.zeros(0),
                            root.objective if ok else float("inf"),
                            root.objective if ok else float("-inf"), 1,
                            0 if root.status in _CONCLUSIVE else 1)
    ordered = sorted(set(integers))
    stack = [(data["l"].copy(), data["u"].copy(), float("-inf"), None)]
    incumbent, inc_val = None, float("inf")
    nodes = unproven = 0
    root_unbounded = False
    while stack and nodes < max_nodes:
        lower, upper, safe_parent, warm = stack.pop()
        if skips_before_solving(safe_parent, inc_val, incumbent is not None):
            continue
        sub = _node_lp(data, lower, upper)
        if warm_start and warm is not None:
            rel = sub.solve(params, 64, warm[0], warm[1], warm[2], warm[3])
        else:
            rel = sub.solve(params)
        nodes += 1
        if rel.status == "PrimalInfeasible":
            continue
        if rel.status == "DualInfeasible":
            if nodes == 1:
                root_unbounded = True
            else:
                unproven += 1
            continue
        conclusive = rel.status in _CONCLUSIVE
        if not conclusive:
            unproven += 1
        slack = pruning_safety_factor * max(abs(rel.objective - rel.dual_objective),
                                            params.eps_abs)
        safe = safe_bound_of(conclusive, rel.objective, slack)
        if conclusive and incumbent is not None and safe >= inc_val:
            continue
        j = most_fractional(rel.primal, ordered, integrality_tolerance)
        if j < 0:
            if conclusive:
                snapped = rel.primal.copy()
                for c in ordered:
                    snapped[c] = java_round(snapped[c])
                value = _primal_objective(data, snapped)
                if value < inc_val:
                    inc_val, incumbent = value, snapped
            continue
        child_warm = None
        if warm_start and np.all(np.isfinite(rel.primal)) and np.all(np.isfinite(rel.dual)):
            ok = lambda v: v if (np.isfinite(v) and v > 0) else float("nan")
            child_warm = (rel.primal.copy(), rel.dual.copy(), ok(rel.final_step_size),
                          ok(rel.final_primal_weight))
        v = rel.primal[j]
        fl, ce = float(np.floor(v)), float(np.ceil(v))
        if ce <= upper[j]:
            lo = lower.copy()
            lo[j] = max(lo[j], ce)
            stack.append((lo, upper, safe, child_warm))
        if fl >= lower[j]:
            up = upper.copy()
            up[j] = min(up[j], fl)
            stack.append((lower, up, safe, child_warm))
    open_bound = inc_val if not stack else min(inc_val, min(s[2] for s in stack))
    bound = inc_val if not stack else open_bound

    def closed(inc, b):
        a = abs(inc - b)
        return a <= gap_abs or a <= gap_rel * max(1.0, abs(inc))

    have = incumbent is not None
    if root_unbounded:
        st = "Unbounded"
    elif not have and not stack and unproven == 0:
        st = "Infeasible"
    elif not have:
        st = "NoSolutionFound"
    elif unproven > 0:
        st = "Feasible"
    elif not stack or closed(inc_val, bound):
        st = "Optimal"
    else:
        st = "Feasible"
    return MilpSolution(st, incumbent if have else np.zeros(0), inc_val,
                        bound if not have else min(bound, inc_val), nodes, unproven)
