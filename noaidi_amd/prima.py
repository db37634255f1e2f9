"""Host-side mirror of Prima's solver API over libprima_hip's C ABI.

Names and argument meaning follow the reference (paths relative to
modules/prima-core/src/main/scala/org/noaidi/prima/):

  HipKernels           trait Kernels (kernels/Kernels.scala:42-125)
  LpProblem            LpProblem (LpProblem.scala:22-104), standard form
  PdhgParams           PdhgParams (PdhgParams.scala:9-59) + ScalingParams
  MixedParams          MixedPrecision.Params (MixedPrecision.scala:37-76)
  LpSolution           LpSolution (SolveResult.scala:44-65)
  WarmStart            Pdhg.WarmStart (Pdhg.scala:72-105)
  solve / solve_with   Pdhg.solve / Pdhg.solveWith (Pdhg.scala:35-55)
  mixed_solve          MixedPrecision.solve (MixedPrecision.scala:104-199)
  solve_all            PrimaZio.solveAll (PrimaZio.scala:61-66): input order kept
  HipPdhgSolver        LpSolver "prima-pdhg" (Pdhg.scala:108-110)
  HipMixedSolver       LpSolver "prima-pdhg-mixed" (MixedPrecision.scala:204-214)

Argument errors raise ValueError (the reference's IllegalArgumentException
from `require`); numerical trouble comes back as a status, never raised.
The library must be built (``python -m noaidi_amd.build``): there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libprima_hip.so")
INSTANCES_PATH = os.path.join(_HERE, "_build", "libprima_instances.so")

STATUS = [
    "Optimal",
    "PrimalInfeasible",
    "DualInfeasible",
    "IterationLimit",
    "TimeLimit",
    "NumericalError",
    "Interrupted",
]
CONCLUSIVE = {"Optimal", "PrimalInfeasible", "DualInfeasible"}

# MilpStatus (BranchAndBound.scala:7-31), same ordinals as the C ABI.
MILP_STATUS = ["Optimal", "Feasible", "Infeasible", "Unbounded", "NoSolutionFound"]
MILP_CONCLUSIVE = {"Optimal", "Infeasible", "Unbounded"}

PRIMA_OK, PRIMA_EINVAL, PRIMA_ENOMEM, PRIMA_EDEVICE, PRIMA_ECLOSED = 0, -1, -2, -3, -4


class DeviceError(RuntimeError):
    pass


# ---------------------------------------------------------------------------
# C structures (include/prima_hip.h)


class _Caps(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("device", C.c_char * 64),
                ("supports_float64", C.c_int32)]


class _Lp(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("n_eq", C.c_int32), ("pad", C.c_int32),
                ("nnz", C.c_int64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
                ("val", C.c_void_p), ("c", C.c_void_p), ("q", C.c_void_p), ("l", C.c_void_p),
                ("u", C.c_void_p), ("offset", C.c_double)]


class _Params(C.Structure):
    _fields_ = [("eps_abs", C.c_double), ("eps_rel", C.c_double),
                ("max_iterations", C.c_int32), ("evaluation_frequency", C.c_int32),
                ("time_limit_ms", C.c_int64), ("ruiz_iterations", C.c_int32),
                ("apply_pock_chambolle", C.c_int32), ("pock_chambolle_alpha", C.c_double),
                ("infeasibility_tolerance", C.c_double), ("restart_sufficient", C.c_double),
                ("restart_necessary", C.c_double), ("restart_artificial", C.c_double),
                ("primal_weight_smoothing", C.c_double), ("workgroups_per_lp", C.c_int32),
                ("chunk_iterations", C.c_int32)]


class _Warm(C.Structure):
    _fields_ = [("primal", C.c_void_p), ("dual", C.c_void_p), ("step_size", C.c_double),
                ("primal_weight", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("restarts", C.c_int32),
                ("used_average", C.c_int32), ("objective", C.c_double),
                ("dual_objective", C.c_double), ("primal_residual", C.c_double),
                ("dual_residual", C.c_double), ("solve_time_ms", C.c_double),
                ("total_time_ms", C.c_double), ("final_step_size", C.c_double),
                ("final_primal_weight", C.c_double), ("trials", C.c_int64),
                ("evaluations", C.c_int32), ("workgroups", C.c_int32)]


class _Mixed(C.Structure):
    _fields_ = [("reduced_tolerance", C.c_double), ("reduced_max_iterations", C.c_int32),
                ("always_refine", C.c_int32), ("reduced_time_share", C.c_double),
                ("refinement", C.c_int32), ("max_refinements", C.c_int32),
                ("inner_tolerance", C.c_double), ("inner_max_iterations", C.c_int32),
                ("stall_iterations", C.c_int32), ("fix_threshold", C.c_double)]


class _MixedInfo(C.Structure):
    _fields_ = [("reduced_iterations", C.c_int32), ("reduced_restarts", C.c_int32),
                ("refinement_iterations", C.c_int32), ("refined", C.c_int32),
                ("reduced_ms", C.c_double), ("refinement_ms", C.c_double),
                ("rounds", C.c_int32), ("fallback_iterations", C.c_int32)]


TUNE_MAX = 9


class _TuneResult(C.Structure):
    _fields_ = [("workgroups", C.c_int32), ("blocks_per_workgroup", C.c_int32),
                ("us_per_iteration", C.c_double), ("candidates", C.c_int32),
                ("cand_workgroups", C.c_int32 * TUNE_MAX),
                ("cand_blocks_per_workgroup", C.c_int32 * TUNE_MAX),
                ("cand_us_per_iteration", C.c_double * TUNE_MAX)]


class _Stats(C.Structure):
    _fields_ = [("launches", C.c_int64), ("kernel_ms", C.c_double),
                ("accepted_iterations", C.c_int64), ("trials", C.c_int64),
                ("evaluations", C.c_int64), ("algorithmic_bytes", C.c_double),
                ("launches_by_precision", C.c_int64 * 2),
                ("kernel_ms_by_precision", C.c_double * 2),
                ("algorithmic_bytes_by_precision", C.c_double * 2)]


class _BnbParams(C.Structure):
    _fields_ = [("lp", _Params), ("integrality_tolerance", C.c_double),
                ("gap_abs", C.c_double), ("gap_rel", C.c_double),
                ("pruning_safety_factor", C.c_double), ("max_nodes", C.c_int32),
                ("warm_start", C.c_int32), ("time_limit_ms", C.c_int64),
                ("frontier", C.c_int32), ("pad", C.c_int32)]


class _MilpResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("nodes_explored", C.c_int32),
                ("unproven_nodes", C.c_int32), ("has_incumbent", C.c_int32),
                ("relaxations", C.c_int32), ("batches", C.c_int32),
                ("objective", C.c_double), ("best_bound", C.c_double),
                ("solve_time_ms", C.c_double)]


class _Instance(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("n_eq", C.c_int32), ("pad", C.c_int32),
                ("nnz", C.c_int64), ("row_ptr", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double)),
                ("c", C.POINTER(C.c_double)), ("q", C.POINTER(C.c_double)),
                ("l", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)),
                ("offset", C.c_double), ("known_objective", C.c_double),
                ("x_star", C.POINTER(C.c_double)), ("y_star", C.POINTER(C.c_double)),
                ("n_int", C.c_int32), ("pad2", C.c_int32),
                ("integers", C.POINTER(C.c_int32))]


# Every symbol include/prima_hip.h declares.
EXPORTED = [
    "prima_open", "prima_close", "prima_destroy", "prima_capabilities_of", "prima_last_error",
    "prima_vec_alloc", "prima_vec_upload", "prima_vec_download", "prima_vec_length",
    "prima_vec_free", "prima_mat_upload_csr", "prima_mat_free", "prima_spmv", "prima_axpby",
    "prima_scale", "prima_copy", "prima_dot", "prima_squared_norm", "prima_primal_step",
    "prima_dual_step", "prima_default_params", "prima_default_mixed_params", "prima_solve",
    "prima_mixed_solve", "prima_solve_batch", "prima_cancel", "prima_stats_get",
    "prima_stats_reset", "prima_mat_set_schedule", "prima_mat_autotune",
    "prima_default_bnb_params", "prima_milp_solve", "prima_bnb_skips_before_solving",
    "prima_bnb_safe_bound_of", "prima_progress", "prima_test_stall_barriers",
    "prima_solve_noalloc", "prima_tune_solver",
]

_lib = None
_ilib = None


def lib():
    """The loaded libprima_hip; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libprima_hip is not built ({LIB_PATH}); run `python -m noaidi_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        pp = C.POINTER(C.c_void_p)
        sig = {
            "prima_open": (i32, [i32, i32, pp]),
            "prima_close": (i32, [vp]),
            "prima_destroy": (None, [vp]),
            "prima_capabilities_of": (i32, [vp, C.POINTER(_Caps)]),
            "prima_last_error": (C.c_char_p, [vp]),
            "prima_vec_alloc": (i32, [vp, i64, pp]),
            "prima_vec_upload": (i32, [vp, vp, i64, pp]),
            "prima_vec_download": (i32, [vp, vp, vp, i64]),
            "prima_vec_length": (i64, [vp]),
            "prima_vec_free": (i32, [vp, vp]),
            "prima_mat_upload_csr": (i32, [vp, i32, i32, i64, vp, vp, vp, pp]),
            "prima_mat_free": (i32, [vp, vp]),
            "prima_spmv": (i32, [vp, vp, vp, vp]),
            "prima_axpby": (i32, [vp, d, vp, d, vp, vp]),
            "prima_scale": (i32, [vp, d, vp, vp]),
            "prima_copy": (i32, [vp, vp, vp]),
            "prima_dot": (i32, [vp, vp, vp, C.POINTER(d)]),
            "prima_squared_norm": (i32, [vp, vp, C.POINTER(d)]),
            "prima_primal_step": (i32, [vp, vp, vp, vp, vp, vp, d, vp]),
            "prima_dual_step": (i32, [vp, vp, vp, vp, d, i32, vp]),
            "prima_default_params": (None, [C.POINTER(_Params)]),
            "prima_default_mixed_params": (None, [C.POINTER(_Mixed)]),
            "prima_solve": (i32, [vp, C.POINTER(_Lp), C.POINTER(_Params), C.POINTER(_Warm),
                                  C.POINTER(_Result), vp, vp, vp]),
            "prima_mixed_solve": (i32, [vp, C.POINTER(_Lp), C.POINTER(_Params),
                                        C.POINTER(_Mixed), C.POINTER(_Result),
                                        C.POINTER(_MixedInfo), vp, vp, vp]),
            "prima_solve_batch": (i32, [vp, i32, C.POINTER(_Lp), C.POINTER(_Params),
                                        C.POINTER(_Mixed), C.POINTER(_Warm), C.POINTER(_Result),
                                        C.POINTER(_MixedInfo), vp, vp, vp]),
            "prima_cancel": (i32, [vp]),
            "prima_stats_get": (i32, [vp, C.POINTER(_Stats)]),
            "prima_stats_reset": (i32, [vp]),
            "prima_mat_set_schedule": (i32, [vp, vp, i32]),
            "prima_mat_autotune": (i32, [vp, vp, i32, C.POINTER(i32), C.POINTER(d)]),
            "prima_default_bnb_params": (None, [C.POINTER(_BnbParams)]),
            "prima_milp_solve": (i32, [vp, C.POINTER(_Lp), i32, vp, C.POINTER(_BnbParams),
                                       C.POINTER(_Mixed), C.POINTER(_MilpResult), vp]),
            "prima_bnb_skips_before_solving": (i32, [d, d, i32]),
            "prima_bnb_safe_bound_of": (d, [i32, d, d]),
            "prima_progress": (i32, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
            "prima_test_stall_barriers": (i32, [vp, i32]),
            "prima_solve_noalloc": (i32, [vp, C.POINTER(_Lp), C.POINTER(_Params),
                                          C.POINTER(_Warm), C.POINTER(_Result), vp, vp, vp]),
            "prima_tune_solver": (i32, [vp, C.POINTER(_Lp), C.POINTER(_Params), i32,
                                        C.POINTER(_TuneResult)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def instances_lib():
    global _ilib
    if _ilib is None:
        if not os.path.exists(INSTANCES_PATH):
            raise RuntimeError(f"instance generators not built ({INSTANCES_PATH})")
        L = C.CDLL(INSTANCES_PATH)
        L.prima_gen_random_known.restype = C.POINTER(_Instance)
        L.prima_gen_random_known.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32,
                                             C.c_double, C.c_double, C.c_int32, C.c_double]
        L.prima_gen_dispatch.restype = C.POINTER(_Instance)
        L.prima_gen_dispatch.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_double, C.c_double, C.c_uint64]
        L.prima_gen_uc.restype = C.POINTER(_Instance)
        L.prima_gen_uc.argtypes = [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_int32, C.c_int32]
        L.prima_instance_free.argtypes = [C.POINTER(_Instance)]
        _ilib = L
    return _ilib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------------------
# problem, parameters, results


@dataclass
class LpProblem:
    """min c'x + offset  s.t.  K x = q (first n_eq rows), K x >= q, l <= x <= u."""

    n: int
    m: int
    n_eq: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    c: np.ndarray
    q: np.ndarray
    l: np.ndarray
    u: np.ndarray
    offset: float = 0.0
    known_objective: float = float("nan")
    x_star: Optional[np.ndarray] = None
    y_star: Optional[np.ndarray] = None
    integers: Optional[np.ndarray] = None  # integer columns of a generated MILP

    def __post_init__(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(self.col, dtype=np.int32)
        self.val = np.ascontiguousarray(self.val, dtype=np.float64)
        for k in ("c", "q", "l", "u"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float64))

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[self.m]) if self.m else 0

    @classmethod
    def from_dict(cls, d: dict) -> "LpProblem":
        keys = ("n", "m", "n_eq", "row_ptr", "col", "val", "c", "q", "l", "u", "offset")
        return cls(**{k: d[k] for k in keys})

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=(self.m, self.n))

    def _c(self) -> _Lp:
        s = _Lp()
        s.n, s.m, s.n_eq, s.nnz = self.n, self.m, self.n_eq, self.nnz
        s.row_ptr, s.col, s.val = _ptr(self.row_ptr), _ptr(self.col), _ptr(self.val)
        s.c, s.q, s.l, s.u = _ptr(self.c), _ptr(self.q), _ptr(self.l), _ptr(self.u)
        s.offset = self.offset
        return s


@dataclass
class PdhgParams:
    eps_abs: float = 1e-8
    eps_rel: float = 1e-8
    max_iterations: int = 100_000
    evaluation_frequency: int = 64
    time_limit_ms: Optional[int] = None
    ruiz_iterations: int = 10
    apply_pock_chambolle: bool = True
    pock_chambolle_alpha: float = 1.0
    infeasibility_tolerance: float = 1e-8
    restart_sufficient: float = 0.2
    restart_necessary: float = 0.8
    restart_artificial: float = 0.36
    primal_weight_smoothing: float = 0.5
    workgroups_per_lp: int = 0
    chunk_iterations: int = 0

    def _c(self) -> _Params:
        p = _Params()
        for f in _Params._fields_:
            name = f[0]
            if name == "time_limit_ms":
                p.time_limit_ms = -1 if self.time_limit_ms is None else int(self.time_limit_ms)
            elif name == "apply_pock_chambolle":
                p.apply_pock_chambolle = 1 if self.apply_pock_chambolle else 0
            else:
                setattr(p, name, getattr(self, name))
        return p


LOOSE = PdhgParams(eps_abs=1e-6, eps_rel=1e-6)

@dataclass
class RowTranslation:
    """How each original two-sided row maps to standard-form rows
    (LpProblem.scala:110-147): ("direct", r), ("negated", r) or
    ("range", lower_row, upper_row)."""

    expansions: list

    @property
    def num_original_rows(self) -> int:
        return len(self.expansions)

    def original_duals(self, standard_duals) -> np.ndarray:
        out = np.zeros(len(self.expansions))
        for i, e in enumerate(self.expansions):
            if e[0] == "direct":
                out[i] = standard_duals[e[1]]
            elif e[0] == "negated":
                out[i] = -standard_duals[e[1]]
            else:
                out[i] = standard_duals[e[1]] - standard_duals[e[2]]
        return out


class LpBuilder:
    """LpBuilder (LpProblem.scala:154-248): rows `lo <= a'x <= hi`, converted
    to equalities first, then `>=` rows; `<=` rows are negated and ranges split
    into two rows."""

    def __init__(self, num_variables: int):
        if num_variables < 0:
            raise ValueError(f"numVariables must be non-negative, got {num_variables}")
        self.n = num_variables
        self.c = np.zeros(num_variables)
        self.l = np.full(num_variables, -np.inf)
        self.u = np.full(num_variables, np.inf)
        self.offset = 0.0
        self.rows: list = []

    def _check(self, v):
        if not (0 <= v < self.n):
            raise ValueError(f"variable index {v} out of range [0, {self.n})")

    def objectiveCoefficient(self, v: int, value: float) -> "LpBuilder":
        self._check(v)
        self.c[v] = value
        return self

    def objectiveOffset(self, value: float) -> "LpBuilder":
        self.offset = value
        return self

    def bounds(self, v: int, lo: float, hi: float) -> "LpBuilder":
        self._check(v)
        if not lo <= hi:
            raise ValueError(f"variable {v} has empty bound interval [{lo}, {hi}]")
        self.l[v], self.u[v] = lo, hi
        return self

    def constraint(self, coefficients, lo: float, hi: float) -> "LpBuilder":
        if not lo <= hi:
            raise ValueError(f"constraint has empty range [{lo}, {hi}]")
        if lo == -np.inf and hi == np.inf:
            raise ValueError("constraint is unbounded on both sides and would have no effect")
        coefficients = list(coefficients)
        for v, _ in coefficients:
            self._check(v)
        self.rows.append((coefficients, lo, hi))
        return self

    def equalityConstraint(self, coefficients, value: float) -> "LpBuilder":
        return self.constraint(coefficients, value, value)

    def greaterThan(self, coefficients, lo: float) -> "LpBuilder":
        return self.constraint(coefficients, lo, np.inf)

    def lessThan(self, coefficients, hi: float) -> "LpBuilder":
        return self.constraint(coefficients, -np.inf, hi)

    def build(self):
        trips_r, trips_c, trips_v, rhs = [], [], [], []
        exp: list = [None] * len(self.rows)

        def emit(coefs, sign, value):
            r = len(rhs)
            for v, a in coefs:
                trips_r.append(r)
                trips_c.append(v)
                trips_v.append(sign * a)
            rhs.append(sign * value)
            return r

        for i, (co, lo, hi) in enumerate(self.rows):
            if lo == hi:
                exp[i] = ("direct", emit(co, 1.0, lo))
        n_eq = len(rhs)
        for i, (co, lo, hi) in enumerate(self.rows):
            if lo == hi:
                continue
            has_lo, has_hi = lo != -np.inf, hi != np.inf
            if has_lo and has_hi:
                exp[i] = ("range", emit(co, 1.0, lo), emit(co, -1.0, hi))
            elif has_lo:
                exp[i] = ("direct", emit(co, 1.0, lo))
            else:
                exp[i] = ("negated", emit(co, -1.0, hi))
        m = len(rhs)
        row_ptr, col, val = csr_from_triplets(m, self.n, trips_r, trips_c, trips_v)
        p = LpProblem(self.n, m, n_eq, row_ptr, col, val, self.c.copy(), np.array(rhs, float),
                      self.l.copy(), self.u.copy(), self.offset)
        return p, RowTranslation(exp)


def csr_from_triplets(rows: int, cols: int, r, c, v):
    """SparseMatrix.fromTriplets (SparseMatrix.scala:204-260): stable sort by
    (row, col), duplicates summed in input order, exact zeros dropped,
    non-finite values rejected."""
    r = np.asarray(r, dtype=np.int64)
    c = np.asarray(c, dtype=np.int64)
    v = np.asarray(v, dtype=np.float64)
    if len(r) and (r.min() < 0 or r.max() >= rows or c.min() < 0 or c.max() >= cols):
        raise ValueError("triplet index out of range")
    if not np.all(np.isfinite(v)):
        raise ValueError("matrix entry is not finite")
    order = np.lexsort((c, r))  # stable
    r, c, v = r[order], c[order], v[order]
    keep_r, keep_c, keep_v = [], [], []
    i = 0
    while i < len(r):
        j = i + 1
        acc = v[i]
        while j < len(r) and r[j] == r[i] and c[j] == c[i]:
            acc += v[j]
            j += 1
        if acc != 0.0:
            keep_r.append(r[i])
            keep_c.append(c[i])
            keep_v.append(acc)
        i = j
    row_ptr = np.zeros(rows + 1, dtype=np.int32)
    np.add.at(row_ptr, np.asarray(keep_r, dtype=np.int64) + 1, 1)
    row_ptr = np.cumsum(row_ptr, dtype=np.int64).astype(np.int32)
    return row_ptr, np.asarray(keep_c, dtype=np.int32), np.asarray(keep_v, dtype=np.float64)


def builder(num_variables: int) -> LpBuilder:
    """LpProblem.builder (LpProblem.scala:105)."""
    return LpBuilder(num_variables)



@dataclass
class MixedParams:
    """MixedPrecision.Params; `refinement` selects the finishing pass:
    "warm" (the reference's float64 warm-started pass) or "iterative" (FP64
    outer iterative refinement around float32 correction solves)."""

    reduced_tolerance: float = 1e-5
    reduced_max_iterations: int = 20_000
    always_refine: bool = False
    reduced_time_share: float = 0.5
    refinement: str = "warm"
    max_refinements: int = 4
    inner_tolerance: float = 1e-4
    inner_max_iterations: int = 20_000
    stall_iterations: int = 0  # 0: the reference contract (no cold fallback)
    fix_threshold: float = 0.0  # iterative refinement: fix strongly active bounds (0: off)

    def _c(self) -> _Mixed:
        m = _Mixed()
        m.reduced_tolerance = self.reduced_tolerance
        m.reduced_max_iterations = self.reduced_max_iterations
        m.always_refine = 1 if self.always_refine else 0
        m.reduced_time_share = self.reduced_time_share
        if self.refinement not in ("warm", "iterative"):
            raise ValueError(f"refinement must be 'warm' or 'iterative', got {self.refinement!r}")
        m.refinement = 1 if self.refinement == "iterative" else 0
        m.max_refinements = self.max_refinements
        m.inner_tolerance = self.inner_tolerance
        m.inner_max_iterations = self.inner_max_iterations
        m.stall_iterations = self.stall_iterations
        m.fix_threshold = self.fix_threshold
        return m


@dataclass
class WarmStart:
    primal: np.ndarray
    dual: np.ndarray
    step_size: Optional[float] = None
    primal_weight: Optional[float] = None

    @classmethod
    def of(cls, sol: "LpSolution") -> "WarmStart":
        """Pdhg.WarmStart.apply(solution) (Pdhg.scala:90-105)."""
        if not np.all(np.isfinite(sol.primal)):
            raise ValueError("warm start rejected: the source solution's primal vector is not finite")
        if not np.all(np.isfinite(sol.dual)):
            raise ValueError("warm start rejected: the source solution has no usable dual vector")
        ok = lambda v: v if (v is not None and math.isfinite(v) and v > 0) else None
        return cls(sol.primal.copy(), sol.dual.copy(), ok(sol.final_step_size),
                   ok(sol.final_primal_weight))

    def _c(self, keep: list) -> _Warm:
        w = _Warm()
        p = np.ascontiguousarray(self.primal, dtype=np.float64)
        d = np.ascontiguousarray(self.dual, dtype=np.float64)
        keep += [p, d]
        w.primal, w.dual = _ptr(p), _ptr(d)
        w.step_size = self.step_size if self.step_size is not None else float("nan")
        w.primal_weight = self.primal_weight if self.primal_weight is not None else float("nan")
        return w


@dataclass
class LpSolution:
    status: str
    primal: np.ndarray
    dual: np.ndarray
    reduced_costs: np.ndarray
    objective_value: float
    dual_objective_value: float
    iterations: int
    restarts: int
    solve_time_ms: float
    primal_residual: float
    dual_residual: float
    final_step_size: float
    final_primal_weight: float
    total_time_ms: float = 0.0
    trials: int = 0
    evaluations: int = 0
    workgroups: int = 0
    used_average: bool = False
    mixed: Optional[dict] = None

    @property
    def duality_gap(self) -> float:
        return abs(self.objective_value - self.dual_objective_value)


def _solution(r: _Result, x, y, rc, info: Optional[_MixedInfo] = None) -> LpSolution:
    s = LpSolution(STATUS[r.status], x, y, rc, r.objective, r.dual_objective, r.iterations,
                   r.restarts, r.solve_time_ms, r.primal_residual, r.dual_residual,
                   r.final_step_size, r.final_primal_weight, r.total_time_ms, r.trials,
                   r.evaluations, r.workgroups, bool(r.used_average))
    if info is not None:
        s.mixed = dict(reduced_iterations=info.reduced_iterations,
                       reduced_restarts=info.reduced_restarts,
                       refinement_iterations=info.refinement_iterations,
                       refined=bool(info.refined), reduced_ms=info.reduced_ms,
                       refinement_ms=info.refinement_ms, rounds=info.rounds,
                       fallback_iterations=info.fallback_iterations)
    return s


@dataclass
class BnbParams:
    """BnbParams (BranchAndBound.scala:57-108). `frontier` open nodes are
    popped together and their relaxations solved in one batched device call;
    frontier = 1 is the reference's sequential depth-first search."""

    lp: PdhgParams = field(default_factory=PdhgParams)
    integrality_tolerance: float = 1e-6
    gap_abs: float = 1e-6
    gap_rel: float = 1e-6
    pruning_safety_factor: float = 10.0
    max_nodes: int = 100_000
    time_limit_ms: Optional[int] = None
    warm_start: bool = True
    frontier: int = 1

    def copy(self, **kw) -> "BnbParams":
        return replace(self, **kw)

    def _c(self) -> _BnbParams:
        b = _BnbParams()
        b.lp = self.lp._c()
        b.integrality_tolerance = self.integrality_tolerance
        b.gap_abs, b.gap_rel = self.gap_abs, self.gap_rel
        b.pruning_safety_factor = self.pruning_safety_factor
        b.max_nodes = self.max_nodes
        b.warm_start = 1 if self.warm_start else 0
        b.time_limit_ms = -1 if self.time_limit_ms is None else int(self.time_limit_ms)
        b.frontier = self.frontier
        return b


@dataclass
class MilpSolution:
    """MilpSolution (BranchAndBound.scala:33-55); `primal` is empty when no
    integer solution was found."""

    status: str
    primal: np.ndarray
    objective_value: float
    best_bound: float
    nodes_explored: int
    unproven_nodes: int
    solve_time_ms: float
    relaxations: int = 0
    batches: int = 0

    @property
    def gap(self) -> float:
        return abs(self.objective_value - self.best_bound)

    @property
    def is_conclusive(self) -> bool:
        return self.status in MILP_CONCLUSIVE


def skips_before_solving(safe_bound: float, incumbent: float, have_incumbent: bool) -> bool:
    """BranchAndBound.skipsBeforeSolving, evaluated by the library."""
    return bool(lib().prima_bnb_skips_before_solving(safe_bound, incumbent,
                                                     1 if have_incumbent else 0))


def safe_bound_of(conclusive: bool, bound: float, slack: float) -> float:
    """BranchAndBound.safeBoundOf, evaluated by the library."""
    return lib().prima_bnb_safe_bound_of(1 if conclusive else 0, bound, slack)


# ---------------------------------------------------------------------------
# the Kernels seam


class HipKernels:
    """trait Kernels on the GPU (kernels/Kernels.scala:42-125).

    precision=64 behaves like ScalaKernels, precision=32 like Float32Kernels
    (supportsFloat64 = false, so MixedPrecision runs its two passes).
    """

    def __init__(self, device: int = 0, precision: int = 64):
        h = C.c_void_p()
        rc = lib().prima_open(device, precision, C.byref(h))
        if rc != PRIMA_OK:
            raise DeviceError(f"prima_open(device={device}, precision={precision}) failed: {rc}")
        self._h = h
        self.precision = precision
        self._closed = False

    # lifecycle --------------------------------------------------------------
    def close(self) -> None:
        if not self._closed:
            lib().prima_close(self._h)
            self._closed = True

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().prima_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc == PRIMA_OK:
            return
        msg = lib().prima_last_error(self._h).decode()
        if rc == PRIMA_EINVAL:
            raise ValueError(f"{what}: {msg}")
        raise DeviceError(f"{what} failed ({rc}): {msg}")

    @property
    def capabilities(self) -> dict:
        c = _Caps()
        self._check(lib().prima_capabilities_of(self._h, C.byref(c)), "capabilities")
        return dict(name=c.name.decode(), device=c.device.decode(),
                    supportsFloat64=bool(c.supports_float64),
                    requiresDoublePrecisionRefinement=not bool(c.supports_float64))

    # memory -----------------------------------------------------------------
    def allocate(self, n: int) -> "DeviceVec":
        h = C.c_void_p()
        self._check(lib().prima_vec_alloc(self._h, n, C.byref(h)), "allocate")
        return DeviceVec(self, h, n)

    def upload(self, data) -> "DeviceVec":
        a = np.ascontiguousarray(data, dtype=np.float64)
        h = C.c_void_p()
        self._check(lib().prima_vec_upload(self._h, _ptr(a), len(a), C.byref(h)), "upload")
        return DeviceVec(self, h, len(a))

    def download(self, v: "DeviceVec", into: Optional[np.ndarray] = None) -> np.ndarray:
        out = np.zeros(v.n) if into is None else into
        if out.dtype != np.float64 or not out.flags.c_contiguous or len(out) != v.n:
            raise ValueError("download target must be a contiguous float64 array of the vector's length")
        self._check(lib().prima_vec_download(self._h, v._h, _ptr(out), v.n), "download")
        return out

    def uploadMatrix(self, rows: int, cols: int, row_ptr, col, val) -> "DeviceMat":
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col, dtype=np.int32)
        va = np.ascontiguousarray(val, dtype=np.float64)
        h = C.c_void_p()
        self._check(lib().prima_mat_upload_csr(self._h, rows, cols, len(va), _ptr(rp), _ptr(ci),
                                               _ptr(va), C.byref(h)), "uploadMatrix")
        return DeviceMat(self, h, rows, cols)

    def set_schedule(self, a: "DeviceMat", schedule: int) -> None:
        self._check(lib().prima_mat_set_schedule(self._h, a._h, schedule), "set_schedule")

    def autotune(self, a: "DeviceMat", reps: int = 5):
        """Time every SpMV schedule on `a`; keeps and returns (schedule, ms)."""
        ch, ms = C.c_int32(), C.c_double()
        self._check(lib().prima_mat_autotune(self._h, a._h, reps, C.byref(ch), C.byref(ms)),
                    "autotune")
        return ch.value, ms.value

    def length(self, v: "DeviceVec") -> int:
        return int(lib().prima_vec_length(v._h))

    # operations -------------------------------------------------------------
    def spmv(self, a: "DeviceMat", x: "DeviceVec", out: "DeviceVec") -> None:
        self._check(lib().prima_spmv(self._h, a._h, x._h, out._h), "spmv")

    def axpby(self, alpha, x, beta, y, out) -> None:
        self._check(lib().prima_axpby(self._h, alpha, x._h, beta, y._h, out._h), "axpby")

    def scale(self, alpha, x, out) -> None:
        self._check(lib().prima_scale(self._h, alpha, x._h, out._h), "scale")

    def copy(self, src, dst) -> None:
        self._check(lib().prima_copy(self._h, src._h, dst._h), "copy")

    def dot(self, x, y) -> float:
        r = C.c_double()
        self._check(lib().prima_dot(self._h, x._h, y._h, C.byref(r)), "dot")
        return r.value

    def squaredNorm(self, x) -> float:
        r = C.c_double()
        self._check(lib().prima_squared_norm(self._h, x._h, C.byref(r)), "squaredNorm")
        return r.value

    def primalStep(self, x, kty, cost, lower, upper, tau, out) -> None:
        self._check(lib().prima_primal_step(self._h, x._h, kty._h, cost._h, lower._h, upper._h,
                                            tau, out._h), "primalStep")

    def dualStep(self, y, kx_bar, rhs, sigma, num_equalities, out) -> None:
        self._check(lib().prima_dual_step(self._h, y._h, kx_bar._h, rhs._h, sigma,
                                          num_equalities, out._h), "dualStep")

    def cancel(self) -> None:
        lib().prima_cancel(self._h)

    def _test_stall_barriers(self, enable: bool) -> None:
        """Test hook (no reference counterpart): while enabled, solves whose
        LPs span several workgroups fail with DeviceError, as when a group
        barrier of the solver kernel gives up."""
        self._check(lib().prima_test_stall_barriers(self._h, 1 if enable else 0),
                    "test_stall_barriers")

    def stats(self) -> dict:
        st = _Stats()
        self._check(lib().prima_stats_get(self._h, C.byref(st)), "stats")
        out = {}
        for f in _Stats._fields_:
            v = getattr(st, f[0])
            out[f[0]] = list(v) if f[0].endswith("_by_precision") else v
        return out

    def progress(self) -> dict:
        """Launches, accepted iterations and finished LPs so far; may be called
        from another thread while a solve runs on this context."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(lib().prima_progress(self._h, C.byref(a), C.byref(b), C.byref(c)), "progress")
        return dict(launches=a.value, accepted_iterations=b.value, finished_lps=c.value)

    def reset_stats(self) -> None:
        self._check(lib().prima_stats_reset(self._h), "reset_stats")

    # coarse solve -------------------------------------------------------------
    def solve(self, problem: LpProblem, params: PdhgParams = PdhgParams(),
              warm: Optional[WarmStart] = None) -> LpSolution:
        """Pdhg.solveWith(problem, params, this, warmStart) on the device."""
        return self.solve_all([problem], params, warms=[warm])[0]

    def solve_noalloc(self, problem: LpProblem, params: PdhgParams = PdhgParams(),
                      warm: Optional[WarmStart] = None) -> LpSolution:
        """The solve of `solve`, with no device allocation (prima_solve_noalloc):
        the LP's device memory comes from the library's per-device workspace."""
        keep: list = []
        lp = problem._c()
        w = warm._c(keep) if warm is not None else None
        res = _Result()
        x, y, rc = np.zeros(max(problem.n, 1)), np.zeros(max(problem.m, 1)), np.zeros(max(problem.n, 1))
        cp = params._c()
        self._check(lib().prima_solve_noalloc(self._h, C.byref(lp), C.byref(cp),
                                              C.byref(w) if w is not None else None,
                                              C.byref(res), _ptr(x), _ptr(y), _ptr(rc)),
                    "solve_noalloc")
        return _solution(res, x[: problem.n], y[: problem.m], rc[: problem.n])

    def tune_solver(self, problem: LpProblem, params: PdhgParams = PdhgParams(),
                    iterations: int = 1024) -> dict:
        """prima_tune_solver: time `iterations` iterations of the LP at each
        candidate group size and row-block cut, keep the fastest for later
        solves of this matrix on this context, and return every timing."""
        lp = problem._c()
        cp = params._c()
        out = _TuneResult()
        self._check(lib().prima_tune_solver(self._h, C.byref(lp), C.byref(cp), int(iterations),
                                            C.byref(out)), "tune_solver")
        n = out.candidates
        return {"workgroups": out.workgroups,
                "blocks_per_workgroup": out.blocks_per_workgroup,
                "us_per_iteration": out.us_per_iteration,
                "candidates": [(out.cand_workgroups[i], out.cand_blocks_per_workgroup[i],
                                out.cand_us_per_iteration[i]) for i in range(n)]}

    def mixed_solve(self, problem: LpProblem, params: PdhgParams = PdhgParams(),
                    mixed: MixedParams = MixedParams()) -> LpSolution:
        """MixedPrecision.solve(problem, params, this, mixed)."""
        return self.solve_all([problem], params, mixed=mixed)[0]

    def solve_all(self, problems: Sequence[LpProblem], params: PdhgParams = PdhgParams(),
                  mixed: Optional[MixedParams] = None,
                  warms: Optional[Sequence[Optional[WarmStart]]] = None) -> list:
        k = len(problems)
        keep: list = []
        lps = (_Lp * max(k, 1))()
        for i, p in enumerate(problems):
            lps[i] = p._c()
        warr = None
        if warms is not None and any(w is not None for w in warms):
            warr = (_Warm * max(k, 1))()
            for i, w in enumerate(warms):
                warr[i] = w._c(keep) if w is not None else _Warm(None, None, float("nan"),
                                                                  float("nan"))
        res = (_Result * max(k, 1))()
        infos = (_MixedInfo * max(k, 1))()
        xs = [np.zeros(p.n) for p in problems]
        ys = [np.zeros(p.m) for p in problems]
        rcs = [np.zeros(p.n) for p in problems]
        px = (C.c_void_p * max(k, 1))(*[_ptr(a) if len(a) else None for a in xs])
        py = (C.c_void_p * max(k, 1))(*[_ptr(a) if len(a) else None for a in ys])
        prc = (C.c_void_p * max(k, 1))(*[_ptr(a) if len(a) else None for a in rcs])
        cp = params._c()
        cm = mixed._c() if mixed is not None else None
        rc = lib().prima_solve_batch(self._h, k, lps, C.byref(cp),
                                     C.byref(cm) if cm is not None else None, warr, res,
                                     infos if mixed is not None else None, px, py, prc)
        self._check(rc, "solve")
        return [_solution(res[i], xs[i], ys[i], rcs[i], infos[i] if mixed is not None else None)
                for i in range(k)]


    def milp_solve(self, problem: LpProblem, integers, params: BnbParams = BnbParams(),
                   mixed: Optional[MixedParams] = None) -> MilpSolution:
        """BranchAndBound.solveWith(problem, integers, params, this): node
        relaxations on the device, `params.frontier` of them per batched call."""
        ints = np.ascontiguousarray(sorted(set(int(j) for j in integers)), dtype=np.int32)
        lp = problem._c()
        cp = params._c()
        cm = mixed._c() if mixed is not None else None
        res = _MilpResult()
        x = np.zeros(max(problem.n, 1))
        rc = lib().prima_milp_solve(self._h, C.byref(lp), len(ints), _ptr(ints) if len(ints) else None,
                                    C.byref(cp), C.byref(cm) if cm is not None else None,
                                    C.byref(res), _ptr(x))
        self._check(rc, "milp_solve")
        primal = x[: problem.n].copy() if res.has_incumbent else np.zeros(0)
        return MilpSolution(MILP_STATUS[res.status], primal, res.objective, res.best_bound,
                            res.nodes_explored, res.unproven_nodes, res.solve_time_ms,
                            res.relaxations, res.batches)


class DeviceVec:
    def __init__(self, k: HipKernels, h, n: int):
        self._k, self._h, self.n = k, h, n

    def __del__(self):
        try:
            if self._h:
                lib().prima_vec_free(self._k._h, self._h)
                self._h = None
        except Exception:
            pass

    def __len__(self):
        return self.n


class DeviceMat:
    def __init__(self, k: HipKernels, h, rows: int, cols: int):
        self._k, self._h, self.rows, self.cols = k, h, rows, cols

    def __del__(self):
        try:
            if self._h:
                lib().prima_mat_free(self._k._h, self._h)
                self._h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# solver-level entry points


def solve(problem: LpProblem, params: PdhgParams = PdhgParams(), device: int = 0) -> LpSolution:
    """Pdhg.solve: a fresh FP64 device backend, closed afterwards."""
    with HipKernels(device, 64) as k:
        return k.solve(problem, params)


def solve_with(problem: LpProblem, params: PdhgParams, kernels: HipKernels,
               warm_start: Optional[WarmStart] = None) -> LpSolution:
    """Pdhg.solveWith: the caller owns the backend's lifetime."""
    return kernels.solve(problem, params, warm_start)


def mixed_solve(problem: LpProblem, params: PdhgParams, device: HipKernels,
                mixed: MixedParams = MixedParams()) -> LpSolution:
    return device.mixed_solve(problem, params, mixed)


def solve_all(problems: Sequence[LpProblem], params: PdhgParams = PdhgParams(),
              kernels: Optional[HipKernels] = None, mixed: Optional[MixedParams] = None) -> list:
    own = kernels is None
    k = kernels or HipKernels(0, 32 if mixed is not None else 64)
    try:
        return k.solve_all(problems, params, mixed=mixed)
    finally:
        if own:
            k.close()


def bnb_solve(problem: LpProblem, integers, params: BnbParams = BnbParams(),
              device: int = 0) -> MilpSolution:
    """BranchAndBound.solve: one FP64 device backend for the whole search."""
    with HipKernels(device, 64) as k:
        return k.milp_solve(problem, integers, params)


def bnb_solve_with(problem: LpProblem, integers, params: BnbParams, kernels: HipKernels,
                   mixed: Optional[MixedParams] = None) -> MilpSolution:
    """BranchAndBound.solveWith against a caller-owned backend."""
    return kernels.milp_solve(problem, integers, params, mixed)


class HipPdhgSolver:
    """LpSolver "prima-pdhg" on the device (Pdhg.scala:108-110)."""

    name = "prima-pdhg"

    def __init__(self, params: PdhgParams = PdhgParams(), device: int = 0):
        self.params, self.device = params, device

    def solve(self, problem: LpProblem) -> LpSolution:
        return solve(problem, self.params, self.device)


class HipMixedSolver:
    """LpSolver "prima-pdhg-mixed" (MixedPrecision.scala:204-214)."""

    name = "prima-pdhg-mixed"

    def __init__(self, params: PdhgParams = PdhgParams(), mixed: MixedParams = MixedParams(),
                 device: int = 0):
        self.params, self.mixed, self.device = params, mixed, device

    def solve(self, problem: LpProblem) -> LpSolution:
        with HipKernels(self.device, 32) as k:
            return k.mixed_solve(problem, self.params, self.mixed)


# ---------------------------------------------------------------------------
# seeded instances


def _from_instance(ptr) -> LpProblem:
    if not ptr:
        raise ValueError("generator rejected its arguments")
    I = ptr.contents
    n, m, nnz = I.n, I.m, I.nnz
    arr = lambda p, k, dt: np.ctypeslib.as_array(p, shape=(max(k, 1),))[:k].astype(dt, copy=True)
    prob = LpProblem(n, m, I.n_eq, arr(I.row_ptr, m + 1, np.int32), arr(I.col, nnz, np.int32),
                     arr(I.val, nnz, np.float64), arr(I.c, n, np.float64),
                     arr(I.q, m, np.float64), arr(I.l, n, np.float64), arr(I.u, n, np.float64),
                     I.offset, I.known_objective)
    if I.x_star:
        prob.x_star = arr(I.x_star, n, np.float64)
    if I.y_star:
        prob.y_star = arr(I.y_star, m, np.float64)
    if I.integers:
        prob.integers = arr(I.integers, I.n_int, np.int32)
    instances_lib().prima_instance_free(ptr)
    return prob


def random_known(seed: int, n: int, m: int, row_len: int = 12, mu: float = 0.0,
                 sigma: float = 0.0, max_len: int = 2048, zero_frac: float = 0.3) -> LpProblem:
    """Random equality LP with a constructed optimum (obj* = c'x* = b'y*).
    row_len > 0 gives fixed row lengths; row_len = 0 draws lognormal(mu, sigma)."""
    return _from_instance(instances_lib().prima_gen_random_known(seed, n, m, row_len, mu, sigma,
                                                                 max_len, zero_frac))


def dispatch(seed: int, buses: int = 200, chords: int = 100, snapshots: int = 24,
             gens_per_bus: int = 3, storage: int = 50, efficiency: float = 0.9,
             perturb: float = 0.0, scenario_seed: int = 0) -> LpProblem:
    """PyPSA-shaped multi-period dispatch LP with DC flows in angle form."""
    return _from_instance(instances_lib().prima_gen_dispatch(seed, buses, chords, snapshots,
                                                             gens_per_bus, storage, efficiency,
                                                             perturb, scenario_seed))


def unit_commitment(seed: int, buses: int = 100, chords: int = 50, snapshots: int = 24,
                    gens_per_bus: int = 3, min_up: int = 3, min_down: int = 3) -> LpProblem:
    """Multi-bus unit commitment over the dispatch network: binary on/off per
    generator and snapshot (`.integers`), minimum up/down times, start-up
    costs U[100,1000], DC flows. The LP is its relaxation."""
    return _from_instance(instances_lib().prima_gen_uc(seed, buses, chords, snapshots,
                                                       gens_per_bus, min_up, min_down))


def config3_args(k: int) -> dict:
    """Instance k of the uneven scenario sweep: buses log-uniform in
    [50, 2000], snapshots in {24, 72, 168}, chords N/2, storage N/4, demand
    perturbed by U[-20%, +20%] per bus; seed 1000 + k."""
    import random
    r = random.Random(1000 + k)
    buses = int(round(math.exp(r.uniform(math.log(50.0), math.log(2000.0)))))
    snapshots = r.choice([24, 72, 168])
    return dict(seed=1000 + k, buses=buses, chords=buses // 2, snapshots=snapshots,
                storage=buses // 4, perturb=0.2, scenario_seed=k)


def config3(k: int) -> LpProblem:
    return dispatch(**config3_args(k))


def config0(seed: int = 0) -> LpProblem:
    """600 vars, 400 equality rows, 12 nnz/row, known optimum."""
    return random_known(seed, 600, 400, row_len=12)


def config1(seed: int = 1) -> LpProblem:
    """200-bus, 24-snapshot dispatch LP (~35k vars)."""
    return dispatch(seed)


def config2(seed: int = 2, n: int = 2_000_000, m: int = 1_000_000,
            mu: float = math.log(20.0) - 0.5) -> LpProblem:
    """Large random LP, lognormal row lengths capped at 2048, known optimum."""
    return random_known(seed, n, m, row_len=0, mu=mu, sigma=1.0, max_len=2048)
