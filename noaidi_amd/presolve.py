"""Exact presolve and postsolve around the device solve (Presolve.scala:27-410).

Four reductions, repeated until nothing changes (at most `max_passes`
passes): columns fixed by their own bounds, rows with no surviving entry
(dropped if they hold, a proof of infeasibility if not), rows with one entry
(folded into that column's bounds), and columns in no surviving row (set from
their cost alone; a cost that runs to an infinite bound is recorded as
unboundedness-if-feasible, which the reduced solve settles). Postsolve
restores primal values, duals and reduced costs of the original problem and
recomputes the objective and KKT error from the restored vectors.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import Optional

import numpy as np

from . import prima


@dataclass
class Stats:
    fixed_columns: int = 0
    empty_columns: int = 0
    empty_rows: int = 0
    singleton_rows: int = 0

    @property
    def total(self) -> int:
        return self.fixed_columns + self.empty_columns + self.empty_rows + self.singleton_rows

    @property
    def is_empty(self) -> bool:
        return self.total == 0


def _csr_T(p: prima.LpProblem):
    """Column access to K: (ptr, row, val) of the transpose."""
    A = p.to_scipy().tocsc()
    return A.indptr, A.indices, A.data


def kkt_evaluate(p: prima.LpProblem, x: np.ndarray, y: np.ndarray):
    """Kkt.evaluate (Kkt.scala:68-125) in float64 on the host: primal residual
    (equality rows Kx - q, inequality shortfall max(q - Kx, 0)), dual residual
    (reduced costs pushing towards an infinite bound), primal and dual
    objectives. Returns (primal_residual, dual_residual, pobj, dobj)."""
    A = p.to_scipy()
    ax = A @ x if p.m else np.zeros(0)
    r = np.concatenate([ax[: p.n_eq] - p.q[: p.n_eq],
                        np.maximum(p.q[p.n_eq:] - ax[p.n_eq:], 0.0)])
    pobj = p.offset + float(np.dot(p.c, x))
    aty = A.T @ y if p.m else np.zeros(p.n)
    rc = p.c - aty
    dobj = p.offset + float(np.dot(p.q, y))
    dsq = 0.0
    for i in range(p.n):
        v = rc[i]
        if v > 0.0:
            if p.l[i] == -math.inf:
                dsq += v * v
            else:
                dobj += p.l[i] * v
        elif v < 0.0:
            if p.u[i] == math.inf:
                dsq += v * v
            else:
                dobj += p.u[i] * v
    return math.sqrt(float(np.dot(r, r))), math.sqrt(dsq), pobj, dobj


def kkt_converged(p: prima.LpProblem, kkt, eps_abs: float, eps_rel: float) -> bool:
    """Kkt.convergence (Kkt.scala:130-143) against the data norms."""
    pres, dres, pobj, dobj = kkt
    return (pres <= eps_abs + eps_rel * float(np.linalg.norm(p.q)) and
            dres <= eps_abs + eps_rel * float(np.linalg.norm(p.c)) and
            abs(pobj - dobj) <= eps_abs + eps_rel * (abs(pobj) + abs(dobj)))


class PresolveResult:
    """PresolveResult (Presolve.scala:67-192)."""

    def __init__(self, original, reduced, stats, undo, column_map, row_map,
                 proven_infeasible, unbounded_if_feasible):
        self.original = original
        self._reduced = reduced
        self.stats = stats
        self._undo = undo  # most recent first
        self._column_map = column_map
        self._row_map = row_map
        self.proven_infeasible = proven_infeasible
        self.unbounded_if_feasible = unbounded_if_feasible

    @property
    def proven_status(self) -> Optional[str]:
        """Only infeasibility can be proven here (Presolve.scala:79-92)."""
        return "PrimalInfeasible" if self.proven_infeasible else None

    @property
    def reduced(self) -> prima.LpProblem:
        if self.proven_status is not None:
            raise RuntimeError(f"presolve settled the problem as {self.proven_status}; "
                               "no reduced problem is left to solve")
        return self._reduced

    def restore(self, sol: prima.LpSolution) -> prima.LpSolution:
        """Map a solution of the reduced problem back (Presolve.scala:108-190)."""
        if self.proven_status is not None:
            raise RuntimeError(f"presolve settled the problem as {self.proven_status}; a "
                               "solution of a reduced problem cannot belong to it")
        p = self.original
        primal = np.zeros(p.n)
        dual = np.zeros(p.m)
        rc = np.zeros(p.n)
        for j, t in enumerate(self._column_map):
            primal[t] = sol.primal[j]
            rc[t] = sol.reduced_costs[j]
        for r, t in enumerate(self._row_map):
            dual[t] = sol.dual[r]
        for u in self._undo:
            kind = u[0]
            if kind == "fixed":
                _, i, v = u
                primal[i] = v
                rc[i] = math.nan  # recomputed from the restored duals below
            elif kind == "empty_col":
                _, i, v, cost = u
                primal[i] = v
                rc[i] = cost
            elif kind == "empty_row":
                dual[u[1]] = 0.0
            else:  # singleton row: priced only where it holds the column
                _, r, c, a, bound = u
                x, red = primal[c], rc[c]
                on_bound = abs(x - bound) <= 1e-7 * max(1.0, abs(bound))
                if on_bound and a != 0.0 and not math.isnan(red) and red != 0.0:
                    price = red / a
                    dual[r] = price
                    rc[c] = red - a * price
                else:
                    dual[r] = 0.0
        if p.m:
            kty = p.to_scipy().T @ dual
        else:
            kty = np.zeros(p.n)
        miss = np.isnan(rc)
        rc[miss] = p.c[miss] - kty[miss]
        status = sol.status
        if self.unbounded_if_feasible and sol.status == "Optimal":
            status = "DualInfeasible"
        pres, dres, pobj, dobj = kkt_evaluate(p, primal, dual)
        return replace(sol, status=status, primal=primal, dual=dual, reduced_costs=rc,
                       objective_value=p.offset + float(np.dot(p.c, primal)),
                       dual_objective_value=dobj, primal_residual=pres, dual_residual=dres)


def presolve(problem: prima.LpProblem, max_passes: int = 20) -> PresolveResult:
    """Presolve.apply (Presolve.scala:202-408)."""
    p = problem
    n, m, n_eq = p.n, p.m, p.n_eq
    col_alive = np.ones(n, bool)
    row_alive = np.ones(m, bool)
    lower, upper, rhs = p.l.copy(), p.u.copy(), p.q.copy()
    fixed_shift = np.zeros(m)
    fixed_value = np.zeros(n)
    undo: list = []
    st = Stats()
    infeasible = False
    unbounded_if_feasible = False
    rptr, rcol, rval = p.row_ptr, p.col, p.val
    cptr, crow, cval = _csr_T(p)

    def row_entries(r):
        return [(int(rcol[k]), float(rval[k])) for k in range(rptr[r], rptr[r + 1])
                if col_alive[rcol[k]]]

    def col_entries(c):
        return [(int(crow[k]), float(cval[k])) for k in range(cptr[c], cptr[c + 1])
                if row_alive[crow[k]]]

    def fix_column(c, value):
        col_alive[c] = False
        fixed_value[c] = value
        st.fixed_columns += 1
        undo.insert(0, ("fixed", c, value))
        for r, a in col_entries(c):
            fixed_shift[r] += a * value

    passes, changed = 0, True
    while changed and passes < max_passes and not infeasible:
        changed = False
        passes += 1
        for c in range(n):
            if infeasible:
                break
            if col_alive[c] and lower[c] == upper[c] and math.isfinite(lower[c]):
                fix_column(c, float(lower[c]))
                changed = True
        for r in range(m):
            if infeasible:
                break
            if not row_alive[r]:
                continue
            entries = row_entries(r)
            adjusted = rhs[r] - fixed_shift[r]
            is_eq = r < n_eq
            if not entries:
                mag = max(abs(rhs[r]), abs(fixed_shift[r]))
                scale = 1e-9 * max(1.0, mag) if math.isfinite(mag) else 1e-9
                holds = abs(adjusted) <= scale if is_eq else adjusted <= scale
                if not holds:
                    infeasible = True
                else:
                    row_alive[r] = False
                    st.empty_rows += 1
                    undo.insert(0, ("empty_row", r))
                    changed = True
            elif len(entries) == 1:
                col, a = entries[0]
                bound = adjusted / a
                lo, hi = lower[col], upper[col]
                if is_eq:
                    lo, hi = max(lo, bound), min(hi, bound)
                elif a > 0.0:
                    lo = max(lo, bound)
                else:
                    hi = min(hi, bound)
                if not math.isfinite(lo) or not math.isfinite(hi):
                    contradictory = lo > hi
                else:
                    contradictory = lo > hi + 1e-9 * max(1.0, max(abs(lo), abs(hi)))
                if contradictory:
                    infeasible = True
                else:
                    lower[col] = lo
                    upper[col] = max(hi, lo)
                    row_alive[r] = False
                    st.singleton_rows += 1
                    undo.insert(0, ("singleton", r, col, a, bound))
                    changed = True
        for c in range(n):
            if infeasible:
                break
            if col_alive[c] and not col_entries(c):
                cost = float(p.c[c])
                if cost > 0.0:
                    at = lower[c]
                elif cost < 0.0:
                    at = upper[c]
                elif math.isfinite(lower[c]):
                    at = lower[c]
                elif math.isfinite(upper[c]):
                    at = upper[c]
                else:
                    at = 0.0
                col_alive[c] = False
                st.empty_columns += 1
                if not math.isfinite(at):
                    unbounded_if_feasible = True
                    undo.insert(0, ("empty_col", c, 0.0, cost))
                else:
                    fixed_value[c] = at
                    undo.insert(0, ("empty_col", c, float(at), cost))
                changed = True

    if infeasible:
        return PresolveResult(p, p, st, [], [], [], True, False)
    column_map = [i for i in range(n) if col_alive[i]]
    row_map = [i for i in range(m) if row_alive[i]]
    old_to_new = {c: j for j, c in enumerate(column_map)}
    b = prima.builder(len(column_map))
    offset = p.offset
    for i in range(n):
        if not col_alive[i]:
            offset += p.c[i] * fixed_value[i]
    b.objectiveOffset(offset)
    for j, old in enumerate(column_map):
        b.objectiveCoefficient(j, float(p.c[old]))
        b.bounds(j, float(lower[old]), float(upper[old]))
    for old in row_map:
        entries = [(old_to_new[c], a) for c, a in row_entries(old)]
        value = float(rhs[old] - fixed_shift[old])
        if old < n_eq:
            b.equalityConstraint(entries, value)
        else:
            b.greaterThan(entries, value)
    reduced, _ = b.build()
    return PresolveResult(p, reduced, st, undo, column_map, row_map, False,
                          unbounded_if_feasible)
