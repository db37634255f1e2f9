"""MPS reader for the linear-programming subset (modules/prima-mps,
MpsReader.scala:62-342), building the standard-form LpProblem through the
host LpBuilder.

The format's conventions are the reference's, each a deliberate choice
(MpsReader.scala:36-60): whitespace-separated tokens, the first N row is the
objective and later N rows are dropped, an RHS on the objective row is the
negated objective constant, MI sets only the lower bound, and a negative UP
bound on a column whose lower bound is still the implicit zero moves that
lower bound to minus infinity. OBJSENSE MAX is rejected.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import Iterable, Optional

import numpy as np

from . import prima


class MpsParseException(ValueError):
    """Malformed or unsupported MPS input (MpsReader.scala:32)."""


@dataclass
class MpsInstance:
    """MpsInstance (MpsReader.scala:14-30): the problem with its names kept."""

    name: str
    problem: prima.LpProblem
    translation: prima.RowTranslation
    row_names: list
    column_names: list
    integer_columns: set = field(default_factory=set)

    @property
    def is_mixed_integer(self) -> bool:
        return bool(self.integer_columns)

    def column_index(self, name: str) -> Optional[int]:
        try:
            return self.column_names.index(name)
        except ValueError:
            return None


# java.lang.Double.parseDouble's grammar (what the reference's toDoubleOption
# accepts), without hexadecimal literals: optional sign, "NaN", "Infinity",
# or a decimal with optional exponent and an optional f/F/d/D suffix.
_JAVA_DOUBLE = re.compile(
    r"[+-]?(NaN|Infinity|((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?))[fFdD]?")


def _number(token: str, line: str) -> float:
    t = token.strip()
    if not _JAVA_DOUBLE.fullmatch(t):
        raise MpsParseException(f"cannot read '{token}' as a number (line '{line}')")
    if t[-1] in "fFdD" and not t.endswith("Infinity"):
        t = t[:-1]
    if t.lstrip("+-") == "NaN":
        return math.nan
    if t.lstrip("+-") == "Infinity":
        return -math.inf if t.startswith("-") else math.inf
    return float(t)


class _Row:
    __slots__ = ("name", "sense", "rhs", "range", "entries")

    def __init__(self, name: str, sense: str):
        self.name, self.sense = name, sense
        self.rhs = 0.0
        self.range: Optional[float] = None
        self.entries: list = []


class _Col:
    __slots__ = ("name", "index", "lower", "upper", "objective", "integer", "lower_explicit")

    def __init__(self, name: str, index: int):
        self.name, self.index = name, index
        self.lower, self.upper = 0.0, math.inf
        self.objective = 0.0
        self.integer = False
        self.lower_explicit = False


def _check_objsense(token: str) -> None:
    if token.upper().startswith("MAX"):
        raise MpsParseException("OBJSENSE MAX: only minimisation problems are read")


def _pairs(fields: list, line: str, section: str):
    if len(fields) % 2:
        raise MpsParseException(f"{section}: name/value fields do not pair up in '{line}'")
    return [(fields[i], _number(fields[i + 1], line)) for i in range(0, len(fields), 2)]


def _drop_set_name(tokens: list) -> list:
    # (row, value) pairs follow an optional set name: odd count -> a name
    return tokens[1:] if len(tokens) % 2 == 1 else tokens


_VALUELESS = {"FR", "MI", "PL", "BV"}


def from_lines(lines: Iterable[str], fallback_name: str = "inline") -> MpsInstance:
    """MpsReader.fromLines (MpsReader.scala:83-262)."""
    name = fallback_name
    objective = ""
    obj_constant = 0.0
    section = ""
    integer_block = False
    rows: dict = {}
    cols: dict = {}
    free_rows: set = set()

    def column(cname: str) -> _Col:
        c = cols.get(cname)
        if c is None:
            c = cols[cname] = _Col(cname, len(cols))
        return c

    def row(rname: str, context: str) -> Optional[_Row]:
        if rname in free_rows or rname == objective:
            return None
        r = rows.get(rname)
        if r is None:
            raise MpsParseException(f"{context} names row '{rname}', which ROWS does not declare")
        return r

    for raw in lines:
        line = raw.rstrip("\r\n").replace("\t", " ")
        if not line.strip() or line.startswith("*"):
            continue
        tokens = line.split()
        if not line.startswith(" "):  # a section header
            section = tokens[0].upper()
            if section == "NAME" and len(tokens) > 1:
                name = tokens[1]
            elif section == "OBJSENSE" and len(tokens) > 1:
                _check_objsense(tokens[1])
            continue
        if section == "ROWS":
            if len(tokens) < 2:
                raise MpsParseException(f"ROWS line has the wrong number of fields: '{line}'")
            kind, rname = tokens[0].upper(), tokens[1]
            if kind == "N":
                if not objective:
                    objective = rname
                else:
                    free_rows.add(rname)
            elif kind in ("L", "G", "E"):
                rows[rname] = _Row(rname, kind)
            else:
                raise MpsParseException(f"unknown row type '{kind}' for row '{rname}'")
        elif section == "COLUMNS":
            if "'MARKER'" in tokens:
                up = [t.upper() for t in tokens]
                if "'INTORG'" in up:
                    integer_block = True
                elif "'INTEND'" in up:
                    integer_block = False
                continue
            if len(tokens) < 3:
                raise MpsParseException(f"COLUMNS line has the wrong number of fields: '{line}'")
            col = column(tokens[0])
            if integer_block:
                col.integer = True
            for rname, value in _pairs(tokens[1:], line, "COLUMNS"):
                if rname == objective:
                    col.objective += value
                elif rname not in free_rows:
                    r = row(rname, "COLUMNS")
                    if r is not None:
                        r.entries.append((col.index, value))
        elif section == "RHS":
            for rname, value in _pairs(_drop_set_name(tokens), line, "RHS"):
                if rname == objective:
                    obj_constant = -value
                elif rname not in free_rows:
                    r = row(rname, "RHS")
                    if r is not None:
                        r.rhs = value
        elif section == "RANGES":
            for rname, value in _pairs(_drop_set_name(tokens), line, "RANGES"):
                r = row(rname, "RANGES")
                if r is not None:
                    r.range = value
        elif section == "BOUNDS":
            if len(tokens) < 2:
                raise MpsParseException(f"BOUNDS line has the wrong number of fields: '{line}'")
            kind = tokens[0].upper()
            # the column is whichever candidate token names a declared column:
            # the set name is optional and valueless kinds may carry a value
            cand = tokens[1:] if kind in _VALUELESS else tokens[1:max(2, len(tokens) - 1)]
            cname = next((t for t in cand if t in cols), None)
            if cname is None:
                raise MpsParseException(
                    f"BOUNDS: none of {', '.join(cand)} is a column of the COLUMNS section "
                    f"('{line}')")
            if len(tokens) > 4:
                raise MpsParseException(f"BOUNDS: too many fields for a single bound in "
                                        f"'{line}'")
            col = column(cname)
            value = (lambda: _number(tokens[-1], line))
            if kind == "UP":
                v = value()
                col.upper = v
                if v < 0.0 and not col.lower_explicit:
                    col.lower = -math.inf
            elif kind == "LO":
                col.lower, col.lower_explicit = value(), True
            elif kind == "FX":
                v = value()
                col.lower = col.upper = v
                col.lower_explicit = True
            elif kind == "FR":
                col.lower, col.upper, col.lower_explicit = -math.inf, math.inf, True
            elif kind == "MI":
                col.lower, col.lower_explicit = -math.inf, True
            elif kind == "PL":
                col.upper = math.inf
            elif kind == "BV":
                col.lower, col.upper, col.integer, col.lower_explicit = 0.0, 1.0, True, True
            elif kind == "LI":
                col.lower, col.integer, col.lower_explicit = value(), True, True
            elif kind == "UI":
                col.upper, col.integer = value(), True
            else:
                raise MpsParseException(f"bound type '{kind}' (column '{cname}') is not one of the MPS types")
        elif section == "OBJSENSE":
            _check_objsense(tokens[0])
        elif section in ("ENDATA", ""):
            pass
        else:
            raise MpsParseException(f"section '{section}' does not take data lines")

    if not objective:
        raise MpsParseException("the ROWS section declares no N row to serve as objective")
    return _build(name, obj_constant, list(rows.values()), list(cols.values()))


def _bracket(r: _Row):
    """Two-sided bounds of a row from its sense, RHS and RANGES entry
    (MpsReader.scala:324-340): G extends up, L down, E by the range's sign."""
    if r.range is None:
        return {"L": (-math.inf, r.rhs), "G": (r.rhs, math.inf), "E": (r.rhs, r.rhs)}[r.sense]
    w = abs(r.range)
    if r.sense == "G":
        return r.rhs, r.rhs + w
    if r.sense == "L":
        return r.rhs - w, r.rhs
    return (r.rhs, r.rhs + w) if r.range >= 0.0 else (r.rhs - w, r.rhs)


def _build(name, obj_constant, rows, cols) -> MpsInstance:
    b = prima.builder(len(cols))
    b.objectiveOffset(obj_constant)
    for c in cols:
        b.objectiveCoefficient(c.index, c.objective)
        if c.lower > c.upper:
            raise MpsParseException(f"column '{c.name}': lower bound {c.lower} exceeds upper bound {c.upper}")
        b.bounds(c.index, c.lower, c.upper)
    for r in rows:
        lo, hi = _bracket(r)
        b.constraint(r.entries, lo, hi)
    problem, translation = b.build()
    return MpsInstance(name, problem, translation, [r.name for r in rows],
                       [c.name for c in cols], {c.index for c in cols if c.integer})


def from_string(content: str, fallback_name: str = "inline") -> MpsInstance:
    return from_lines(content.splitlines(), fallback_name)


def from_file(path: str) -> MpsInstance:
    with open(path) as f:
        return from_lines(f, os.path.basename(path))
