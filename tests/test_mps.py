"""The MPS reader against MpsReaderSuite (modules/prima-mps/src/test/.../
MpsReaderSuite.scala): every convention pinned by solving the parsed problem
to a hand-derived optimum. CPU tests solve with the oracle; the GPU test
solves the same instances on the device."""
import math

import numpy as np
import pytest

from noaidi_amd import mps, prima
from oracle.oracle import OracleLp, default_params

INF = math.inf
BASIC = """NAME          BASIC
ROWS
 N  COST
 G  LIM1
 L  LIM2
COLUMNS
    X         COST      1.0        LIM1      1.0
    X         LIM2      1.0
    Y         COST      1.0        LIM1      1.0
RHS
    RHS       LIM1      2.0        LIM2      3.0
ENDATA
"""


def oracle_solve(text):
    p = mps.from_string(text).problem
    o = OracleLp.from_csr(p.n, p.m, p.n_eq, p.row_ptr, p.col, p.val, p.c, p.q, p.l, p.u,
                          p.offset)
    return o.solve(default_params(eps_abs=1e-9, eps_rel=1e-9))


def range_case(kind, rhs, rng, cost, free=True):
    return f"""NAME          RG
ROWS
 N  COST
 {kind}  R1
COLUMNS
    X         COST      {cost}       R1        1.0
RHS
    RHS       R1        {rhs}
RANGES
    RNG       R1        {rng}
""" + ("BOUNDS\n FR BND       X\n" if free else "") + "ENDATA\n"


FREEROW = """NAME          FREEROW
ROWS
 N  COST
 N  IGNORED
 G  LIM1
COLUMNS
    X         COST      1.0        IGNORED   99.0
    X         LIM1      1.0
RHS
    RHS       LIM1      4.0        IGNORED   1234.0
ENDATA
"""

DUP = """NAME          DUP
ROWS
 N  COST
 G  R1
COLUMNS
    X         COST      1.0
    X         COST      2.0        R1        1.0
RHS
    RHS       R1        5.0
ENDATA
"""

INFEAS = """NAME          INFEAS
ROWS
 N  COST
 G  R1
 L  R2
COLUMNS
    X         COST      1.0        R1        1.0
    X         R2        1.0
RHS
    RHS       R1        5.0        R2        1.0
BOUNDS
 FR BND       X
ENDATA
"""

# (text, status, objective) -- MpsReaderSuite.scala:30-276
SOLVED = {
    "basic": (BASIC, "Optimal", 2.0),
    "decorated": ("\n".join(x for l in BASIC.splitlines()
                            for x in ("* a comment", "", l.replace("    ", "\t"))),
                  "Optimal", 2.0),
    "free-row": (FREEROW, "Optimal", 4.0),
    "objective-constant": (BASIC.replace("    RHS       LIM1      2.0        LIM2      3.0",
                                         "    RHS       LIM1      2.0        LIM2      3.0\n"
                                         "    RHS       COST      10.0"), "Optimal", -8.0),
    "range-G": (range_case("G", 2.0, 3.0, -1.0, free=False), "Optimal", -5.0),
    "range-L": (range_case("L", 5.0, 3.0, 1.0), "Optimal", 2.0),
    "range-E+min": (range_case("E", 10.0, "2.0", 1.0), "Optimal", 10.0),
    "range-E+max": (range_case("E", 10.0, "2.0", -1.0), "Optimal", -12.0),
    "range-E-min": (range_case("E", 10.0, "-2.0", 1.0), "Optimal", 8.0),
    "range-E-max": (range_case("E", 10.0, "-2.0", -1.0), "Optimal", -10.0),
    "duplicate-objective": (DUP, "Optimal", 15.0),
    "infeasible": (INFEAS, "PrimalInfeasible", None),
}


@pytest.mark.parametrize("case", list(SOLVED))
def test_solved_by_the_oracle(case):
    text, status, obj = SOLVED[case]
    s = oracle_solve(text)
    assert s.status == status
    if obj is not None:
        assert abs(s.objective - obj) <= 1e-6


@pytest.mark.gpu
def test_solved_on_the_device(hip64):
    probs = [mps.from_string(t).problem for t, _, _ in SOLVED.values()]
    sols = hip64.solve_all(probs, prima.PdhgParams(eps_abs=1e-9, eps_rel=1e-9))
    for (name, (_, status, obj)), s in zip(SOLVED.items(), sols):
        assert s.status == status, name
        if obj is not None:
            assert abs(s.objective_value - obj) <= 1e-6, name


def test_names_and_shape():
    inst = mps.from_string(BASIC)
    assert inst.name == "BASIC"
    assert inst.column_names == ["X", "Y"] and inst.row_names == ["LIM1", "LIM2"]
    assert not inst.is_mixed_integer and inst.column_index("Y") == 1
    assert inst.column_index("Z") is None
    fr = mps.from_string(FREEROW)
    assert fr.row_names == ["LIM1"] and fr.problem.m == 1


def bounds_of(b):
    text = f"""NAME          BND
ROWS
 N  COST
 G  R1
COLUMNS
    X         COST      1.0        R1        1.0
RHS
    RHS       R1        -1000.0
BOUNDS
{b}
ENDATA
"""
    p = mps.from_string(text).problem
    return (p.l[0], p.u[0])


def test_bound_types():
    # MpsReaderSuite.scala:160-200
    assert bounds_of(" UP BND       X         4.0") == (0.0, 4.0)
    assert bounds_of(" LO BND       X         -2.0") == (-2.0, INF)
    assert bounds_of(" FX BND       X         7.0") == (7.0, 7.0)
    assert bounds_of(" FR BND       X") == (-INF, INF)
    assert bounds_of(" MI BND       X") == (-INF, INF)
    assert bounds_of(" PL BND       X") == (0.0, INF)
    assert bounds_of(" BV BND       X") == (0.0, 1.0)
    assert bounds_of(" MI BND       X         0.0") == (-INF, INF)
    assert bounds_of(" FR BND       X         0.0") == (-INF, INF)
    assert bounds_of(" MI X         0.0") == (-INF, INF)
    assert bounds_of(" FR X") == (-INF, INF)
    assert bounds_of(" MI BND       X\n UP BND2      X         5.0") == (-INF, 5.0)
    assert bounds_of(" UP BND       X         -3.0") == (-INF, -3.0)
    assert bounds_of(" LO BND       X         -9.0\n UP BND       X         -3.0") == (-9.0, -3.0)


def test_integer_markers():
    text = """NAME          INT
ROWS
 N  COST
 G  R1
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    X         COST      1.0        R1        1.0
    MARKER                 'MARKER'                 'INTEND'
    Y         COST      1.0        R1        1.0
RHS
    RHS       R1        1.0
ENDATA
"""
    inst = mps.from_string(text)
    assert inst.is_mixed_integer and inst.integer_columns == {0}


@pytest.mark.parametrize("text", [
    "ROWS\n X  BAD\nENDATA\n",
    "ROWS\n G  R1\nENDATA\n",
    "ROWS\n N  C\n G  R1\nCOLUMNS\n    X    NOSUCHROW   1.0\nENDATA\n",
    "ROWS\n N  C\n G  R1\nCOLUMNS\n    X    R1   notanumber\nENDATA\n",
    "ROWS\n N  C\n G  R1\nCOLUMNS\n    X    R1   1.0\nBOUNDS\n UP BND  NOSUCHCOL  4.0\nENDATA\n",
    "ROWS\n N  C\n G  R1\nCOLUMNS\n    X    R1   1.0\n    Y    R1   1.0\n"
    "BOUNDS\n UP BND  X  4.0  Y  5.0\nENDATA\n",
    "OBJSENSE\n    MAX\nROWS\n N  C\n G  R1\nCOLUMNS\n    X    R1   1.0\nENDATA\n",
    "OBJSENSE MAXIMIZE\nROWS\n N  C\n G  R1\nCOLUMNS\n    X    R1   1.0\nENDATA\n",
])
def test_malformed_input_is_rejected(text):
    with pytest.raises(mps.MpsParseException):
        mps.from_string(text)


def test_numbers_follow_java_parsing():
    assert mps._number("1.5d", "") == 1.5 and mps._number("-2e3", "") == -2000.0
    assert mps._number("Infinity", "") == INF and mps._number("-Infinity", "") == -INF
    for bad in ("inf", "nan", "1_000", "0x10"):
        with pytest.raises(mps.MpsParseException):
            mps._number(bad, "")


def test_file_round_trip(tmp_path):
    f = tmp_path / "basic.mps"
    f.write_text(BASIC)
    inst = mps.from_file(str(f))
    assert inst.name == "BASIC"
    assert np.array_equal(inst.problem.c, mps.from_string(BASIC).problem.c)
