"""MPS reader and writer (sovlp.io), on the CPU.

Known answers: the reference's fixtures (tests/golden/models/, copied from the
reference's tests/fixtures) and the conventions its MPS tests state; every
other case is a small file written here whose model follows from the format's
rules by hand.
"""

import gzip

import numpy as np
import pytest

from sovlp.core.problem import ObjSense, Problem, VarKind
from sovlp.core.sparse import SparseMatrix
from sovlp.core.tolerances import INF
from sovlp.io import MPSError, read_model, read_mps, write_mps

MODELS = __file__.rsplit("/", 1)[0] + "/golden/models/"


def test_testprob_reads_as_written():
    p = read_model(MODELS + "testprob.mps")
    assert p.name == "TESTPROB" and p.sense == ObjSense.MINIMISE
    assert p.row_names == ["LIM1", "LIM2", "MYEQN"]
    assert p.col_names == ["XONE", "YTWO", "ZTHREE"]
    assert list(p.c) == [1.0, 2.0, 3.0]
    np.testing.assert_array_equal(p.A.to_dense(), [[1, 1, 0], [1, 0, 1], [0, -1, 1]])
    assert list(p.row_lb) == [-INF, 1.0, 7.0] and list(p.row_ub) == [4.0, INF, 7.0]
    assert list(p.col_lb) == [0.0, -1.0, 0.0] and list(p.col_ub) == [4.0, 1.0, INF]


def test_the_blending_fixture_is_a_maximisation():
    p = read_model(MODELS + "blend_plan_6h.mps.gz")
    assert p.sense == ObjSense.MAXIMISE          # OBJSENSE on its own line, MAX below
    assert (p.m, p.n, p.nnz) == (654, 1104, 11472)
    assert not p.integer_mask.any()


QUIRKS = """NAME          QUIRKS
OBJSENSE MAX
ROWS
 N  COST
 N  SPARE
 L  RL
 G  RG
 E  REP
 E  REN
 L  RFREE
COLUMNS
    A         COST      1.0   RL        1.0
    A         SPARE     9.0   RG        2.0
    B         COST      -2.5  REP       1.0
    B         REN       1.0   RFREE     1.0
    MARKER    'MARKER'  'INTORG'
    I1        COST      1.0   RL        1.0
    I2        COST      1.0   RG        1.0
    MARKER    'MARKER'  'INTEND'
    C         COST      0.5   REP       3.0
RHS
    RHS       COST      -4.0  RL        10.0
    RHS       RG        2.0   REP       5.0
    RHS       REN       6.0   SPARE     1.0
RANGES
    RNG       RL        3.0   RG        -4.0
    RNG       REP       2.0   REN       -1.5
BOUNDS
 UP BND       A         -1.0
 MI BND       B
 UP BND       B         8.0
 UI BND       I2        7
 FX BND       C         2.5
 UP NEWCOL     5.0
 BV NEWBIN     1
ENDATA
"""


def test_the_format_rules(tmp_path):
    f = tmp_path / "quirks.mps"
    f.write_text(QUIRKS, encoding="utf-8")
    p = read_mps(str(f))
    assert p.sense == ObjSense.MAXIMISE
    assert p.obj_offset == 4.0                       # RHS on the objective, negated
    assert p.row_names == ["RL", "RG", "REP", "REN", "RFREE"]     # SPARE dropped
    assert p.col_names == ["A", "B", "I1", "I2", "C", "NEWCOL", "NEWBIN"]
    assert list(p.c) == [1.0, -2.5, 1.0, 1.0, 0.5, 0.0, 0.0]
    assert p.nnz == 8                                # the SPARE entry is not a row
    # L: [rhs - |r|, rhs]; G: [rhs, rhs + |r|]; E: r >= 0 moves ub, r < 0 moves lb
    assert list(p.row_lb) == [7.0, 2.0, 5.0, 4.5, -INF]
    assert list(p.row_ub) == [10.0, 6.0, 7.0, 6.0, 0.0]
    b = dict(zip(p.col_names, zip(p.col_lb, p.col_ub, p.integer_mask)))
    assert b["A"] == (-INF, -1.0, False)             # negative UP on a default lower bound
    assert b["B"] == (-INF, 8.0, False)              # MI sets only the lower side
    assert b["I1"] == (0.0, 1.0, True)               # integer, no bound line: binary
    assert b["I2"] == (0.0, 7.0, True)
    assert b["C"] == (2.5, 2.5, False)
    assert b["NEWCOL"] == (0.0, 5.0, False)          # created by its bound line (no set name)
    assert b["NEWBIN"] == (0.0, 1.0, True)
    np.testing.assert_array_equal(p.A.to_dense()[:, 0], [1.0, 2.0, 0.0, 0.0, 0.0])


def _fixed(*fields):
    """A data line with fields starting at 1-based columns 2, 5, 15, 25, 40, 50."""
    line = ""
    for start, text in zip((1, 4, 14, 24, 39, 49), fields):
        if text:
            line = line.ljust(start) + text
    return line + "\n"


def test_names_with_spaces_switch_to_the_fixed_columns(tmp_path):
    text = ("NAME          FIXED\nROWS\n" + _fixed("N", "OBJ") + _fixed("G", "CAP A")
            + _fixed("E", "BAL 2") + "COLUMNS\n"
            + _fixed("", "PROD 1", "OBJ", "3.0", "CAP A", "1.0")
            + _fixed("", "PROD 1", "BAL 2", "2.0")
            + _fixed("", "PROD 2", "OBJ", "1.0", "BAL 2", "1.0")
            + "RHS\n" + _fixed("", "RHS", "CAP A", "1.0", "BAL 2", "4.0")
            + "BOUNDS\n" + _fixed("UP", "BND", "PROD 2", "3.0") + "ENDATA\n")
    f = tmp_path / "fixed.mps"
    f.write_text(text, encoding="utf-8")
    p = read_mps(str(f))
    assert p.row_names == ["CAP A", "BAL 2"] and p.col_names == ["PROD 1", "PROD 2"]
    np.testing.assert_array_equal(p.A.to_dense(), [[1.0, 0.0], [2.0, 1.0]])
    assert list(p.row_lb) == [1.0, 4.0] and list(p.row_ub) == [INF, 4.0]
    assert list(p.col_ub) == [INF, 3.0]


def test_compressed_files_are_read(tmp_path):
    plain = (MODELS + "testprob.mps")
    f = tmp_path / "t.mps.gz"
    with gzip.open(f, "wb") as g:
        g.write(open(plain, "rb").read())
    a, b = read_model(plain), read_model(str(f))
    assert b.name == "TESTPROB"
    np.testing.assert_array_equal(a.A.to_dense(), b.A.to_dense())
    assert read_mps(str(f), name="given").name == "TESTPROB"   # the NAME line wins


def _awkward():
    A = SparseMatrix.from_dense(np.array([[1.0, -2.5, 0.0, 0.0],
                                          [0.1, 700.0, 1.0 / 3.0, 0.0],
                                          [1.0, 0.0, 2.0, 1e-300],
                                          [0.0, 3.0, 0.0, 0.0]]))
    return Problem(A=A, c=np.array([700.0, -0.25, 1e-7, 0.0]),
                   row_lb=np.array([-INF, 0.1, 1.5, -INF]),
                   row_ub=np.array([10.0, 0.7, 1.5, INF]),     # L, ranged, E, free
                   col_lb=np.array([0.0, -3.0, 0.0, -INF]),
                   col_ub=np.array([INF, 4.5, 1.0, -2.0]),
                   kind=np.array([VarKind.CONTINUOUS, VarKind.CONTINUOUS, VarKind.INTEGER,
                                  VarKind.CONTINUOUS], dtype=np.uint8),
                   obj_offset=-12.75, sense=ObjSense.MAXIMISE, name="rt",
                   row_names=["L", "RNG", "EQ", "FREE"], col_names=["X", "Y", "Z", "W"])


def test_a_written_model_reads_back_exactly(tmp_path):
    p = _awkward()
    f = tmp_path / "rt.mps"
    write_mps(p, str(f))
    text = f.read_text(encoding="utf-8")
    assert "np.float64" not in text
    q = read_model(str(f))
    assert q.sense == p.sense and q.name == "rt" and q.obj_offset == p.obj_offset
    assert q.row_names == p.row_names and q.col_names == p.col_names
    np.testing.assert_array_equal(q.A.to_dense(), p.A.to_dense())
    for attr in ("c", "row_lb", "row_ub", "col_lb", "col_ub", "kind"):
        np.testing.assert_array_equal(getattr(q, attr), getattr(p, attr), err_msg=attr)
    # an integer column at [0, +inf) must not read back as binary
    p.col_ub[2] = INF
    write_mps(p, str(f))
    assert read_model(str(f)).col_ub[2] == INF


def test_the_blending_fixture_survives_a_round_trip(tmp_path):
    p = read_model(MODELS + "blend_plan_6h.mps.gz")
    f = tmp_path / "b.mps"
    write_mps(p, str(f))
    q = read_model(str(f))
    assert q.sense == p.sense and (q.m, q.n, q.nnz) == (p.m, p.n, p.nnz)
    for attr in ("c", "row_lb", "row_ub", "col_lb", "col_ub"):
        np.testing.assert_array_equal(getattr(q, attr), getattr(p, attr), err_msg=attr)
    np.testing.assert_array_equal(q.A.cx, p.A.cx)


def test_malformed_files_name_their_line(tmp_path):
    f = tmp_path / "bad.mps"
    f.write_text("NAME X\nROWS\n N  OBJ\n L  R1\nCOLUMNS\n    X1  R9  1.0\nENDATA\n")
    with pytest.raises(MPSError) as e:
        read_mps(str(f))
    assert e.value.lineno == 6 and "R9" in str(e.value)
    f.write_text("NAME X\nROWS\n L  R1\nCOLUMNS\n    X1  R1  1.0\nENDATA\n")
    with pytest.raises(MPSError, match="no objective"):
        read_mps(str(f))
    f.write_text("NAME X\nROWS\n N  OBJ\n Q  R1\nENDATA\n")
    with pytest.raises(MPSError, match="unknown row type"):
        read_mps(str(f))
    with pytest.raises(FileNotFoundError):
        read_model(str(tmp_path / "model.lp"))          # .lp goes to the LP reader
