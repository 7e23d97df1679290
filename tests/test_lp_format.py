"""CPLEX LP reader (sovlp.io.read_lp), on the CPU.

Known answers: the reference's fixture testprob.lp is the same model as
testprob.mps (reference tests/test_lp_format.py:34-46), and the format rules
its tests state, on small files written here.
"""

import numpy as np
import pytest

from sovlp.core.problem import ObjSense, VarKind
from sovlp.core.tolerances import INF
from sovlp.io import LPFormatError, read_lp, read_model

MODELS = __file__.rsplit("/", 1)[0] + "/golden/models/"


def lp(tmp_path, text, name="m.lp"):
    f = tmp_path / name
    f.write_text(text, encoding="utf-8")
    return read_lp(str(f))


def test_the_fixture_is_the_mps_model():
    a, b = read_model(MODELS + "testprob.lp"), read_model(MODELS + "testprob.mps")
    assert a.col_names == b.col_names and a.row_names == b.row_names and a.sense == b.sense
    np.testing.assert_array_equal(a.A.to_dense(), b.A.to_dense())
    for attr in ("c", "row_lb", "row_ub", "col_lb", "col_ub", "kind"):
        np.testing.assert_array_equal(getattr(a, attr), getattr(b, attr), err_msg=attr)


def test_objective_sense_constant_and_terms_on_both_sides(tmp_path):
    p = lp(tmp_path, "\\ header\nMaximise\n profit: 4 a + 1.5 b - 2\n + c\n"
                     "Subject To\n cap: 3 a - b + 2 c >= a + 5\n tight: a + a + b <= 9\nEnd\n")
    assert p.sense == ObjSense.MAXIMISE and p.obj_offset == -2.0
    assert p.col_names == ["a", "b", "c"] and list(p.c) == [4.0, 1.5, 1.0]
    np.testing.assert_array_equal(p.A.to_dense(), [[2.0, -1.0, 2.0], [2.0, 1.0, 0.0]])
    assert list(p.row_lb) == [5.0, -INF] and list(p.row_ub) == [INF, 9.0]


def test_ranged_rows_both_ways_and_unnamed_rows(tmp_path):
    p = lp(tmp_path, "min\n x + y\nst\n 2 <= x + 4 y <= 7\n 7 >= x - y >= -1\n x + y = 3\n"
                     " x\n - y\n >= 0\nend\n")
    assert p.m == 4 and p.row_names == ["R0", "R1", "R2", "R3"]
    assert list(p.row_lb) == [2.0, -1.0, 3.0, 0.0] and list(p.row_ub) == [7.0, 7.0, 3.0, INF]
    np.testing.assert_array_equal(p.A.to_dense()[3], [1.0, -1.0])      # wrapped over lines


def test_bounds_integer_binary_and_defaults(tmp_path):
    p = lp(tmp_path, "Minimize\n obj: a + b + c + d + e + f + g\nSubject To\n"
                     " r: a + b + c + d + e + f + g <= 50\nBounds\n -1 <= a <= 3\n b free\n"
                     " c >= -infinity\n c <= 8\n d = 2.5\n 4 >= e\n f >= +inf\n"
                     "General\n a e\nBinary\n g\nEnd\n")
    b = dict(zip(p.col_names, zip(p.col_lb, p.col_ub, p.kind)))
    assert b["a"] == (-1.0, 3.0, VarKind.INTEGER)
    assert b["b"][:2] == (-INF, INF)
    assert b["c"][:2] == (-INF, 8.0)
    assert b["d"][:2] == (2.5, 2.5)
    assert b["e"] == (0.0, 4.0, VarKind.INTEGER)        # "4 >= e": an upper bound
    assert b["f"][0] >= INF
    assert b["g"] == (0.0, 1.0, VarKind.INTEGER)


def test_malformed_and_empty_models_raise(tmp_path):
    with pytest.raises(LPFormatError) as e:
        lp(tmp_path, "Minimize\n obj: x\nSubject To\n c: x + y\nEnd\n")
    assert e.value.lineno == 4
    with pytest.raises(LPFormatError):
        lp(tmp_path, "Minimize\nSubject To\nEnd\n")
    with pytest.raises(LPFormatError):
        lp(tmp_path, "Minimize\n obj: x\nSubject To\n c: 1 <= x <= y\nEnd\n")
