"""CPLEX LP files: reader (reference ``src/sovopt/io/lp_format.py:173-357``).

The format, as the reference reads it:

* sections ``Maximize``/``Maximise``/``Max``, ``Minimize``/``Minimise``/``Min``,
  ``Subject To``/``st``/``s.t.``/``such that``, ``Bounds``, ``General``/
  ``Generals``/``Integer(s)``, ``Binary``/``Binaries``, ``Semi-Continuous``/
  ``Semis``, ``End`` (case-insensitive, each on a line of its own or at the
  start of one);
* ``\\`` starts a comment that runs to the end of the line;
* an objective ``[name:] expr`` whose constant is the offset;
* constraints ``[name:] lhs op rhs`` with terms and constants on both sides
  (moved to the left and right respectively), operators ``<= =< < >= => > =``,
  and ranged rows ``lo <= expr <= hi`` or ``hi >= expr >= lo``; a repeated
  variable's coefficients are summed; unnamed rows are called ``R<k>``;
* bounds ``lo <= x <= hi``, ``x <= v``, ``x >= v``, ``v <= x``, ``x = v``,
  ``x free``, with ``inf``/``infinity`` (signed) for an infinite value;
* General makes columns integer, Binary integer in [0, 1], Semi-Continuous
  semi-continuous.

Columns are numbered in order of first appearance.  A statement may wrap
over lines: it ends at a line end once it has an operator and something after
its last operator, unless the next line starts with an operator.
"""

from __future__ import annotations

import re

import numpy as np

from ..core.problem import ObjSense, Problem, VarKind
from ..core.sparse import VAL, SparseMatrix
from ..core.tolerances import INF
from .mps import _model_name, _text

__all__ = ["read_lp", "LPFormatError"]

_KEYWORDS = [
    ("maximize", "max"), ("maximise", "max"), ("maximum", "max"), ("max", "max"),
    ("minimize", "min"), ("minimise", "min"), ("minimum", "min"), ("min", "min"),
    ("subject to", "st"), ("such that", "st"), ("s.t.", "st"), ("st.", "st"), ("st", "st"),
    ("bounds", "bounds"), ("bound", "bounds"),
    ("generals", "int"), ("general", "int"), ("integers", "int"), ("integer", "int"),
    ("gen", "int"), ("binaries", "bin"), ("binary", "bin"), ("bin", "bin"),
    ("semi-continuous", "semi"), ("semicontinuous", "semi"), ("semis", "semi"),
    ("semi", "semi"), ("end", "end"),
]
_TOKEN = re.compile(r"\s*(<=|=<|>=|=>|<|>|=|:|[+-]|"
                    r"(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|[A-Za-z_!\"#$%&()/,.;?@'`{}|~\[\]]"
                    r"[A-Za-z0-9_!\"#$%&()/,.;?@'`{}|~\[\]^]*)")
_OPS = {"<=": "L", "=<": "L", "<": "L", ">=": "G", "=>": "G", ">": "G", "=": "E"}
_INF_WORDS = {"inf", "infinity"}


class LPFormatError(ValueError):
    """Malformed LP input, with the line number."""

    def __init__(self, message: str, lineno: int = -1):
        self.lineno = lineno
        super().__init__(message + (f" at line {lineno}" if lineno >= 0 else ""))


def _keyword(line: str):
    """(section, rest of the line) when the line opens a section, else None."""
    low = line.lower()
    for word, sec in _KEYWORDS:
        if low.startswith(word):
            rest = line[len(word):]
            if not rest or not (rest[0].isalnum() or rest[0] in "_."):
                if sec == "st" and word in ("st", "st.") and rest.strip().startswith(":"):
                    return None                  # a row named "st"
                if sec in ("max", "min") and rest.strip().startswith(":"):
                    return sec, rest.strip()[1:]   # "max: ..." is the objective's name
                return sec, rest
    return None


def _tokens(text: str, lineno: int) -> list[str]:
    out, pos = [], 0
    text = text.rstrip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise LPFormatError(f"cannot read {text[pos:].strip()[:20]!r}", lineno)
        out.append(m.group(1))
        pos = m.end()
    return out


def _is_number(tok: str) -> bool:
    return tok[0].isdigit() or (tok[0] == "." and len(tok) > 1)


class _Model:
    def __init__(self):
        self.cols: dict[str, int] = {}

    def col(self, name: str) -> int:
        j = self.cols.get(name)
        if j is None:
            j = len(self.cols)
            self.cols[name] = j
        return j


def _expression(toks: list[str], lineno: int, model: _Model):
    """A linear expression: ({col: coef}, constant)."""
    terms: dict[int, float] = {}
    const = 0.0
    k = 0
    while k < len(toks):
        sign = 1.0
        while k < len(toks) and toks[k] in "+-":
            if toks[k] == "-":
                sign = -sign
            k += 1
        if k >= len(toks):
            raise LPFormatError("expression ends with a sign", lineno)
        coef = None
        if _is_number(toks[k]):
            coef = float(toks[k])
            k += 1
        if k < len(toks) and not _is_number(toks[k]) and toks[k] not in "+-":
            name = toks[k]
            if name.lower() in _INF_WORDS:
                value = INF if coef is None else coef * INF
                const += sign * value
                k += 1
                continue
            j = model.col(name)
            terms[j] = terms.get(j, 0.0) + sign * (1.0 if coef is None else coef)
            k += 1
        elif coef is not None:
            const += sign * coef
        else:
            raise LPFormatError(f"unexpected {toks[k]!r}", lineno)
    return terms, const


def _split_ops(toks):
    sides, ops, cur = [], [], []
    for t in toks:
        if t in _OPS:
            sides.append(cur)
            ops.append(_OPS[t])
            cur = []
        else:
            cur.append(t)
    sides.append(cur)
    return sides, ops


def _label(toks):
    if len(toks) >= 2 and toks[1] == ":":
        return toks[0], toks[2:]
    return None, toks


def read_lp(path, name: str | None = None) -> Problem:
    """Read a CPLEX LP file (optionally compressed)."""
    model = _Model()
    sense = None
    obj_terms: dict[int, float] = {}
    offset = 0.0
    rows: list[tuple[dict, float, float, str]] = []       # terms, lb, ub, name
    lo: dict[int, float] = {}
    hi: dict[int, float] = {}
    kinds: dict[int, int] = {}
    section = None
    pending: list[str] = []
    pending_line = 0

    def finish(lineno):
        nonlocal pending, offset
        if not pending:
            return
        toks, at = pending, pending_line
        pending = []
        if section in ("max", "min"):
            label, toks = _label(toks)
            terms, const = _expression(toks, at, model)
            for j, v in terms.items():
                obj_terms[j] = obj_terms.get(j, 0.0) + v
            offset += const
        elif section == "st":
            label, toks = _label(toks)
            sides, ops = _split_ops(toks)
            if len(ops) == 1:
                lt, lc = _expression(sides[0], at, model)
                rt, rc = _expression(sides[1], at, model)
                terms = dict(lt)
                for j, v in rt.items():
                    terms[j] = terms.get(j, 0.0) - v
                rhs = rc - lc
                lb = -INF if ops[0] == "L" else rhs
                ub = INF if ops[0] == "G" else rhs
            elif len(ops) == 2 and ops[0] == ops[1] and ops[0] in "LG":
                a = _expression(sides[0], at, model)
                terms, mc = _expression(sides[1], at, model)
                b = _expression(sides[2], at, model)
                if a[0] or b[0]:
                    raise LPFormatError("a ranged row's outer sides must be constants", at)
                first, last = a[1] - mc, b[1] - mc
                lb, ub = (first, last) if ops[0] == "L" else (last, first)
            else:
                raise LPFormatError("a constraint needs one operator, or two alike", at)
            rows.append((terms, lb, ub, label or f"R{len(rows)}"))
        elif section == "bounds":
            low = [t.lower() for t in toks]
            if len(toks) == 2 and low[1] == "free":
                j = model.col(toks[0])
                lo[j], hi[j] = -INF, INF
                return
            sides, ops = _split_ops(toks)
            exprs = [_expression(s, at, model) for s in sides]
            var = [k for k, e in enumerate(exprs) if e[0]]
            if len(var) != 1 or len(exprs[var[0]][0]) != 1:
                raise LPFormatError("a bound names exactly one variable", at)
            (j, coef), = exprs[var[0]][0].items()
            if coef != 1.0:
                raise LPFormatError("a bound's variable has no coefficient", at)
            for k, op in enumerate(ops):
                left, right = exprs[k], exprs[k + 1]
                if k == var[0]:            # x op value
                    v, rel = right[1], op
                else:                      # value op x: flip
                    v, rel = left[1], {"L": "G", "G": "L", "E": "E"}[op]
                if rel == "L":
                    hi[j] = v
                elif rel == "G":
                    lo[j] = v
                else:
                    lo[j] = hi[j] = v
        else:
            for t in toks:
                j = model.col(t)
                if section == "int":
                    kinds[j] = int(VarKind.INTEGER)
                elif section == "bin":
                    kinds[j] = int(VarKind.INTEGER)
                    lo[j], hi[j] = 0.0, 1.0
                elif section == "semi":
                    kinds[j] = int(VarKind.SEMI_CONTINUOUS)

    with _text(path) as fh:
        lines = fh.read().splitlines()
    lineno = 0
    for lineno, raw in enumerate(lines, 1):
        line = raw.split("\\", 1)[0].strip()
        if not line:
            continue
        kw = _keyword(line)
        if kw is not None:
            finish(lineno)
            section, rest = kw
            if section in ("max", "min"):
                sense = ObjSense.MAXIMISE if section == "max" else ObjSense.MINIMISE
            if section == "end":
                break
            line = rest.strip()
            if not line:
                continue
        if section is None:
            raise LPFormatError("text before the objective section", lineno)
        toks = _tokens(line, lineno)
        if section in ("int", "bin", "semi"):
            pending, pending_line = toks, lineno
            finish(lineno)
            continue
        if section in ("st", "bounds") and pending:
            has_op = any(t in _OPS for t in pending)
            tail = _split_ops(pending)[0][-1]
            starts_op = toks[0] in _OPS
            bound_word = section == "bounds" and pending[-1].lower() == "free"
            if (has_op and tail and not starts_op) or bound_word:
                finish(lineno)
        if not pending:
            pending_line = lineno
        pending = pending + toks
    finish(lineno)
    if section in ("st", "bounds") and pending:
        raise LPFormatError("statement left open", lineno)
    n = len(model.cols)
    if sense is None or n == 0:
        raise LPFormatError("no objective or no variables")
    m = len(rows)
    ri, ci, vx = [], [], []
    for i, (terms, _lb, _ub, _n) in enumerate(rows):
        for j, v in terms.items():
            ri.append(i)
            ci.append(j)
            vx.append(v)
    for i, (terms, lb, ub, _n) in enumerate(rows):
        if not (lb <= ub):
            raise LPFormatError(f"row {_n!r} has lower bound above upper bound")
    A = SparseMatrix.from_triplets(np.array(ri, np.int64), np.array(ci, np.int64),
                                   np.array(vx, np.float64), m, n)
    c = np.zeros(n, dtype=VAL)
    for j, v in obj_terms.items():
        c[j] = v
    col_lb = np.zeros(n, dtype=VAL)
    col_ub = np.full(n, INF, dtype=VAL)
    for j, v in lo.items():
        col_lb[j] = max(v, -INF)
    for j, v in hi.items():
        col_ub[j] = min(v, INF)
    kind = np.zeros(n, dtype=np.uint8)
    for j, k in kinds.items():
        kind[j] = k
    names = [None] * n
    for nm, j in model.cols.items():
        names[j] = nm
    return Problem(A=A, c=c, row_lb=np.array([r[1] for r in rows], dtype=VAL),
                   row_ub=np.array([r[2] for r in rows], dtype=VAL), col_lb=col_lb,
                   col_ub=col_ub, kind=kind, obj_offset=offset, sense=sense,
                   name=name if name is not None else _model_name(path),
                   col_names=names, row_names=[r[3] for r in rows])
