"""MPS files: reader and writer (reference ``src/sovopt/io/mps.py:104-550``).

The file conventions are those the reference reads, which are the ones
CPLEX, Gurobi, SCIP and MIPLIB keep:

* free format (whitespace-separated) and fixed format.  A ROWS line that
  does not split into exactly a type and a name has a name with a space in it,
  and from then on every data line is cut on the fixed column positions
  (2-3, 5-12, 15-22, 25-36, 40-47, 50-61);
* plain, ``.gz``, ``.bz2``, ``.xz`` / ``.lzma`` files;
* ``OBJSENSE`` with ``MAX``/``MIN`` on the same line or the next one;
* the first N row is the objective, later N rows are dropped;
* an RHS entry on the objective row is the NEGATED objective constant;
* RANGES: on an L row lb = rhs - |r|, on a G row ub = rhs + |r|, on an E row
  the sign of r picks the side that moves;
* BOUNDS UP, LO, FX, FR, MI (lower only), PL, BV, LI, UI, SC/SI; a negative
  UP on a continuous column with no lower bound line sets lb = -inf; the
  bound-set name may be missing; a bound on a column absent from COLUMNS
  creates it;
* MARKER INTORG/INTEND integer blocks; an integer column with no bound line
  of any kind is binary;
* QUADOBJ / QMATRIX / QSECTION (lower triangle, mirrored).

``MPSError`` carries the line number.  The reader builds the model in one
pass: rows and columns are numbered in order of appearance.
"""

from __future__ import annotations

import bz2
import gzip
import io
import lzma
import os

import numpy as np

from ..core.problem import ObjSense, Problem, VarKind
from ..core.sparse import VAL, SparseMatrix
from ..core.tolerances import INF

__all__ = ["read_mps", "write_mps", "MPSError"]

_HEADERS = {"NAME", "ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS", "ENDATA", "OBJSENSE",
            "OBJSENS", "QUADOBJ", "QMATRIX", "QSECTION", "SOS", "INDICATORS"}
_QUAD = {"QUADOBJ", "QMATRIX", "QSECTION"}
_FIELDS = ((1, 3), (4, 12), (14, 22), (24, 36), (39, 47), (49, 61))
_COMPRESSED = {".gz": gzip.open, ".bz2": bz2.open, ".xz": lzma.open, ".lzma": lzma.open}


class MPSError(ValueError):
    """Malformed MPS input (``lineno`` = -1 when no line is to blame)."""

    def __init__(self, message: str, lineno: int = -1, line: str = ""):
        self.lineno, self.line = lineno, line
        at = f" at line {lineno}" if lineno >= 0 else ""
        shown = f": {line.rstrip()!r}" if line else ""
        super().__init__(message + at + shown)


def _text(path):
    ext = os.path.splitext(str(path))[1].lower()
    opener = _COMPRESSED.get(ext)
    if opener is not None:
        return io.TextIOWrapper(opener(path, "rb"), encoding="utf-8", errors="replace")
    return open(path, "r", encoding="utf-8", errors="replace")


def _format_extension(path) -> str:
    """Lower-case extension of the model held in ``path``, looked up through one
    compression suffix of ``_COMPRESSED`` (``plan.lp.gz`` -> ``.lp``)."""
    stem, ext = os.path.splitext(os.path.basename(str(path)))
    if ext.lower() in _COMPRESSED:
        ext = os.path.splitext(stem)[1]
    return ext.lower()


def _model_name(path) -> str:
    base = os.path.basename(str(path))
    stem, ext = os.path.splitext(base)
    if ext.lower() in _COMPRESSED:
        base = stem
    return os.path.splitext(base)[0]


def _columns_of(raw: str) -> list[str]:
    out = []
    for a, b in _FIELDS:
        if a >= len(raw):
            break
        field = raw[a:b].strip()
        if field:
            out.append(field)
    return out


class _Builder:
    """Accumulates one model while the file is read."""

    def __init__(self, name: str):
        self.name = name
        self.sense = ObjSense.MINIMISE
        self.objective: str | None = None
        self.dropped: set[str] = set()      # N rows after the first
        self.rows: dict[str, int] = {}
        self.row_type: list[str] = []
        self.cols: dict[str, int] = {}
        self.col_int: list[bool] = []
        self.entries: list[tuple[int, int, float]] = []
        self.cost: dict[int, float] = {}
        self.rhs: dict[int, float] = {}
        self.rng: dict[int, float] = {}
        self.offset = 0.0
        self.lo: dict[int, float] = {}
        self.hi: dict[int, float] = {}
        self.has_bound: set[int] = set()
        self.integer: set[int] = set()
        self.semi: dict[int, float] = {}
        self.quad: list[tuple[int, int, float]] = []
        self.in_int_block = False

    def column(self, cname: str, create_int: bool) -> int:
        j = self.cols.get(cname)
        if j is None:
            j = len(self.col_int)
            self.cols[cname] = j
            self.col_int.append(create_int)
        return j

    def row_of(self, rname: str, lineno: int, raw: str, where: str) -> int | None:
        """Index of a constraint row; None for the objective or a dropped N row."""
        if rname == self.objective or rname in self.dropped:
            return None
        i = self.rows.get(rname)
        if i is None:
            raise MPSError(f"unknown row {rname!r}{where}", lineno, raw)
        return i


def _number(tok: str, lineno: int, raw: str) -> float:
    try:
        return float(tok)
    except ValueError:
        raise MPSError(f"not a number: {tok!r}", lineno, raw) from None


def _pairs(tok: list[str], lineno: int, raw: str):
    """(name, value) pairs after an optional set name (told apart by parity)."""
    first = 1 if len(tok) % 2 == 1 else 0
    for k in range(first, len(tok) - 1, 2):
        yield tok[k], _number(tok[k + 1], lineno, raw)


def _rows_line(b: _Builder, tok, lineno, raw):
    kind, rname = tok[0].upper(), tok[1]
    if kind == "N":
        if b.objective is None:
            b.objective = rname
        else:
            b.dropped.add(rname)
        return
    if kind not in ("L", "G", "E"):
        raise MPSError(f"unknown row type {kind!r}", lineno, raw)
    if rname in b.rows:
        raise MPSError(f"duplicate row name {rname!r}", lineno, raw)
    b.rows[rname] = len(b.row_type)
    b.row_type.append(kind)


def _columns_line(b: _Builder, tok, lineno, raw):
    if "'MARKER'" in raw.upper():
        up = raw.upper()
        if "INTORG" in up:
            b.in_int_block = True
        elif "INTEND" in up:
            b.in_int_block = False
        return
    if len(tok) < 3:
        tok = _columns_of(raw)
    if len(tok) < 3:
        raise MPSError("COLUMNS entry needs name, row, value", lineno, raw)
    j = b.column(tok[0], b.in_int_block)
    for rname, v in _pairs([""] + tok[1:], lineno, raw):
        if rname == b.objective:
            b.cost[j] = b.cost.get(j, 0.0) + v
            continue
        i = b.row_of(rname, lineno, raw, "")
        if i is not None:
            b.entries.append((i, j, v))


def _rhs_line(b: _Builder, tok, lineno, raw):
    if len(tok) < 3:
        tok = _columns_of(raw)
    for rname, v in _pairs(tok, lineno, raw):
        if rname == b.objective:
            b.offset = -v
            continue
        i = b.row_of(rname, lineno, raw, " in RHS")
        if i is not None:
            b.rhs[i] = v


def _ranges_line(b: _Builder, tok, lineno, raw):
    if len(tok) < 3:
        tok = _columns_of(raw)
    for rname, v in _pairs(tok, lineno, raw):
        i = b.row_of(rname, lineno, raw, " in RANGES")
        if i is not None:
            b.rng[i] = v


def _bounds_line(b: _Builder, tok, lineno, raw):
    if len(tok) < 3:
        tok = _columns_of(raw)
    if len(tok) < 3:
        raise MPSError("BOUNDS entry too short", lineno, raw)
    btype = tok[0].upper()
    # with a bound-set name the column is the third token, without it the second
    if tok[2] in b.cols:
        cname, rest = tok[2], tok[3:]
    else:
        cname, rest = tok[1], tok[2:]
    j = b.column(cname, False)
    v = _number(rest[0], lineno, raw) if rest else 0.0
    b.has_bound.add(j)
    if btype == "UP":
        b.hi[j] = v
        if v < 0.0 and j not in b.lo and not b.col_int[j]:
            b.lo[j] = -INF
    elif btype == "LO":
        b.lo[j] = v
    elif btype == "FX":
        b.lo[j] = b.hi[j] = v
    elif btype == "FR":
        b.lo[j], b.hi[j] = -INF, INF
    elif btype == "MI":
        b.lo[j] = -INF
    elif btype in ("PL", "BV_PL"):
        b.hi[j] = INF
    elif btype == "BV":
        b.lo[j], b.hi[j] = 0.0, 1.0
        b.integer.add(j)
    elif btype == "LI":
        b.lo[j] = v
        b.integer.add(j)
    elif btype == "UI":
        b.hi[j] = v
        b.integer.add(j)
    elif btype in ("SC", "SI"):
        b.semi[j] = v if v != 0.0 else INF
        b.hi[j] = b.semi[j]
        if btype == "SI":
            b.integer.add(j)
    else:
        raise MPSError(f"unknown bound type {btype!r}", lineno, raw)


def _quad_line(b: _Builder, tok, lineno, raw):
    if len(tok) < 3:
        tok = _columns_of(raw)
    if len(tok) < 3:
        return
    ja, jb = b.cols.get(tok[0]), b.cols.get(tok[1])
    if ja is None or jb is None:
        raise MPSError("unknown column in QUADOBJ", lineno, raw)
    b.quad.append((ja, jb, _number(tok[2], lineno, raw)))


_HANDLERS = {"COLUMNS": _columns_line, "RHS": _rhs_line, "RANGES": _ranges_line,
             "BOUNDS": _bounds_line, "QUADOBJ": _quad_line}


def _assemble(b: _Builder) -> Problem:
    if b.objective is None:
        raise MPSError("no objective (type N) row found")
    n, m = len(b.col_int), len(b.row_type)
    if n == 0:
        raise MPSError("model has no columns")
    ent = np.array(b.entries, dtype=np.float64).reshape(-1, 3)
    A = SparseMatrix.from_triplets(ent[:, 0].astype(np.int64), ent[:, 1].astype(np.int64),
                                   ent[:, 2], m, n)
    c = np.zeros(n, dtype=VAL)
    for j, v in b.cost.items():
        c[j] = v
    rhs = np.zeros(m, dtype=VAL)
    for i, v in b.rhs.items():
        rhs[i] = v
    types = np.array(b.row_type, dtype="<U1")
    row_lb = np.where(types == "L", -INF, rhs)
    row_ub = np.where(types == "G", INF, rhs)
    for i, r in b.rng.items():
        if types[i] == "L":
            row_lb[i] = rhs[i] - abs(r)
        elif types[i] == "G":
            row_ub[i] = rhs[i] + abs(r)
        elif r >= 0.0:
            row_ub[i] = rhs[i] + r
        else:
            row_lb[i] = rhs[i] + r
    col_lb = np.zeros(n, dtype=VAL)
    col_ub = np.full(n, INF, dtype=VAL)
    is_int = np.array(b.col_int, dtype=bool)
    for j in np.flatnonzero(is_int):
        if j not in b.has_bound:
            col_ub[j] = 1.0                     # MPSX: a bare integer column is binary
    for j, v in b.lo.items():
        col_lb[j] = v
    for j, v in b.hi.items():
        col_ub[j] = v
    kind = np.zeros(n, dtype=np.uint8)
    kind[is_int] = VarKind.INTEGER
    for j in b.integer:
        kind[j] = VarKind.INTEGER
    for j in b.semi:
        kind[j] = VarKind.SEMI_INTEGER if j in b.integer else VarKind.SEMI_CONTINUOUS
    Q = None
    if b.quad:
        q = np.array(b.quad, dtype=np.float64).reshape(-1, 3)
        off = q[q[:, 0] != q[:, 1]]
        qr = np.concatenate([q[:, 0], off[:, 1]]).astype(np.int64)
        qc = np.concatenate([q[:, 1], off[:, 0]]).astype(np.int64)
        Q = SparseMatrix.from_triplets(qr, qc, np.concatenate([q[:, 2], off[:, 2]]), n, n)
    names_c = [None] * n
    for name, j in b.cols.items():
        names_c[j] = name
    names_r = [None] * m
    for name, i in b.rows.items():
        names_r[i] = name
    return Problem(A=A, c=c, row_lb=row_lb, row_ub=row_ub, col_lb=col_lb, col_ub=col_ub,
                   kind=kind, Q=Q, obj_offset=b.offset, sense=b.sense, name=b.name,
                   col_names=names_c, row_names=names_r)


def read_mps(path, name: str | None = None) -> Problem:
    """Read an MPS file (free or fixed format, optionally compressed)."""
    b = _Builder(name if name is not None else _model_name(path))
    section = None
    fixed = False
    sense_next = False
    with _text(path) as fh:
        for lineno, raw in enumerate(fh, 1):
            if not raw.strip() or raw[0] in "*$":
                continue
            if not raw[0].isspace():
                head = raw.split()
                key = head[0].upper()
                if key in _HEADERS:
                    if key == "ENDATA":
                        break
                    section = "QUADOBJ" if key in _QUAD else key
                    if key == "NAME" and len(head) > 1:
                        b.name = head[1]
                    if key in ("OBJSENSE", "OBJSENS"):
                        sense_next = len(head) == 1
                        if len(head) > 1:
                            b.sense = (ObjSense.MAXIMISE if head[1].upper().startswith("MAX")
                                       else ObjSense.MINIMISE)
                    continue
                # an unindented line that is no header is data (OBJSENSE values)
            tok = _columns_of(raw) if fixed else raw.split()
            if not tok:
                continue
            if sense_next:
                b.sense = (ObjSense.MAXIMISE if tok[0].upper().startswith("MAX")
                           else ObjSense.MINIMISE)
                sense_next = False
                continue
            if section == "ROWS":
                if len(tok) != 2:
                    tok = _columns_of(raw)
                    fixed = len(tok) == 2
                if len(tok) < 2:
                    raise MPSError("ROWS entry needs a type and a name", lineno, raw)
                _rows_line(b, tok, lineno, raw)
            elif section in _HANDLERS:
                _HANDLERS[section](b, tok, lineno, raw)
    return _assemble(b)


def _fmt(v: float) -> str:
    return repr(float(v))                     # shortest text that reads back exactly


def write_mps(prob: Problem, path) -> None:
    """Write ``prob`` as a free-format MPS file that ``read_mps`` reads back exactly
    (names without spaces are kept, others become R<i> / C<j>)."""
    m, n = prob.m, prob.n

    def names(given, k, prefix):
        ok = (given is not None and len(given) == k
              and all(isinstance(s, str) and s and not any(ch.isspace() for ch in s)
                      for s in given) and len(set(given)) == k)
        return list(given) if ok else [f"{prefix}{i}" for i in range(k)]

    rn, cn = names(prob.row_names, m, "R"), names(prob.col_names, n, "C")
    obj = "OBJ"
    while obj in rn:
        obj += "_"
    lb, ub = prob.row_lb, prob.row_ub
    lines = [f"NAME          {prob.name or 'model'}"]
    if int(prob.sense) == int(ObjSense.MAXIMISE):
        lines += ["OBJSENSE", "    MAX"]
    lines += ["ROWS", f" N  {obj}"]
    rtype, rhs, rng = [], [], {}
    for i in range(m):
        lo_inf, hi_inf = lb[i] <= -INF, ub[i] >= INF
        if lb[i] == ub[i]:
            t, r = "E", lb[i]
        elif lo_inf and hi_inf:
            t, r = "G", -INF                     # a free row: G with the -inf sentinel
        elif lo_inf:
            t, r = "L", ub[i]
        elif hi_inf:
            t, r = "G", lb[i]
        else:
            # a ranged row: L with lb = ub - r, or E with ub = lb + r, whichever
            # the reader's arithmetic gives back exactly
            t, r = "L", ub[i]
            rng[i] = ub[i] - lb[i]
            if ub[i] - rng[i] != lb[i] and lb[i] + rng[i] == ub[i]:
                t, r = "E", lb[i]
        rtype.append(t)
        rhs.append(r)
        lines.append(f" {t}  {rn[i]}")
    lines.append("COLUMNS")
    A = prob.A
    is_int = np.asarray(prob.kind) != VarKind.CONTINUOUS
    in_block = False
    marks = 0
    for j in range(n):
        if bool(is_int[j]) != in_block:
            tag = "INTORG" if not in_block else "INTEND"
            lines.append(f"    MARKER{marks}  'MARKER'  '{tag}'")
            marks += 1
            in_block = not in_block
        if prob.c[j] != 0.0:
            lines.append(f"    {cn[j]}  {obj}  {_fmt(prob.c[j])}")
        for q in range(int(A.cp[j]), int(A.cp[j + 1])):
            lines.append(f"    {cn[j]}  {rn[int(A.ci[q])]}  {_fmt(A.cx[q])}")
        if prob.c[j] == 0.0 and A.cp[j] == A.cp[j + 1]:
            lines.append(f"    {cn[j]}  {obj}  0.0")     # the column must appear
    if in_block:
        lines.append(f"    MARKER{marks}  'MARKER'  'INTEND'")
    lines.append("RHS")
    if prob.obj_offset != 0.0:
        lines.append(f"    RHS  {obj}  {_fmt(-prob.obj_offset)}")
    for i in range(m):
        if rhs[i] != 0.0:
            lines.append(f"    RHS  {rn[i]}  {_fmt(rhs[i])}")
    if rng:
        lines.append("RANGES")
        for i, r in rng.items():
            lines.append(f"    RNG  {rn[i]}  {_fmt(r)}")
    lines.append("BOUNDS")
    for j in range(n):
        lo, hi = float(prob.col_lb[j]), float(prob.col_ub[j])
        integer = bool(is_int[j])
        if lo <= -INF and hi >= INF:
            lines.append(f" FR BND  {cn[j]}")
            continue
        if lo == hi:
            lines.append(f" FX BND  {cn[j]}  {_fmt(lo)}")
            continue
        # integer columns always get both sides (a bare one reads as binary);
        # LO before UP, so a negative UP never reads as lb = -inf
        if lo <= -INF:
            lines.append(f" MI BND  {cn[j]}")
        elif lo != 0.0 or integer or hi < 0.0:
            lines.append(f" LO BND  {cn[j]}  {_fmt(lo)}")
        if hi < INF:
            lines.append(f" UP BND  {cn[j]}  {_fmt(hi)}")
        elif integer:
            lines.append(f" PL BND  {cn[j]}")
    lines.append("ENDATA")
    with open(path, "w", encoding="utf-8") as fh:
        fh.write("\n".join(lines) + "\n")
