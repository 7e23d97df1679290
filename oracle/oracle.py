"""ctypes loader for the CPU oracle (``liboracle_pdlp.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, the smoke entry point and the
``cpu_baseline`` leg of ``bench.py``, never by the ``sovlp`` package.

``solve(prob, ...)`` restates ``solve_pdlp`` (reference ``src/sovopt/lp/pdlp.py:238-494``)
and ``scale(prob)`` restates ``scale_problem(..., "pdlp")``
(``src/sovopt/numerics/scaling.py:374-478``); see ``pdlp_oracle.c`` for the line map.
``residual`` and ``certify`` restate ``compensated_residual``
(``src/sovopt/numerics/refine.py:48-110``) and ``certified_bound``
(``src/sovopt/mip/safebound.py:93-143, 175-186, 252-307``) with the feasibility lines of
``bench/verify.py:208-224``; see ``certify_oracle.c``.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_pdlp.so")
# the same source with fused multiply-adds (rounding study only)
FMA_LIB_PATH = os.path.join(HERE, "liboracle_pdlp_fma.so")

_libs: dict = {}


class _Params(C.Structure):
    _fields_ = [("eps_abs", C.c_double), ("eps_rel", C.c_double), ("feas_cap", C.c_double),
                ("max_iter", C.c_longlong), ("time_limit", C.c_double),
                ("check_every", C.c_int), ("stepsize_every", C.c_int),
                ("beta_sufficient", C.c_double), ("beta_necessary", C.c_double),
                ("beta_artificial", C.c_double), ("primal_weight_smoothing", C.c_double)]


class _Info(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong), ("restarts", C.c_longlong),
                ("rejected_steps", C.c_longlong), ("primal_res", C.c_double),
                ("dual_res", C.c_double), ("gap", C.c_double), ("primal_obj", C.c_double),
                ("dual_obj", C.c_double), ("step_size", C.c_double),
                ("primal_weight", C.c_double), ("objective", C.c_double),
                ("scale_obj", C.c_double), ("history_len", C.c_longlong)]


def build(force: bool = False, fma: bool = False) -> str:
    path = FMA_LIB_PATH if fma else LIB_PATH
    src = max(os.path.getmtime(os.path.join(HERE, f)) for f in ("pdlp_oracle.c", "certify_oracle.c"))
    if force or not os.path.exists(path) or os.path.getmtime(path) < src:
        subprocess.run(["make", "-C", HERE] + (["-B"] if force else []) + [os.path.basename(path)],
                       check=True, capture_output=True)
    return path


def lib(fma: bool = False):
    """The oracle library; ``fma``: the build with contracted multiply-adds."""
    if fma not in _libs:
        path = FMA_LIB_PATH if fma else LIB_PATH
        if not os.path.exists(path):
            build(fma=fma)
        L = C.CDLL(path)
        P = C.c_void_p
        L.oracle_pdlp_solve.restype = C.c_int
        L.oracle_pdlp_solve.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P,
                                        C.c_double, C.c_int, P, C.POINTER(_Params), P, P, P,
                                        C.POINTER(_Info), P, C.c_longlong, C.c_int]
        L.oracle_scale.restype = C.c_int
        L.oracle_scale.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P]
        L.oracle_spmv.restype = None
        L.oracle_spmv.argtypes = [C.c_int, P, P, P, P, P]
        L.oracle_residual.restype = None
        L.oracle_residual.argtypes = [C.c_int, P, P, P, P, P, P]
        L.oracle_certify.restype = None
        L.oracle_certify.argtypes = [C.c_int, C.c_int, P, P, P, P, P, P, P, P, P, P, P,
                                     C.c_double, C.c_int, P, P, C.c_double, P, P, P, P, P,
                                     C.POINTER(_Certificate)]
        _libs[fma] = L
    return _libs[fma]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _Certificate(C.Structure):
    _fields_ = [("bound", C.c_double), ("perturbation", C.c_double), ("objective", C.c_double),
                ("col_violation", C.c_double), ("row_violation", C.c_double),
                ("vacuous", C.c_int), ("pad", C.c_int)]


@dataclass
class Certificate:
    bound: float
    perturbation: float
    objective: float
    col_violation: float
    row_violation: float
    vacuous: bool
    d: np.ndarray           # c - A^T yy, minimisation sense
    act: np.ndarray         # compensated row activities A x


def residual(rp, ri, rx, x, b) -> np.ndarray:
    """``b - A x``, compensated, as the reference's ``compensated_residual``."""
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    rx = np.ascontiguousarray(rx, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(rp.size - 1)
    lib().oracle_residual(rp.size - 1, _p(rp), _p(ri), _p(rx), _p(x), _p(b), _p(out))
    return out


def certify(prob, x, y, dual_tol: float = 1e-6) -> Certificate:
    """The reference's ``certified_bound(prob, x, y, dual_tol)`` and the feasibility
    figures of its ``verify``."""
    A = prob.A
    m, n = A.m, A.n
    arrs = [np.ascontiguousarray(a) for a in (A.rp, A.ri, A.rx, A.cp, A.ci, A.cx, prob.c,
                                              prob.row_lb, prob.row_ub, prob.col_lb,
                                              prob.col_ub)]
    arrs[0] = arrs[0].astype(np.int64)
    arrs[3] = arrs[3].astype(np.int64)
    arrs[1] = arrs[1].astype(np.int32)
    arrs[4] = arrs[4].astype(np.int32)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    yy, d, act, rt, ct = np.empty(m), np.empty(n), np.empty(m), np.empty(m), np.empty(n)
    out = _Certificate()
    lib().oracle_certify(m, n, *[_p(a) for a in arrs], float(prob.obj_offset),
                         int(int(prob.sense) == -1), _p(x), _p(y), float(dual_tol), _p(yy),
                         _p(d), _p(act), _p(rt), _p(ct), C.byref(out))
    return Certificate(out.bound, out.perturbation, out.objective, out.col_violation,
                       out.row_violation, bool(out.vacuous), d, act)


@dataclass
class OracleResult:
    status: int
    x: np.ndarray
    y: np.ndarray
    reduced_costs: np.ndarray
    objective: float
    iterations: int
    restarts: int
    rejected_steps: int
    step_size: float
    primal_weight: float
    primal_res: float
    dual_res: float
    gap: float
    scale_obj: float
    history: list = field(default_factory=list)


def power_start(n: int) -> np.ndarray:
    """``np.random.default_rng(0).standard_normal(n)`` (pdlp.py:142-143)."""
    return np.random.default_rng(0).standard_normal(n)


def solve(prob, eps_abs=1e-8, eps_rel=1e-8, feas_cap=1e-6, max_iter=200_000, time_limit=600.0,
          check_every=64, stepsize_every=8, threads=0, hist_cap=100_000,
          fma: bool = False) -> OracleResult:
    """``fma``: run the build with contracted multiply-adds (same algorithm, other
    rounding), to measure how rounding alone moves the trajectory."""
    A = prob.A
    m, n = A.m, A.n
    par = _Params(eps_abs, eps_rel, feas_cap, max_iter, time_limit, check_every, stepsize_every,
                  0.2, 0.8, 0.36, 0.5)
    info = _Info()
    x = np.zeros(n)
    y = np.zeros(m)
    d = np.zeros(n)
    hist = np.zeros((hist_cap, 5))
    arrs = [np.ascontiguousarray(a) for a in (A.cp, A.ci, A.cx, A.rp, A.ri, A.rx, prob.c,
                                              prob.row_lb, prob.row_ub, prob.col_lb, prob.col_ub)]
    v0 = power_start(n)
    rc = lib(fma).oracle_pdlp_solve(m, n, *[_p(a) for a in arrs], float(prob.obj_offset),
                                 int(int(prob.sense) == -1), _p(v0), C.byref(par), _p(x), _p(y),
                                 _p(d), C.byref(info), _p(hist), hist_cap, threads)
    if rc != 0:
        raise RuntimeError(f"oracle_pdlp_solve failed ({rc})")
    h = min(info.history_len, hist_cap)
    return OracleResult(info.status, x, y, d, info.objective, info.iterations, info.restarts,
                        info.rejected_steps, info.step_size, info.primal_weight,
                        info.primal_res, info.dual_res, info.gap, info.scale_obj,
                        [tuple(r) for r in hist[:h].tolist()])


def scale(prob, c=None):
    """``(row, col, obj)`` factors of ``scale_problem(prob, "pdlp")``; ``c`` defaults
    to ``prob.c`` (pass the negated costs for a maximisation, as solve_pdlp does)."""
    A = prob.A
    row = np.zeros(A.m)
    col = np.zeros(A.n)
    obj = C.c_double()
    cc = np.ascontiguousarray(prob.c if c is None else c, dtype=np.float64)
    arrs = [np.ascontiguousarray(a) for a in (A.cp, A.ci, A.cx, A.rp, A.ri, A.rx)]
    rc = lib().oracle_scale(A.m, A.n, *[_p(a) for a in arrs], _p(cc), _p(row), _p(col),
                            C.byref(obj), None)
    if rc != 0:
        raise RuntimeError("oracle_scale failed")
    return row, col, obj.value


def spmv(rp, ri, rx, v, rows):
    out = np.zeros(rows)
    rp = np.ascontiguousarray(rp, dtype=np.int64)
    ri = np.ascontiguousarray(ri, dtype=np.int32)
    rx = np.ascontiguousarray(rx, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    lib().oracle_spmv(rows, _p(rp), _p(ri), _p(rx), _p(v), _p(out))
    return out
