"""Certificate and verifier of an LP point on the GPU.

Mirrors the reference's certificate path with the arithmetic on the device
(``sovlp_certify_f64`` / ``sovlp_residual_f64`` of ``include/sovlp.h``):

* ``compensated_residual(rp, ri, rx, x, b)`` -- ``b - A x`` to about
  double-double accuracy, bit for bit the reference's
  (``src/sovopt/numerics/refine.py:48-84, 103-110``).
* ``certified_bound(prob, x, y, dual_tol)`` -- the Neumaier-Shcherbina bound
  of the optimum from any dual vector, in the model's own sense, and the cost
  perturbation it needed (``src/sovopt/mip/safebound.py:252-307``).
* ``verify(prob, x, claimed_obj, ...)`` -- the independent verdict of
  ``bench/verify.py:55-113, 194-276``: dimension, finiteness, column bounds,
  compensated row activities, integrality, objective, and with duals the
  optimality gap against the certified bound.  ``Verdict.headline`` gives
  ACCEPTED / FEASIBLE / REJECTED in the reference's words.

``DeviceCertifier`` keeps one LP on the device for repeated certificates (a
serving loop certifies every report).  Linear programs only: a model with a
quadratic term is refused.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..core.problem import ObjSense


def _dev(a, dt):
    import torch
    from ..core.backend import get_backend
    get_backend("gpu")        # RuntimeError when no GPU or no library: no CPU path
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to("cuda")


def compensated_residual(rp, ri, rx, x, b, out=None) -> np.ndarray:
    """``b - A x`` (``A`` in CSR) to roughly double-double accuracy, on the GPU."""
    import torch
    rows = int(len(rp) - 1)
    if out is None:
        out = np.empty(rows, dtype=np.float64)
    if rows == 0:
        return out
    t = [_dev(rp, np.int64), _dev(ri, np.int32), _dev(rx, np.float64), _dev(x, np.float64),
         _dev(b, np.float64)]
    r = torch.empty(rows, dtype=torch.float64, device=t[0].device)
    nb = C.c_int64()
    _lib.call("sovlp_residual_workspace_bytes", rows, C.byref(nb))
    ws = torch.empty(int(nb.value), dtype=torch.uint8, device=t[0].device)
    stream = torch.cuda.current_stream(t[0].device).cuda_stream
    _lib.call("sovlp_residual_f64", rows, *[a.data_ptr() for a in t], r.data_ptr(),
              ws.data_ptr(), ws.numel(), stream)
    out[:] = r.cpu().numpy()
    return out


class DeviceCertifier:
    """One LP's original data on the device; ``certify(x, y)`` per point."""

    def __init__(self, prob):
        import torch
        if getattr(prob, "Q", None) is not None:
            raise NotImplementedError("certificates of quadratic models are not provided")
        A = prob.A
        self.prob = prob
        self.m, self.n = int(A.m), int(A.n)
        self.maximise = int(prob.sense) == int(ObjSense.MAXIMISE)
        self.offset = float(prob.obj_offset)
        self.t = {
            "rp": _dev(A.rp, np.int64), "ri": _dev(A.ri, np.int32), "rx": _dev(A.rx, np.float64),
            "cp": _dev(A.cp, np.int64), "ci": _dev(A.ci, np.int32), "cx": _dev(A.cx, np.float64),
            "c": _dev(prob.c, np.float64), "rl": _dev(prob.row_lb, np.float64),
            "ru": _dev(prob.row_ub, np.float64), "lo": _dev(prob.col_lb, np.float64),
            "hi": _dev(prob.col_ub, np.float64),
        }
        dev = self.t["c"].device
        nb = C.c_int64()
        _lib.call("sovlp_certify_workspace_bytes", self.m, self.n, C.byref(nb))
        self.ws = torch.empty(max(int(nb.value), 8), dtype=torch.uint8, device=dev)
        self.d = torch.empty(max(self.n, 1), dtype=torch.float64, device=dev)
        self.act = torch.empty(max(self.m, 1), dtype=torch.float64, device=dev)

    def certify(self, x, y, dual_tol: float = 1e-6) -> _lib.sovlp_certificate:
        """The certificate of (x, y); x and y are host arrays or device tensors."""
        import torch
        xs = x if isinstance(x, torch.Tensor) else _dev(x, np.float64)
        ys = y if isinstance(y, torch.Tensor) else _dev(y, np.float64)
        if xs.numel() != self.n or ys.numel() != self.m:
            raise ValueError(f"x has {xs.numel()} entries and y {ys.numel()}; "
                             f"the model has {self.n} columns and {self.m} rows")
        xs = xs.to(torch.float64).contiguous()
        ys = ys.to(torch.float64).contiguous()
        out = _lib.sovlp_certificate()
        t = self.t
        _lib.call("sovlp_certify_f64", self.m, self.n, t["rp"].data_ptr(), t["ri"].data_ptr(),
                  t["rx"].data_ptr(), t["cp"].data_ptr(), t["ci"].data_ptr(), t["cx"].data_ptr(),
                  t["c"].data_ptr(), t["rl"].data_ptr(), t["ru"].data_ptr(), t["lo"].data_ptr(),
                  t["hi"].data_ptr(), self.offset, int(self.maximise), xs.data_ptr(),
                  ys.data_ptr(), float(dual_tol), self.ws.data_ptr(), self.ws.numel(),
                  self.d.data_ptr(), self.act.data_ptr(), C.byref(out),
                  torch.cuda.current_stream(xs.device).cuda_stream)   # where xs, ys were made
        return out

    def reduced_costs(self) -> np.ndarray:
        """``c - Aᵀy`` of the last certificate (minimisation sense)."""
        return self.d[: self.n].cpu().numpy()

    def activities(self) -> np.ndarray:
        """The compensated ``A x`` of the last certificate."""
        return self.act[: self.m].cpu().numpy()


def certified_bound(prob, x, y, dual_tol: float = 1e-6) -> tuple[float, float]:
    """``(bound, perturbation)`` as the reference's ``certified_bound``: a lower
    bound on a minimisation's optimum (upper bound on a maximisation's) valid
    for any ``y`` (textbook signs); ``-inf`` (``+inf``) when vacuous."""
    cert = DeviceCertifier(prob).certify(x, y, dual_tol)
    return float(cert.bound), float(cert.perturbation)


class Verdict:
    """The checks of a point, as the reference's ``bench.verify.Verdict``."""

    def __init__(self):
        self.checks: list[tuple[str, bool, str]] = []
        self.certified_bound = None
        self.gap_rel = None
        self.feas_tol = None
        self.opt_tol = None

    def add(self, name, ok, detail=""):
        self.checks.append((name, bool(ok), detail))

    @property
    def ok(self) -> bool:
        return all(c[1] for c in self.checks)

    def headline(self, extra=None):
        """``(verdict, detail)``: ACCEPTED, FEASIBLE or REJECTED (verify.py:70-103)."""
        checks = list(self.checks) + ([extra] if extra is not None else [])
        failed = [c for c in checks if not c[1]]
        hard = [c for c in failed if c[0] != "optimality"]
        if hard:
            return "REJECTED", "; ".join(f"{name}: {detail}" for name, _ok, detail in hard)
        feas = f"feasible to {self.feas_tol:.0e}" if self.feas_tol else "feasible"
        opt = f"{self.opt_tol:.0e}" if self.opt_tol else "the certificate's tolerance"
        if not failed:
            if any(c[0] == "optimality" for c in checks):
                return "ACCEPTED", f"{feas}, and certified optimal to {opt}"
            return "ACCEPTED", f"{feas}; optimality not checked"
        gap = self.gap_rel
        if gap is None or not np.isfinite(gap):
            return "FEASIBLE", f"{feas}; its duals certify no bound"
        if gap < 0:
            return "FEASIBLE", (f"{feas}; the objective sits {-gap:.1e} past the certified "
                                f"bound -- rounding at the bound's resolution -- so it is "
                                f"not certified at {opt}")
        return "FEASIBLE", f"{feas}, and optimal to {gap:.1e} -- short of the {opt} certificate"

    def report(self) -> str:
        w = max(len(c[0]) for c in self.checks) if self.checks else 10
        lines = [f"  [{'PASS' if ok else 'FAIL'}] {name:<{w}}  {detail}"
                 for name, ok, detail in self.checks]
        lines.append("")
        lines.append(f"  VERDICT: {'ACCEPTED' if self.ok else 'REJECTED'}")
        return "\n".join(lines)


def verify(prob, x, claimed_obj=None, feas_tol=1e-6, int_tol=1e-6, y=None, opt_tol=1e-9,
           certifier: DeviceCertifier | None = None) -> Verdict:
    """The reference's ``verify`` (bench/verify.py:194-276) with the device
    certificate.  ``certifier``: a ``DeviceCertifier`` of ``prob`` to reuse."""
    v = Verdict()
    v.feas_tol, v.opt_tol = feas_tol, opt_tol
    x = np.asarray(x, dtype=np.float64)
    if x.shape[0] != prob.n:
        v.add("dimension", False, f"got {x.shape[0]}, model has {prob.n} columns")
        return v
    v.add("dimension", True, f"{prob.n} columns")
    if not np.isfinite(x).all():
        v.add("finite", False, "solution contains NaN or Inf")
        return v
    v.add("finite", True, "")

    mask = prob.integer_mask
    with_duals = y is not None and not mask.any()
    yv = np.zeros(prob.m)
    dual_ok = True
    if with_duals:
        yv = np.asarray(y, dtype=np.float64)
        dual_ok = yv.shape[0] == prob.m and bool(np.isfinite(yv).all())
        if not dual_ok:
            yv = np.zeros(prob.m)
    cz = certifier or DeviceCertifier(prob)
    cert = cz.certify(x, yv)

    bnd = float(cert.col_violation)
    v.add("column bounds", bnd <= feas_tol, f"max violation {bnd:.3e}")
    rowv = float(cert.row_violation)
    v.add("row constraints", rowv <= feas_tol, f"max violation {rowv:.3e}")
    if mask.any():
        iv = float(np.abs(x[mask] - np.round(x[mask])).max(initial=0.0))
        v.add("integrality", iv <= int_tol,
              f"max distance to integer {iv:.3e} over {int(mask.sum())} vars")
    else:
        v.add("integrality", True, "no integer variables")
    obj = float(cert.objective)
    if claimed_obj is not None:
        d = abs(obj - claimed_obj) / max(1.0, abs(obj))
        v.add("objective", d <= 1e-6,
              f"recomputed {obj:.12g}, claimed {claimed_obj:.12g}, rel diff {d:.2e}")
    else:
        v.add("objective", True, f"recomputed {obj:.12g}")

    if with_duals:
        if not dual_ok:
            v.add("optimality", False, "dual vector has the wrong length or is not finite")
            return v
        bound, pert = float(cert.bound), float(cert.perturbation)
        v.certified_bound = bound
        if not np.isfinite(bound):
            v.gap_rel = float("inf")
            v.add("optimality", False,
                  f"the duals certify nothing: a reduced cost of {pert:.2e} "
                  f"points at an infinite column bound")
        else:
            minimise = int(prob.sense) != int(ObjSense.MAXIMISE)
            gap = (obj - bound) if minimise else (bound - obj)
            rel = gap / max(1.0, abs(obj))
            v.gap_rel = rel
            if rel < -opt_tol:
                v.add("optimality", False,
                      f"objective {gap:.3e} BELOW the certified bound {bound:.12g} "
                      f"({-rel:.2e} relative): the point is infeasible at the "
                      f"bound's resolution, not optimal")
            else:
                v.add("optimality", rel <= opt_tol,
                      f"certified bound {bound:.12g}, gap {gap:.3e} ({rel:.2e} relative)"
                      + (f", costs perturbed by {pert:.1e}" if pert else ""))
    return v
