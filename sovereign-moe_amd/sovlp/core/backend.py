"""Device backend: PyTorch-ROCm buffers plus the hand-written gfx950 kernels.

Mirrors the reference ``Backend`` (``src/sovopt/core/backend.py:317-551``):
``spmv``, ``spmm``, ``primal_step``, ``dual_step``, ``pdhg_primal``, ``pdhg_dual``,
``pdhg_accum``, ``dot``, ``nrm2``, ``nrm_inf``, ``to_device``, ``to_host``,
``index``, ``pointer``, ``zeros``, ``empty``, ``sync``, and the module-level
``get_backend``, ``gpu_available``, ``gpu_selftest`` and ``GPU_ERROR``.

Device arrays are ``torch`` tensors on ``cuda`` (HIP); every operation is a call
into ``libsovlp.so`` on torch's current stream. This engine runs on the GPU
only: ``Backend("cpu")`` and ``get_backend("cpu")`` raise, and ``"auto"``
raises when no usable GPU is present (``backend.py:538-551`` falls back to a
CPU backend instead; this package has none by design).
"""

from __future__ import annotations

import numpy as np

from .. import _lib

__all__ = ["Backend", "get_backend", "gpu_available", "gpu_selftest", "GPU_ERROR"]

GPU_ERROR: str | None = None

try:
    import torch as _torch
    _HAVE_GPU = bool(_torch.cuda.is_available())
    if not _HAVE_GPU:
        GPU_ERROR = "torch.cuda.is_available() is False (no HIP device visible)"
except Exception as _e:   # pragma: no cover - environment dependent
    _torch = None
    _HAVE_GPU = False
    GPU_ERROR = f"{type(_e).__name__}: {_e}"


def gpu_available() -> bool:
    """True when a HIP device is visible to PyTorch."""
    return _HAVE_GPU


_SELFTEST: tuple[bool, str | None] | None = None


def gpu_selftest() -> tuple[bool, str | None]:
    """Load ``libsovlp.so`` and check ``[[1,0,2],[0,3,0]]·[1,2,3] = [7,6]`` through
    every SpMV flavour (``backend.py:499-535``). Cached."""
    global _SELFTEST
    if _SELFTEST is not None:
        return _SELFTEST
    if not _HAVE_GPU:
        _SELFTEST = (False, GPU_ERROR)
        return _SELFTEST
    try:
        L = _lib.lib()
        _torch.zeros(1, device="cuda")        # create the context through torch
        rc = L.sovlp_selftest()
        if rc != 0:
            _SELFTEST = (False, L.sovlp_last_error().decode())
        else:
            _SELFTEST = (True, None)
    except Exception as e:   # pragma: no cover - environment dependent
        _SELFTEST = (False, f"{type(e).__name__}: {e}")
    return _SELFTEST


def _stream():
    return _torch.cuda.current_stream().cuda_stream


def _p(t):
    return t.data_ptr()


class Backend:
    """Array namespace plus sparse kernels on the GPU."""

    def __init__(self, device: str = "gpu", dtype=np.float64):
        if device not in ("cpu", "gpu"):
            raise ValueError("device must be 'cpu' or 'gpu'")
        if device == "cpu":
            raise ValueError("this engine runs on the GPU only; device='cpu' is not provided")
        ok, err = gpu_selftest()
        if not ok:
            raise RuntimeError(f"GPU backend unavailable ({err})")
        self.device = device
        self.dtype = np.dtype(dtype)
        self.xp = _torch
        self._tdtype = _torch.float64 if self.dtype == np.float64 else _torch.float32

    # -- transfer
    def to_device(self, a, dtype=None):
        dt = self.dtype if dtype is None else np.dtype(dtype)
        arr = np.ascontiguousarray(np.asarray(a), dtype=dt)
        return _torch.from_numpy(arr).to("cuda")

    def to_host(self, a):
        return a.detach().cpu().numpy()

    def index(self, a):
        return self.to_device(a, np.int32)

    def pointer(self, a):
        return self.to_device(a, np.int64)

    def sync(self):
        _torch.cuda.current_stream().synchronize()

    # -- allocation
    def zeros(self, shape, dtype=None):
        dt = self._tdtype if dtype is None else _torch.from_numpy(np.zeros(0, dtype)).dtype
        return _torch.zeros(shape, dtype=dt, device="cuda")

    def empty(self, shape, dtype=None):
        dt = self._tdtype if dtype is None else _torch.from_numpy(np.zeros(0, dtype)).dtype
        return _torch.empty(shape, dtype=dt, device="cuda")

    # -- sparse products
    @staticmethod
    def lanes_for(rows: int, nnz: int) -> int:
        """Lanes per row from the mean row length (a power of two in 1..64)."""
        mean = nnz / max(rows, 1)
        lanes = 1
        while lanes < 64 and lanes < mean:
            lanes *= 2
        return lanes

    def spmv(self, rp, ri, rx, x, out, m, lanes_per_row: int | None = None):
        """``out = A x`` with ``A`` in CSR (int64 pointers, int32 indices)."""
        lanes = lanes_per_row or self.lanes_for(m, rx.numel())
        fn = "sovlp_spmv_csr_f32" if rx.dtype == _torch.float32 else "sovlp_spmv_csr_f64"
        _lib.call(fn, int(m), _p(rp), _p(ri), _p(rx), _p(x), _p(out), lanes, _stream())
        return out

    def spmm(self, rp, ri, rx, X, out, m, k_batch):
        """``out = A X`` with ``X`` dense ``(n, k)``, batch index fastest
        (``backend.py:386-400``): the reference's CPU ``_csr_spmm`` bit for bit."""
        fn = "sovlp_spmm_csr_f32" if rx.dtype == _torch.float32 else "sovlp_spmm_csr_f64"
        _lib.call(fn, int(m), _p(rp), _p(ri), _p(rx), _p(X), _p(out), int(k_batch), _stream())
        return out

    # -- fused steps
    def primal_step(self, xnew, x, c, aty, lo, hi, tau, n):
        _lib.call("sovlp_primal_step_f64", _p(xnew), _p(x), _p(c), _p(aty), _p(lo), _p(hi),
                  float(tau), int(n), _stream())
        return xnew

    def dual_step(self, ynew, y, ax, rl, ru, sigma, m):
        _lib.call("sovlp_dual_step_f64", _p(ynew), _p(y), _p(ax), _p(rl), _p(ru),
                  float(sigma), int(m), _stream())
        return ynew

    def pdhg_primal(self, x, dx, c, aty, lo, hi, tau, n):
        _lib.call("sovlp_pdhg_primal_f64", _p(x), _p(dx), _p(c), _p(aty), _p(lo), _p(hi),
                  float(tau), int(n), _stream())

    def pdhg_dual(self, y, dy, ax, adx, rl, ru, sigma, m):
        _lib.call("sovlp_pdhg_dual_f64", _p(y), _p(dy), _p(ax), _p(adx), _p(rl), _p(ru),
                  float(sigma), int(m), _stream())

    def pdhg_accum(self, acc, inc, run_sum, v, eta, n):
        _lib.call("sovlp_pdhg_accum_f64", _p(acc), _p(inc), _p(run_sum), _p(v), float(eta),
                  int(n), _stream())

    # -- reductions (FP64, fixed order; one device scalar, then a host float)
    def _reduce(self, fn, *arrs):
        out = _torch.empty(1, dtype=_torch.float64, device="cuda")
        n = arrs[0].numel()
        _lib.call(fn, *[_p(a) for a in arrs], int(n), _p(out), _stream())
        return float(out.item())

    def dot(self, a, b) -> float:
        f32 = a.dtype == _torch.float32
        return self._reduce("sovlp_dot_f32" if f32 else "sovlp_dot_f64", a, b)

    def nrm2(self, a) -> float:
        f32 = a.dtype == _torch.float32
        return self._reduce("sovlp_nrm2_f32" if f32 else "sovlp_nrm2_f64", a)

    def nrm_inf(self, a) -> float:
        if a.numel() == 0:
            return 0.0
        return self._reduce("sovlp_nrminf_f64", a.to(_torch.float64))

    def __repr__(self):
        return f"Backend({self.device}, {self.dtype.name})"


def get_backend(device: str = "auto", dtype=np.float64) -> Backend:
    """``'gpu'`` or ``'auto'``: the GPU backend, or ``RuntimeError`` when it is not
    usable. ``'cpu'``: ``ValueError`` (no CPU engine in this package)."""
    if device == "auto":
        device = "gpu"
    if device == "gpu":
        ok, err = gpu_selftest()
        if not ok:
            raise RuntimeError(f"GPU backend unavailable ({err})")
    return Backend(device, dtype)
