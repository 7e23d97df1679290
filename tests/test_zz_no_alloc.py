"""The allocation-free compensated residual (``sovlp_residual_f64_noalloc``)
against the workspace entry point (``sovlp_residual_f64``), through the C ABI:
bit for bit, on two calls in a row (the second smaller, so stale scratch
contents from the first would show), with long rows that go through the
scratch's list; sizes above the scratch's limit are refused."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr(rng, m, n, long_rows):
    lens = rng.integers(0, 12, m)
    lens[rng.choice(m, long_rows, replace=False)] = rng.integers(129, 4000, long_rows)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ri = rng.integers(0, n, rp[-1]).astype(np.int32)
    rx = rng.standard_normal(rp[-1]) * 10.0 ** rng.integers(-5, 6, rp[-1])
    return rp, ri, rx


def _run(fn, args):
    from sovlp import _lib
    _lib.check(fn(*args), "residual")


def test_noalloc_residual_equals_the_workspace_entry_point():
    import torch
    from sovlp import _lib
    L = _lib.lib()
    rng = np.random.default_rng(7)
    st = torch.cuda.current_stream().cuda_stream
    for m, n, nl in ((200_000, 50_000, 40), (3_000, 900, 7)):
        print(f"[gpu] no-alloc residual m={m} n={n} long rows {nl}", flush=True)
        rp, ri, rx = _csr(rng, m, n, nl)
        x = rng.standard_normal(n)
        b = rng.standard_normal(m)
        d = [torch.from_numpy(a).cuda() for a in (rp, ri, rx, x, b)]
        ptrs = [a.data_ptr() for a in d]
        ref = torch.empty(m, dtype=torch.float64, device="cuda")
        got = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
        nb = ctypes.c_int64()
        _lib.call("sovlp_residual_workspace_bytes", m, ctypes.byref(nb))
        ws = torch.empty(int(nb.value), dtype=torch.uint8, device="cuda")
        _run(L.sovlp_residual_f64, [m] + ptrs + [ref.data_ptr(), ws.data_ptr(), nb.value, st])
        _run(L.sovlp_residual_f64_noalloc, [m] + ptrs + [got.data_ptr(), st])
        torch.cuda.synchronize()
        a, g = ref.cpu().numpy(), got.cpu().numpy()
        assert np.array_equal(a.view(np.int64), g.view(np.int64))


def test_noalloc_residual_refuses_more_rows_than_its_scratch():
    from sovlp import _lib
    L = _lib.lib()
    cap = L.sovlp_residual_noalloc_max_rows()
    assert cap == 1 << 20
    rc = L.sovlp_residual_f64_noalloc(cap + 1, None, None, None, None, None, None, None)
    assert rc == 1
    assert b"at most 1048576 rows" in L.sovlp_last_error()
