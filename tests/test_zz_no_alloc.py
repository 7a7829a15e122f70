"""The allocation-free triangular solve, mom_getrs_noalloc: the solve of
mom_getrs with its workspace in the library's static device array instead of
a caller buffer.  Checked through the C-ABI against mom_getrs on the same
factors: the two calls run the same kernels in the same summation order, so
the solutions must be bit-identical; two calls in a row, the second on a
smaller system, show that the static workspace carries nothing over."""

import numpy as np
import pytest
import torch

from mom25d_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"
MAX_N = 16384  # the static workspace's limit (csrc/lu.hip kNoAllocMaxN)


def _noalloc():
    return _lib.lib().mom_getrs_noalloc  # bound with its signature by mom25d_amd._lib


def _factor(prec, n, seed):
    rng = np.random.default_rng(seed)
    dt = torch.complex128 if prec == 64 else torch.complex64
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Ad = torch.from_numpy(A).to(DEV, dt)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    P = _lib.ptr
    s = _lib.stream_ptr()
    _lib.call("mom_getrf", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    torch.cuda.synchronize()
    assert int(info[0]) == 0 and int(info[1]) == 0
    return A, Ad, perm


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_solve_equals_getrs_twice(prec):
    """Two calls in a row (n = 3001 with 3 right-hand sides, then n = 700
    with 2), each bit-identical to mom_getrs on the same factors, with the
    caller's status word left at zero."""
    f = _noalloc()
    P = _lib.ptr
    s = _lib.stream_ptr()
    for n, nrhs, seed in ((3001, 3, 1), (700, 2, 2)):
        A, Ad, perm = _factor(prec, n, seed)
        rng = np.random.default_rng(seed + 100)
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        Bd = torch.from_numpy(B).to(DEV, Ad.dtype)
        X0 = torch.empty_like(Bd)
        sw = torch.zeros(int(_lib.lib().mom_getrs_work_bytes(prec, n)), dtype=torch.uint8, device=DEV)
        _lib.call("mom_getrs", prec, n, nrhs, P(Ad), n, P(perm), P(Bd), P(X0), P(sw), s)
        X1 = torch.full_like(Bd, float("nan"))
        status = torch.zeros(1, dtype=torch.int32, device=DEV)
        rc = f(prec, n, nrhs, P(Ad), n, P(perm), P(Bd), P(X1), P(status), s)
        assert rc == 0, _lib.lib().mom_last_error().decode()
        torch.cuda.synchronize()
        assert int(status[0]) == 0
        x0 = X0.cpu().numpy()
        x1 = X1.cpu().numpy()
        assert np.array_equal(x0.view(np.uint8), x1.view(np.uint8)), (n, np.max(np.abs(x1 - x0)))
        if prec == 64:  # and it is the solution (backward error at round-off)
            r = np.linalg.norm(A @ x1 - B) / (np.linalg.norm(A) * np.linalg.norm(x1))
            assert r < 1e-14


def test_noalloc_rejects_oversize():
    """A system larger than the static workspace is refused with a clear
    error before anything is launched."""
    f = _noalloc()
    d = torch.zeros(4, dtype=torch.complex128, device=DEV)
    p = torch.zeros(4, dtype=torch.int32, device=DEV)
    st = torch.zeros(1, dtype=torch.int32, device=DEV)
    P = _lib.ptr
    rc = f(64, MAX_N + 1, 1, P(d), MAX_N + 1, P(p), P(d), P(d[2:]), P(st), _lib.stream_ptr())
    assert rc != 0
    assert "static workspace" in _lib.lib().mom_last_error().decode()
