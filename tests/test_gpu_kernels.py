"""GPU unit tests of the building blocks: Bessel J0, the complex GEMM on the
matrix cores, LU + triangular solves (complex128 / complex64)."""

import numpy as np
import pytest
import torch
from scipy.special import j0 as sp_j0
from scipy.special import jv

from mom25d_amd import _lib

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _j0_dev(z, real_axis):
    zd = torch.from_numpy(np.ascontiguousarray(z, complex)).to(DEV)
    out = torch.empty_like(zd)
    _lib.call("mom_j0", z.size, _lib.ptr(zd), _lib.ptr(out), int(real_axis), _lib.stream_ptr())
    return out.cpu().numpy()


def test_j0_complex_matches_scipy():
    rng = np.random.default_rng(0)
    x = np.concatenate([np.linspace(0, 120, 20000), rng.uniform(0, 3, 2000), [0.0, 3.999, 4.0, 19.999, 20.0]])
    y = np.concatenate([rng.uniform(0, 1.1, x.size - 5), [0.0, 0.3, 1.0, 0.7, 0.01]])
    z = x + 1j * y
    got = _j0_dev(z, False)
    ref = jv(0, z)
    amp = np.sqrt(2 / (np.pi * np.maximum(np.abs(z), 1))) * np.cosh(y)
    assert np.max(np.abs(got - ref) / amp) < 5e-14


def test_j0_real_matches_scipy():
    """Below x = 100 both agree to 5e-14 of the amplitude; beyond, scipy's
    j0 rounds its phase x - pi/4 (error ~ x ulp), which bounds the agreement."""
    x = np.concatenate([np.linspace(0, 30, 30001), np.linspace(30, 6000, 200001)])
    got = _j0_dev(x.astype(complex), True).real
    ref = sp_j0(x)
    amp = np.sqrt(2 / (np.pi * np.maximum(x, 1)))
    err = np.abs(got - ref) / amp
    assert np.max(err[x < 100]) < 5e-14
    assert np.all(err <= 5e-14 + 4 * x * np.finfo(float).eps)


def test_j0_matches_mpmath():
    """Independent high-precision check (50 digits) over all three branches."""
    import mpmath

    mpmath.mp.dps = 50
    rng = np.random.default_rng(3)
    r = np.concatenate([rng.uniform(0, 4, 300), rng.uniform(4, 20, 300), rng.uniform(20, 150, 300),
                        rng.uniform(150, 6000, 100)])
    y = np.concatenate([rng.uniform(0, 1.1, 900), np.zeros(100)])
    z = r + 1j * y
    got = _j0_dev(z, False)
    ref = np.array([complex(mpmath.besselj(0, mpmath.mpc(a.real, a.imag))) for a in z])
    amp = np.sqrt(2 / (np.pi * np.maximum(np.abs(z), 1))) * np.cosh(y)
    assert np.max(np.abs(got - ref) / amp) < 1e-14
    xr = r[-100:]
    gotr = _j0_dev(xr.astype(complex), True).real
    refr = np.array([float(mpmath.besselj(0, mpmath.mpf(a))) for a in xr])
    assert np.max(np.abs(gotr - refr) / np.sqrt(2 / (np.pi * xr))) < 1e-14


@pytest.mark.parametrize("tile", [(64, 64), (64, 128), (128, 64), (128, 128)])
@pytest.mark.parametrize("prec,tol", [(64, 1e-13), (32, 2e-5)])
@pytest.mark.parametrize("m,n,k", [(64, 32, 64), (130, 77, 40), (257, 300, 64), (1000, 513, 128)])
def test_cgemm_sub(prec, tol, m, n, k, tile):
    from mom25d_amd.autotune import get_tile, set_tile

    rng = np.random.default_rng(m + n + k)
    dt = torch.complex128 if prec == 64 else torch.complex64
    A = rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    C = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    Ad, Bd, Cd = (torch.from_numpy(a).to(DEV, dt) for a in (A, B, C))
    old = get_tile(prec)
    set_tile(prec, tile)
    try:
        _lib.call("mom_cgemm_sub", prec, m, n, k, _lib.ptr(Ad), k, _lib.ptr(Bd), n, _lib.ptr(Cd), n,
                  _lib.stream_ptr())
        got = Cd.cpu().numpy()
    finally:
        set_tile(prec, old)
    ref = C - A @ B
    assert np.max(np.abs(got - ref)) < tol * np.max(np.abs(ref)) * np.sqrt(k)


def test_autotuned_lu_solves_like_the_default():
    """The autotuner times every panel width / tile and returns the fastest;
    the factorisation with that choice solves a system like the default."""
    from mom25d_amd.autotune import get_tile, set_tile, tune_lu

    best = tune_lu(700, 64, reps=1)
    print(f"autotune n={best['n']} complex128:", best["jb"], best["nb"], best["tile"], f"{best['ms']:.2f} ms",
          len(best["table"]), "runs")
    assert len(best["table"]) == 16 and best["ms"] == min(r[3] for r in best["table"])
    rng = np.random.default_rng(5)
    n = 700
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    from mom25d_amd.autotune import get_block, set_block

    old, old_nb = get_tile(64), get_block(64)
    set_tile(64, best["tile"])
    set_block(64, best["nb"])
    try:
        X, _, info = _lu_solve(A, B, 64, best["jb"])
    finally:
        set_tile(64, old)
        set_block(64, old_nb)
    assert info == 0
    assert np.linalg.norm(A @ X - B) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(X)


def _lu_solve(A, B, prec, jb, status=False):
    n = A.shape[0]
    dt = torch.complex128 if prec == 64 else torch.complex64
    Ad = torch.from_numpy(A).to(DEV, dt)
    Bd = torch.from_numpy(B).to(DEV, dt)
    X = torch.empty_like(Bd)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, jb))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    sw = torch.zeros(int(_lib.lib().mom_getrs_work_bytes(prec, n)), dtype=torch.uint8, device=DEV)
    s = _lib.stream_ptr()
    P = _lib.ptr
    _lib.call("mom_getrf", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, jb, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    _lib.call("mom_getrs", prec, n, B.shape[1], P(Ad), n, P(perm), P(Bd), P(X), P(sw), s)
    if status:
        return X.cpu().numpy(), info.cpu().numpy(), sw[:20].cpu().numpy().view(np.uint32)
    return X.cpu().numpy(), ipiv.cpu().numpy(), int(info[0])


@pytest.mark.parametrize("prec,n,jb", [(32, 1, 8), (32, 200, 32), (32, 777, 64), (32, 2049, 128), (64, 300, 128),
                                       (64, 1029, 64), (64, 130, 8)])
def test_lu_nopiv_reconstructs_the_matrix(prec, n, jb):
    """The LU without row interchanges: identity ipiv, no zero pivot, and
    L U = A to the precision's round-off (a diagonally dominant matrix, whose
    unpivoted factors exist and are well conditioned), over panels that
    leave partial blocks (n not a multiple of jb, complex128 panels of 64
    inside 128-wide blocks)."""
    rng = np.random.default_rng(n + jb)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    dt = torch.complex128 if prec == 64 else torch.complex64
    Ad = torch.from_numpy(A).to(DEV, dt)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.full((4,), 7, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, jb))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    _lib.call("mom_getrf_nopiv", prec, n, _lib.ptr(Ad), n, _lib.ptr(ipiv), _lib.ptr(info), _lib.ptr(work), wb, jb,
              _lib.stream_ptr())
    LU = Ad.cpu().numpy().astype(np.complex128)
    assert int(info[0]) == 0 and int(info[1]) == 0
    np.testing.assert_array_equal(ipiv.cpu().numpy(), np.arange(n))
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    Ar = Ad.new_tensor(A).cpu().numpy().astype(np.complex128)  # A in the factor's precision
    err = np.linalg.norm(L @ U - Ar) / np.linalg.norm(Ar)
    assert err < (3e-6 if prec == 32 else 1e-14), err


@pytest.mark.parametrize("prec,n,jb", [(32, 1, 8), (32, 200, 32), (32, 777, 64), (32, 2049, 128), (32, 3000, 64),
                                       (64, 300, 128), (64, 1029, 64), (64, 130, 8)])
def test_ldlt_nopiv_factors(prec, n, jb):
    """The complex-symmetric factorisation A = L D L^T stored as L and
    U = D L^T: only the lower triangle is read (the upper one holds large
    garbage that must not reach the factors), identity ipiv, U equal to the
    pivots times L^T, L U = A to the precision's round-off, and the same
    factors as the general unpivoted LU of the symmetric matrix up to
    rounding (the two compute the same quantities in different orders)."""
    rng = np.random.default_rng(3 * n + jb)
    G = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = G + G.T + 4 * np.sqrt(n) * np.eye(n)
    dt = torch.complex128 if prec == 64 else torch.complex64
    Ar = torch.from_numpy(A).to(dt).numpy().astype(np.complex128)  # A in the factor's precision
    junk = 1e6 * (rng.standard_normal((n, n)) + 1j)
    Ain = np.tril(A) + np.triu(junk, 1)
    P = _lib.ptr
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, jb))
    out = {}
    for entry, M in (("mom_sytrf_nopiv", Ain), ("mom_getrf_nopiv", A)):
        Ad = torch.from_numpy(M).to(DEV, dt)
        ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
        info = torch.full((4,), 7, dtype=torch.int32, device=DEV)
        work = torch.empty(wb, dtype=torch.uint8, device=DEV)
        _lib.call(entry, prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, jb, _lib.stream_ptr())
        assert int(info[0]) == 0 and int(info[1]) == 0, entry
        np.testing.assert_array_equal(ipiv.cpu().numpy(), np.arange(n))
        out[entry] = Ad.cpu().numpy().astype(np.complex128)
    LU = out["mom_sytrf_nopiv"]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    eps = 6e-8 if prec == 32 else 1.2e-16
    # outside the diagonal blocks U is d_i L_ji itself; inside them it comes from the block's elimination,
    # which agrees to a few units of round-off of U's off-diagonal size
    dU = np.abs(np.triu(U, 1) - np.triu(np.diag(U)[:, None] * L.T, 1)).max()
    assert dU <= 32 * eps * max(np.abs(np.triu(U, 1)).max(), 1e-300), dU
    err = np.linalg.norm(L @ U - Ar) / np.linalg.norm(Ar)
    assert err < (3e-6 if prec == 32 else 1e-14), err
    gen = out["mom_getrf_nopiv"]
    assert np.max(np.abs(LU - gen)) < 1e3 * eps * np.max(np.abs(gen))


def test_lu_nopiv_reports_zero_pivot():
    """A zero leading entry: the unpivoted LU reports pivot 1 in info[0] and
    leaves finite factors (the zero pivot is replaced by 1), no fault."""
    A = np.array([[0, 1, 2], [1, 0, 1], [2, 1, 5]], complex)
    Ad = torch.from_numpy(A).to(DEV, torch.complex64)
    ipiv = torch.empty(3, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(32, 3, 8))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    _lib.call("mom_getrf_nopiv", 32, 3, _lib.ptr(Ad), 3, _lib.ptr(ipiv), _lib.ptr(info), _lib.ptr(work), wb, 8,
              _lib.stream_ptr())
    assert int(info[0]) == 1
    assert np.all(np.isfinite(Ad.cpu().numpy()))


@pytest.mark.parametrize("n,jb,nrhs", [(1, 8, 1), (37, 8, 3), (200, 32, 2), (777, 64, 4), (1500, 64, 2),
                                       (2049, 64, 1), (129, 32, 6), (300, 32, 8)])
def test_lu_fp64_matches_lapack(n, jb, nrhs):
    import scipy.linalg as sla

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    X, ipiv, info = _lu_solve(A, B, 64, jb)
    assert info == 0
    ref = np.linalg.solve(A, B)
    assert np.max(np.abs(X - ref)) < 1e-10 * np.max(np.abs(ref))
    # same pivots as LAPACK zgetrf (|re| + |im| first maximum); rounding of
    # the updated columns may split a near-tie differently on large matrices
    _, piv = sla.lu_factor(A)
    assert np.mean(ipiv == piv) > (0.999 if n > 200 else 1.0 - 1e-12)


@pytest.mark.parametrize("n,jb,nb", [(777, 64, 256), (1500, 32, 128), (2049, 64, 128), (300, 128, 256)])
def test_blocked_lu_matches_lapack(n, jb, nb):
    """Blocks of nb columns factored as panels of jb (two-level blocking):
    the solution and the pivot sequence of zgetrf."""
    import scipy.linalg as sla

    from mom25d_amd.autotune import get_block, set_block

    rng = np.random.default_rng(n + nb)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))
    old = get_block(64)
    set_block(64, nb)
    try:
        X, ipiv, info = _lu_solve(A, B, 64, jb)
    finally:
        set_block(64, old)
    assert info == 0
    ref = np.linalg.solve(A, B)
    assert np.max(np.abs(X - ref)) < 1e-10 * np.max(np.abs(ref))
    _, piv = sla.lu_factor(A)
    assert np.mean(ipiv == piv) > 0.999


@pytest.mark.parametrize("n", [300, 1024])
def test_lu_fp32_backward_error(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    X, _, info = _lu_solve(A, B, 32, 64)
    assert info == 0
    r = np.linalg.norm(A @ X - B) / (np.linalg.norm(A, 2) * np.linalg.norm(X))
    assert r < 1e-5


def test_bounded_waits_report_failure():
    """With a wait limit of one poll the panel's workgroups and the solve
    chain give up instead of waiting: every kernel still ends, the calls flag
    the failure, and with the default limit back the next factorisation is
    correct (the failure path forced from the host; no kernel hangs)."""
    rng = np.random.default_rng(11)
    n = 3000
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    _lib.call("mom_set_spin_limit", 1)
    try:
        _, info, sw = _lu_solve(A, B, 64, 64, status=True)
    finally:
        _lib.call("mom_set_spin_limit", 1 << 21)
    assert info[1] == 1 or sw[4] == 1
    X, _, info0 = _lu_solve(A, B, 64, 64)
    assert info0 == 0
    assert np.linalg.norm(A @ X - B) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(X)


def test_solve_failure_word_is_sticky():
    """A solve whose wait ran out sets the failure word of its workspace; a
    later solve with the same workspace that succeeds does not clear it (a
    GMRES solve runs hundreds of preconditioner solves before its one check),
    and the caller's own clear does."""
    rng = np.random.default_rng(12)
    n = 3000
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    B = rng.standard_normal((n, 1)) + 1j * rng.standard_normal((n, 1))
    Ad = torch.from_numpy(A).to(DEV)
    Bd = torch.from_numpy(B).to(DEV)
    X = torch.empty_like(Bd)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(64, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    sw = torch.zeros(int(_lib.lib().mom_getrs_work_bytes(64, n)), dtype=torch.uint8, device=DEV)
    s = _lib.stream_ptr()
    P = _lib.ptr
    _lib.call("mom_getrf", 64, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    torch.cuda.synchronize()
    assert int(info[1]) == 0
    _lib.call("mom_set_spin_limit", 1)
    try:
        _lib.call("mom_getrs", 64, n, 1, P(Ad), n, P(perm), P(Bd), P(X), P(sw), s)
        torch.cuda.synchronize()
    finally:
        _lib.call("mom_set_spin_limit", 1 << 21)
    if sw[:20].cpu().numpy().view(np.uint32)[4] != 1:
        pytest.skip("every block of the one-poll solve found its dependency at once (nothing to check)")
    _lib.call("mom_getrs", 64, n, 1, P(Ad), n, P(perm), P(Bd), P(X), P(sw), s)
    torch.cuda.synchronize()
    assert sw[:20].cpu().numpy().view(np.uint32)[4] == 1  # the successful solve left the failure standing
    Xh = X.cpu().numpy()
    assert np.linalg.norm(A @ Xh - B) <= 1e-12 * np.linalg.norm(A) * np.linalg.norm(Xh)
    sw[16:20].zero_()
    _lib.call("mom_getrs", 64, n, 1, P(Ad), n, P(perm), P(Bd), P(X), P(sw), s)
    torch.cuda.synchronize()
    assert sw[:20].cpu().numpy().view(np.uint32)[4] == 0


@pytest.mark.parametrize("prec,n,nrhs,rows", [(32, 1, 1, 64), (32, 300, 2, 64), (32, 2049, 4, 64), (64, 777, 3, 64),
                                            (32, 1, 1, 128), (32, 300, 2, 128), (32, 2049, 4, 128), (32, 130, 3, 128),
                                            (32, 1000, 8, 64), (64, 500, 5, 64), (32, 700, 7, 128)])
def test_getrs_with_inverted_blocks(prec, n, nrhs, rows):
    """The solve with explicitly inverted diagonal blocks (64 rows, or in
    complex64 128 rows composed from them): the same solution as the
    substitution solve up to cond(block) rounding, and in complex128 a
    backward error near round-off (a diagonally dominant matrix, so the
    blocks are well conditioned); the inverses are reused by a second solve
    with new right-hand sides.  Up to 8 right-hand sides; with more than 4
    the 128-row setting solves on 64-row blocks."""
    _lib.call("mom_set_solve_rows", rows)
    try:
        _getrs_inv_case(prec, n, nrhs)
    finally:
        _lib.call("mom_set_solve_rows", 64)


@pytest.mark.parametrize("prec,n,nrhs,pivot", [(32, 1, 1, False), (32, 700, 2, False), (32, 1024, 3, False),
                                               (32, 1025, 4, False), (32, 3001, 2, False), (64, 1500, 2, False),
                                               (64, 2048, 1, True), (32, 2600, 2, True), (64, 513, 4, False),
                                               (32, 2600, 3, "ldlt"), (64, 1100, 2, "ldlt")])
def test_superblock_solve(prec, n, nrhs, pivot):
    """The super-block solve (inverses of the 512-row diagonal blocks of L
    and U, couplings scaled by them, one chain of super-blocks per factor):
    backward error near the factor precision's round-off on a diagonally
    dominant matrix, the solution of the 64-row chain up to cond(block)
    rounding, partial last super-blocks, pivoted and unpivoted factors, and
    a second solve with the same set-up; pivot="ldlt": the L D L^T factors
    of a complex-symmetric matrix."""
    rng = np.random.default_rng(n + 10 * nrhs)
    dt = torch.complex128 if prec == 64 else torch.complex64
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    if pivot == "ldlt":
        A = A + A.T
    Ad = torch.from_numpy(A).to(DEV, dt)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    s = _lib.stream_ptr()
    P = _lib.ptr
    if pivot == "ldlt":
        _lib.call("mom_sytrf_nopiv", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, s)
    elif pivot:
        _lib.call("mom_getrf", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    else:
        _lib.call("mom_getrf_nopiv", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    sbb = int(_lib.lib().mom_sbsolve_work_bytes(prec, n, 512))
    sbw = torch.zeros(sbb, dtype=torch.uint8, device=DEV)
    _lib.call("mom_sbsolve_setup", prec, n, P(Ad), n, 512, P(sbw), sbb, s)
    inv = torch.zeros(int(_lib.lib().mom_getrs_inv_work_bytes(prec, n)), dtype=torch.uint8, device=DEV)
    _lib.call("mom_trtri_blocks", prec, n, P(Ad), n, P(inv), s)
    for rep in range(2):
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        Bd = torch.from_numpy(B).to(DEV)
        X = torch.full_like(Bd, float("nan"))
        _lib.call("mom_sbsolve_z", prec, n, nrhs, 512, P(perm), P(Bd), P(X), P(sbw), s)
        X64 = torch.empty_like(Bd)
        _lib.call("mom_getrs_inv_z", prec, n, nrhs, P(Ad), n, P(perm), P(Bd), P(X64), P(inv), s)
        got, ref = X.cpu().numpy(), X64.cpu().numpy()
        assert int(sbw[16:20].view(torch.int32).cpu()[0]) == 0
        assert np.all(np.isfinite(got))
        r = np.linalg.norm(A @ got - B) / (np.linalg.norm(A, 2) * np.linalg.norm(got))
        assert r < (1e-6 if prec == 32 else 1e-14), (rep, r)
        assert np.max(np.abs(got - ref)) < (2e-5 if prec == 32 else 1e-12) * np.max(np.abs(ref)), rep


@pytest.mark.parametrize("prec,n,nrhs", [(32, 2049, 2), (64, 777, 3), (32, 1000, 8), (32, 65, 1), (64, 130, 6)])
def test_getrs_inv_chain_and_plain_forms(prec, n, nrhs):
    """Both forms of the 64-row inverted-block solve -- the chain form
    x_k = y_k - M_k x_prev (y_k formed before the previous block arrives)
    and the plain x_k = D_k^-1 (b_k - all couplings) -- meet the bars of the
    substitution solve, with partial last blocks (65 = 64 + 1 rows)."""
    try:
        for chain in (0, 1):
            _lib.call("mom_set_solve_chain", chain)
            assert _lib.lib().mom_get_solve_chain() == chain
            _getrs_inv_case(prec, n, nrhs)
    finally:
        _lib.call("mom_set_solve_chain", 1)


def _getrs_inv_case(prec, n, nrhs):
    rng = np.random.default_rng(n + nrhs)
    dt = torch.complex128 if prec == 64 else torch.complex64
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    Ad = torch.from_numpy(A).to(DEV, dt)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    sw = torch.zeros(int(_lib.lib().mom_getrs_inv_work_bytes(prec, n)), dtype=torch.uint8, device=DEV)
    assert sw.numel() > int(_lib.lib().mom_getrs_work_bytes(prec, n))
    s = _lib.stream_ptr()
    P = _lib.ptr
    _lib.call("mom_getrf", prec, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    _lib.call("mom_trtri_blocks", prec, n, P(Ad), n, P(sw), s)
    for rep in range(2):
        B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        Bd = torch.from_numpy(B).to(DEV, dt)
        X = torch.empty_like(Bd)
        _lib.call("mom_getrs_inv", prec, n, nrhs, P(Ad), n, P(perm), P(Bd), P(X), P(sw), s)
        got = X.cpu().numpy()
        assert int(sw[16:20].view(torch.int32).cpu()[0]) == 0
        r = np.linalg.norm(A @ got - B) / (np.linalg.norm(A, 2) * np.linalg.norm(got))
        assert r < (1e-6 if prec == 32 else 1e-14), (rep, r)
        Xs = torch.empty_like(Bd)
        sw2 = torch.zeros(int(_lib.lib().mom_getrs_work_bytes(prec, n)), dtype=torch.uint8, device=DEV)
        _lib.call("mom_getrs", prec, n, nrhs, P(Ad), n, P(perm), P(Bd), P(Xs), P(sw2), s)
        ref = Xs.cpu().numpy()
        assert np.max(np.abs(got - ref)) < (1e-5 if prec == 32 else 1e-12) * np.max(np.abs(ref))


@pytest.mark.parametrize("n,nrhs,rowsum", [(1, 1, False), (301, 3, True), (1000, 2, True), (517, 4, False),
                                            (64, 1, True)])
def test_residual_matches_numpy(n, nrhs, rowsum):
    """R = B - Z X in FP64 (the refinement residual), and the row sums
    d_r sum_k |z_rk|_1 d_k of the diagonally scaled matrix."""
    rng = np.random.default_rng(n * 7 + nrhs)
    Z = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    X = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    d = rng.uniform(0.5, 2.0, n)
    Zd, Xd, Bd = (torch.from_numpy(a).to(DEV) for a in (Z, X, B))
    R = torch.empty_like(Bd)
    rs = torch.empty(n, dtype=torch.float64, device=DEV) if rowsum else None
    dd = torch.from_numpy(d).to(DEV) if rowsum else None
    _lib.call("mom_residual", n, nrhs, _lib.ptr(Zd), n, _lib.ptr(Xd), _lib.ptr(Bd), _lib.ptr(R), _lib.ptr(rs),
              _lib.ptr(dd), _lib.stream_ptr())
    ref = B - Z @ X
    assert np.max(np.abs(R.cpu().numpy() - ref)) < 1e-13 * np.sqrt(n) * np.max(np.abs(Z)) * np.max(np.abs(X))
    if rowsum:
        rref = d * ((np.abs(Z.real) + np.abs(Z.imag)) @ d)
        np.testing.assert_allclose(rs.cpu().numpy(), rref, rtol=1e-13)


@pytest.mark.parametrize("n,nrhs", [(1, 1), (333, 2), (1500, 4)])
def test_getrs_inv_z_and_scaled_matvec(n, nrhs):
    """The GMRES building blocks: the complex64 LU solve taking and giving
    complex128 vectors equals the complex64 solve of the rounded vector,
    widened; and W = D Z D X equals the numpy product."""
    rng = np.random.default_rng(n + 17)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + 4 * np.sqrt(n) * np.eye(n)
    Ad = torch.from_numpy(A).to(DEV, torch.complex64)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(32, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    sw = torch.zeros(int(_lib.lib().mom_getrs_inv_work_bytes(32, n)), dtype=torch.uint8, device=DEV)
    s = _lib.stream_ptr()
    P = _lib.ptr
    _lib.call("mom_getrf", 32, n, P(Ad), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    _lib.call("mom_trtri_blocks", 32, n, P(Ad), n, P(sw), s)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    B64 = torch.from_numpy(B).to(DEV)
    X64 = torch.empty_like(B64)
    _lib.call("mom_getrs_inv_z", 32, n, nrhs, P(Ad), n, P(perm), P(B64), P(X64), P(sw), s)
    B32 = B64.to(torch.complex64)
    X32 = torch.empty_like(B32)
    _lib.call("mom_getrs_inv", 32, n, nrhs, P(Ad), n, P(perm), P(B32), P(X32), P(sw), s)
    np.testing.assert_array_equal(X64.cpu().numpy(), X32.cpu().numpy().astype(complex))
    d = rng.uniform(0.5, 2.0, n)
    Zd = torch.from_numpy(A).to(DEV)
    W = torch.empty_like(B64)
    _lib.call("mom_scaled_matvec", n, nrhs, P(Zd), n, P(torch.from_numpy(d).to(DEV)), P(B64), P(W), s)
    ref = d[:, None] * (A @ (d[:, None] * B))
    assert np.max(np.abs(W.cpu().numpy() - ref)) < 1e-13 * np.sqrt(n) * np.max(np.abs(ref))


@pytest.mark.parametrize("n,nrhs,ld", [(1, 1, 1), (63, 2, 63), (64, 3, 70), (65, 2, 66), (333, 4, 333),
                                       (1500, 2, 1500), (2049, 1, 2052)])
def test_scaled_symv_matches_full_product(n, nrhs, ld):
    """W = D Z D X from the tiles on and above the diagonal of a symmetric Z
    equals the numpy product and the row-wise kernel over every row (partial
    last tiles, leading dimensions past n, 1 to 4 right-hand sides in passes
    of at most 2, block rows shorter and longer than a workgroup's strip of 8
    tiles), and a second call on a smaller size reuses the workspace."""
    rng = np.random.default_rng(n + 7 * nrhs)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Z = A + A.T  # complex symmetric
    Zp = np.zeros((n, ld), complex)
    Zp[:, :n] = Z
    Zp[:, n:] = np.nan  # columns past n are never read
    Zd = torch.from_numpy(Zp).to(DEV)
    d = rng.uniform(0.5, 2.0, n)
    dd = torch.from_numpy(d).to(DEV)
    X = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    Xd = torch.from_numpy(X).to(DEV)
    wb = int(_lib.lib().mom_scaled_symv_work_bytes(n, nrhs))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    W = torch.empty_like(Xd)
    P = _lib.ptr
    s = _lib.stream_ptr()
    _lib.call("mom_scaled_symv", n, nrhs, P(Zd), ld, P(dd), P(Xd), P(W), P(work), wb, s)
    ref = d[:, None] * (Z @ (d[:, None] * X))
    got = W.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) < 1e-13 * np.sqrt(n) * np.max(np.abs(ref))
    Wf = torch.empty_like(Xd)
    _lib.call("mom_scaled_matvec", n, nrhs, P(Zd), ld, P(dd), P(Xd), P(Wf), s)
    assert np.max(np.abs(got - Wf.cpu().numpy())) < 1e-13 * np.sqrt(n) * np.max(np.abs(ref))
    if n > 64:  # a smaller problem through the same workspace
        m = n // 2
        Zm = torch.from_numpy(np.ascontiguousarray(Z[:m, :m])).to(DEV)
        Xm = Xd[:m].contiguous()
        Wm = torch.empty_like(Xm)
        _lib.call("mom_scaled_symv", m, nrhs, P(Zm), m, P(dd), P(Xm), P(Wm), P(work), wb, s)
        refm = d[:m, None] * (Z[:m, :m] @ (d[:m, None] * X[:m]))
        assert np.max(np.abs(Wm.cpu().numpy() - refm)) < 1e-13 * np.sqrt(m) * np.max(np.abs(refm))
    with pytest.raises(RuntimeError):
        _lib.call("mom_scaled_symv", n, nrhs, P(Zd), ld, P(dd), P(Xd), P(W), P(work), wb - 16, s)


@pytest.mark.parametrize("n,nrhs,symv", [(300, 2, True), (700, 1, False), (129, 3, True)])
def test_gmres_arnoldi_step(n, nrhs, symv):
    """The fused GMRES step: after k steps from a unit start block, the basis
    is orthonormal per column and each step satisfies the Arnoldi relation
    D Z D W_j = sum_i V_i H[i, j] with W_j = M^-1 V_j (the complex64 LU solve
    of V_j), column by column; H[j+1, j] is real and positive.  The host
    buffer receives the same column as the device workspace."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Z = A + A.T + 4 * np.sqrt(n) * np.eye(n)  # symmetric, well conditioned
    d = rng.uniform(0.5, 2.0, n)
    DZD = d[:, None] * Z * d[None, :]
    P = _lib.ptr
    s = _lib.stream_ptr()
    LUd = torch.from_numpy(DZD).to(DEV, torch.complex64)
    ipiv = torch.empty(n, dtype=torch.int32, device=DEV)
    perm = torch.empty(n, dtype=torch.int32, device=DEV)
    info = torch.zeros(4, dtype=torch.int32, device=DEV)
    wb = int(_lib.lib().mom_getrf_work_bytes(32, n, 64))
    work = torch.empty(wb, dtype=torch.uint8, device=DEV)
    sw = torch.zeros(int(_lib.lib().mom_getrs_inv_work_bytes(32, n)), dtype=torch.uint8, device=DEV)
    _lib.call("mom_getrf", 32, n, P(LUd), n, P(ipiv), P(info), P(work), wb, 64, 0, s)
    _lib.call("mom_build_perm", n, P(ipiv), P(perm), s)
    _lib.call("mom_trtri_blocks", 32, n, P(LUd), n, P(sw), s)
    m, k = 8, 6
    Zd = torch.from_numpy(Z).to(DEV)
    dd = torch.from_numpy(d).to(DEV)
    V = torch.zeros((m + 1, n, nrhs), dtype=torch.complex128, device=DEV)
    W = torch.zeros((m, n, nrhs), dtype=torch.complex128, device=DEV)
    v0 = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    V[0] = torch.from_numpy(v0 / np.linalg.norm(v0, axis=0)).to(DEV)
    H = torch.empty(int(_lib.lib().mom_gmres_work_elems(m, nrhs)), dtype=torch.complex128, device=DEV)
    Hh = torch.empty((m + 1) * nrhs, dtype=torch.complex128, pin_memory=True)
    swb = int(_lib.lib().mom_scaled_symv_work_bytes(n, nrhs))
    swk = torch.empty(swb, dtype=torch.uint8, device=DEV)
    cols = []
    for j in range(k):
        _lib.call("mom_gmres_arnoldi", 32, n, nrhs, j, m, P(LUd), n, P(perm), P(sw), P(Zd), n, P(dd),
                  P(swk) if symv else None, swb if symv else 0, P(V), P(W), P(H), P(Hh), s)
        torch.cuda.synchronize()
        hc = Hh.numpy()[:(j + 2) * nrhs].reshape(j + 2, nrhs).copy()
        np.testing.assert_array_equal(hc, H[:(j + 2) * nrhs].cpu().numpy().reshape(j + 2, nrhs))
        cols.append(hc)
    Vh = V.cpu().numpy()
    Wh = W.cpu().numpy()
    for c in range(nrhs):
        Q = Vh[:k + 1, :, c].T
        assert np.abs(Q.conj().T @ Q - np.eye(k + 1)).max() < 1e-12
        for j in range(k):
            h = cols[j][:, c]
            assert h[j + 1].imag == 0 and h[j + 1].real > 0
            lhs = DZD @ Wh[j, :, c]
            assert np.linalg.norm(lhs - Q[:, :j + 2] @ h) < 1e-12 * np.linalg.norm(lhs)
        # W_j is the complex64 LU solve of V_j: its residual against D Z D is at complex64 level
        for j in range(k):
            r = DZD @ Wh[j, :, c] - Vh[j, :, c]
            assert np.linalg.norm(r) < 1e-4 * np.linalg.norm(Vh[j, :, c])
