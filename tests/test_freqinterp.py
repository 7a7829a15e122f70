"""Frequency interpolation of Z (Engine/mom25d/freqinterp.py): the window and
Lagrange weights (the reference's own test), the oracle pinned by closed
forms, and on the GPU the kernels against the oracle, the interpolated sweep
against the exact one, the via-via block and the "auto" default (the
reference's tests, Engine/tests/test_freqinterp.py:18-76)."""

import numpy as np
import pytest

from oracle import freqinterp_ref as ref

GHz = 1e9
mm = 1e-3


def test_lagrange_weights_and_window():
    """(Engine/tests/test_freqinterp.py:18-27)"""
    from mom25d_amd.freqinterp import FreqInterp

    fi = FreqInterp(step=1e9, order=3)
    fi.anchors = np.arange(1, 10) * 1e9
    for f in (1e9, 2.3e9, 5.5e9, 9e9):
        idx = fi._window(f)
        assert len(idx) == 4 and idx == sorted(idx)
        assert fi.anchors[idx[0]] <= f <= fi.anchors[idx[-1]]
        L = fi.weights(f, idx)
        assert L.sum() == pytest.approx(1)
        assert L @ fi.anchors[idx] == pytest.approx(f)
        assert idx == ref.window(fi.anchors, 3, f)
        np.testing.assert_array_equal(L, ref.lagrange_weights(fi.anchors, idx, f))


def _cubic_in_f_case(n=40, nb=31, seed=0):
    """Z(f) = P(f) exp(-j k_e R) / (j w) with P a cubic in f per entry: W is
    then exactly P(f), so cubic interpolation must reproduce Z(f) to
    round-off."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.02, 0.02, (n, 3))
    C = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for _ in range(4)]
    P = lambda f: sum(c * (f / GHz) ** k for k, c in enumerate(C))  # noqa: E731
    kr = 1.7
    R = ref.centre_distance(pos, nb)

    def Z(f):
        w = 2 * np.pi * f
        return P(f) * np.exp(-1j * kr * w / ref.C0 * R) / (1j * w)

    return pos, nb, kr, P, Z


def test_oracle_interpolation_exact_on_cubics():
    pos, nb, kr, P, Z = _cubic_in_f_case()
    anchors = np.linspace(1, 10, 6) * GHz
    for f in (1.3e9, 4.44e9, 9.9e9):
        idx = ref.window(anchors, 3, f)
        Ws = [ref.anchor_transform(Z(anchors[i]), anchors[i], pos, nb, kr) for i in idx]
        np.testing.assert_allclose(Ws[0], P(anchors[idx[0]]), rtol=1e-12, atol=1e-12 * np.abs(Ws[0]).max())
        Zi = ref.combine(Ws, ref.lagrange_weights(anchors, idx, f), f, pos, nb, kr)
        assert np.max(np.abs(Zi - Z(f))) < 1e-11 * np.max(np.abs(Z(f)))
    # a wrong phase sign breaks it (the via-via block has none, so only the rest moves)
    bad = ref.combine([ref.anchor_transform(Z(a), a, pos, nb, -kr) for a in anchors[:4]],
                      ref.lagrange_weights(anchors, [0, 1, 2, 3], 2e9), 2e9, pos, nb, -kr)
    assert np.max(np.abs(bad[:nb, :nb] - Z(2e9)[:nb, :nb])) > 1e-3 * np.max(np.abs(Z(2e9)))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [64, 32])
def test_interp_kernels_match_oracle(prec):
    import ctypes

    import torch

    from mom25d_amd import _lib

    pos, nb, kr, P, Z = _cubic_in_f_case(n=300, nb=271, seed=3)
    anchors = np.linspace(1, 10, 6) * GHz
    dev = "cuda"
    posd = torch.from_numpy(pos).to(dev)
    s = _lib.stream_ptr()
    n = len(pos)
    tdt = torch.complex128 if prec == 64 else torch.complex64
    Wd = {}
    for i, fa in enumerate(anchors):
        Zd = torch.from_numpy(Z(fa)).to(dev)
        W = torch.empty((n, n), dtype=tdt, device=dev)
        w = 2 * np.pi * fa
        _lib.call("mom_interp_anchor", prec, n, nb, _lib.ptr(Zd), n, _lib.ptr(posd), w, kr * w / ref.C0, _lib.ptr(W),
                  n, s)
        Wd[i] = W
        Wr = ref.anchor_transform(Z(fa), fa, pos, nb, kr)
        tol = 1e-13 if prec == 64 else 1e-6
        assert np.max(np.abs(W.cpu().numpy() - Wr)) < tol * np.max(np.abs(Wr))
    for f in (1.3e9, 4.44e9, 9.9e9):
        idx = ref.window(anchors, 3, f)
        L = ref.lagrange_weights(anchors, idx, f)
        out = torch.empty((n, n), dtype=torch.complex128, device=dev)
        ptrs = (ctypes.c_void_p * 4)(*[Wd[i].data_ptr() for i in idx])
        w = 2 * np.pi * f
        _lib.call("mom_interp_combine", prec, n, nb, 4, ptrs, (ctypes.c_double * 4)(*L), n, _lib.ptr(posd), w,
                  kr * w / ref.C0, _lib.ptr(out), n, s)
        got = out.cpu().numpy()
        if prec == 64:
            Wr = [ref.anchor_transform(Z(anchors[i]), anchors[i], pos, nb, kr) for i in idx]
            zr = ref.combine(Wr, L, f, pos, nb, kr)
            assert np.max(np.abs(got - zr)) < 1e-12 * np.max(np.abs(zr))
        assert np.max(np.abs(got - Z(f))) < (1e-11 if prec == 64 else 2e-6) * np.max(np.abs(Z(f)))


def _notch_solver(precision="fp64"):
    """(Engine/tests/test_freqinterp.py:7-15)"""
    from mom25d_amd import Solver, Substrate
    from mom25d_amd.mesh import Layout

    W, Ls = 0.6 * mm, 12 * mm
    lay = Layout()
    lay.add_rect(-W / 2 - 1 * mm, -W / 2, W / 2 + 1 * mm, W / 2)
    lay.add_rect(-W / 2, W / 2, W / 2, W / 2 + Ls)
    lay.add_port("P1", (-W / 2 - 1 * mm, 0), "-x", W, 12 * mm)
    lay.add_port("P2", (W / 2 + 1 * mm, 0), "+x", W, 12 * mm)
    return Solver(lay.mesh(0.5 * mm, 0.05 * mm), Substrate(3.66, 0.254 * mm, 0.003, 5.8e7), verbose=False,
                  precision=precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_interpolated_sweep_matches_exact(precision):
    """Notch filter (stub resonance in band), 1-9 GHz: cubic interpolation from
    4 anchor fills reproduces the exact sweep to |dS| < 1e-3; at the anchors
    it is exact (Engine/tests/test_freqinterp.py:30-45)."""
    from mom25d_amd.freqinterp import FreqInterp

    sol = _notch_solver(precision)
    f = np.linspace(1, 9, 17) * GHz
    ex = sol.sweep(f, interp=None).renormalize(50)
    fi = FreqInterp(step=4e9, order=3)
    S = sol.sweep(f, interp=fi).renormalize(50)
    assert fi.n_fills == 4
    assert np.abs(S - ex).max() < 1e-3
    on = [0, -1]
    assert np.abs(S[on] - ex[on]).max() < 1e-5
    assert len(fi._W) <= 2 * (fi.order + 1)  # at most one window per lane kept


@pytest.mark.gpu
def test_interp_with_vias():
    """Grounding vias: the via-via block gets no phase extraction, and every
    block interpolates to 1e-5 (Engine/tests/test_freqinterp.py:48-66)."""
    from mom25d_amd import Solver, Substrate
    from mom25d_amd.freqinterp import FreqInterp
    from mom25d_amd.mesh import Layout

    W = 0.5561 * mm
    lay = Layout()
    lay.add_rect(-3 * mm, -W / 2, 3 * mm, W / 2)
    lay.add_port("P1", (-3 * mm, 0), "-x", W, 6 * mm)
    lay.add_port("P2", (3 * mm, 0), "+x", W, 6 * mm)
    for x in (-1.5, 0.0, 1.5):
        lay.add_via(x * mm, 0, drill=0.3 * mm, pad=W)
    sol = Solver(lay.mesh(0.3 * mm, 0.05 * mm), Substrate(3.66, 0.254 * mm, 0.0037, 5.8e7), verbose=False)
    fi = FreqInterp(step=1.5e9, order=3).setup(sol, np.array([1e9, 5.5e9]))
    assert fi.report_error(1.6e9) < 1e-5
    assert fi.report_error(4.2e9) < 1e-5


@pytest.mark.gpu
def test_auto_interp_default():
    """sweep() interpolates by default only when it saves fills
    (Engine/tests/test_freqinterp.py:69-76)."""
    sol = _notch_solver()
    few = sol.sweep(np.array([2, 4]) * GHz)
    assert sol.last_interp is None
    many = sol.sweep(np.linspace(2, 4, 9) * GHz)
    assert sol.last_interp is not None and len(sol.last_interp.anchors) == 4
    np.testing.assert_allclose(many.S[[0, -1]], few.S, atol=1e-6)
