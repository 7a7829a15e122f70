"""The package's de-embedding (mom25d_amd.deembed: Levenberg-Marquardt wave
fit on the separable objective, graph-Laplacian reference impedances)
against the oracle's restatement of the reference (oracle.deembed_ref: the
same objective under scipy's Nelder-Mead / bounded Brent with the
reference's options, numpy lstsq of the log-ratios).  Both minimise the same
function, so they must meet at one optimum: the package's to round-off, the
reference's within its stopping tolerances (xatol 1e-10 on gamma / scale)."""

import numpy as np
import pytest

from mom25d_amd.deembed import FitContext, _WaveProblem, anchor_z0, fit_waves, reciprocal_z0, voltage_z0
from oracle.deembed_ref import fit_waves_ref, z0_anchor_ref, z0_reciprocal_ref, z0_voltage_ref


def _wave_data(seed, lossless, noise=1e-4):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(30, 400)
    g = (0 if lossless else rng.uniform(0.1, 5)) + 1j * beta
    z0 = rng.uniform(20, 90) * (1 + (0 if lossless else 0.01j))
    s_list, I_list, V_list = [], [], []
    for _port in range(2):
        s = np.sort(rng.uniform(0.001, 0.012, 12))
        for _exc in range(2):
            A, B = rng.standard_normal(2) + 1j * rng.standard_normal(2)
            I = A * np.exp(g * s) + B * np.exp(-g * s)
            V = z0 * (A * np.exp(g * s) - B * np.exp(-g * s))
            n = noise * (rng.standard_normal(s.size) + 1j * rng.standard_normal(s.size))
            s_list.append(s)
            I_list.append(I + n * np.abs(I).max())
            V_list.append(V + n[::-1] * np.abs(V).max())
    return s_list, I_list, V_list, g, z0, beta * 1.1 * (1e-3 + 1j)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("lossless", [False, True])
def test_wave_fit_meets_reference_optimum(seed, lossless):
    s, I, V, _g, _z0, g0 = _wave_data(seed, lossless)
    a = fit_waves(s, I, V, g0, lossless=lossless)
    b = fit_waves_ref(s, I, V, g0, lossless=lossless)
    assert abs(a[0] - b[0]) < 3e-8 * abs(b[0])
    assert abs(a[1] - b[1]) < 3e-8 * abs(b[1])
    for (pa, pb) in zip(a[2], b[2]):
        assert np.max(np.abs(np.asarray(pa) - pb)) < 1e-6 * np.max(np.abs(pb))
    # never a worse objective than the reference's stopping point
    prob = _WaveProblem(s, I, V, 1.0)
    fa = prob.evaluate(a[0])[0]
    fb = prob.evaluate(b[0])[0]
    assert fa @ fa <= fb @ fb * (1 + 1e-12)
    assert abs(a[3] - b[3]) < 1e-9


@pytest.mark.parametrize("lossless", [False, True])
def test_wave_fit_exact_data(lossless):
    """Noise-free waves: gamma and Z0 recovered to round-off, residual ~0."""
    s, I, V, g, z0, g0 = _wave_data(7, lossless, noise=0.0)
    gf, zf, _amps, r = fit_waves(s, I, V, g0, lossless=lossless)
    assert abs(gf - g) < 1e-9 * abs(g)
    assert abs(zf - z0) < 1e-9 * abs(z0)
    assert r < 1e-9


def _reciprocal_S(z, rng, n):
    """A current-normalised SI consistent with reference impedances z (the
    power-wave S = D SI D^-1 symmetric), perturbed by 1e-3."""
    S = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    S = 0.3 * (S + S.T)
    root = np.sqrt(z)
    SI = S * np.outer(1 / root, root)
    return SI * (1 + 1e-3 * rng.standard_normal((n, n)))


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_reference_impedances_match_oracle(seed):
    rng = np.random.default_rng(seed)
    n = 5
    z = rng.uniform(30, 70, n)
    SI = _reciprocal_S(z, rng, n)
    zphi = z * (1 + 0.05 * rng.standard_normal(n))
    groups = [[0, 1], [2], [3, 4]]
    for z0_from in (["voltage"] * 5, ["voltage", "voltage", "reciprocity", "voltage", "voltage"]):
        ctx = FitContext(groups=groups, width=None, h=None, eps=None, covered=None, z0_from=z0_from, s=None,
                         lossless=False, v_weight=1.0)
        got = reciprocal_z0(ctx, SI, zphi)
        ref = z0_reciprocal_ref(SI, zphi, groups, z0_from)
        np.testing.assert_allclose(got, ref, rtol=1e-12)
    np.testing.assert_allclose(voltage_z0(SI, zphi, groups, 1e-4), z0_voltage_ref(SI, zphi, groups), rtol=1e-12)
    # two feeds (one group) anchored by two internal ports of exact z_ref
    zi = [50.0, 25.0]
    SIa = _reciprocal_S(np.r_[z[:2], zi], rng, 4)
    ctx = FitContext(groups=[[0, 1]], width=None, h=None, eps=None, covered=None, z0_from=["voltage"] * 2, s=None,
                     lossless=False, v_weight=1.0)
    np.testing.assert_allclose(anchor_z0(ctx, SIa, zphi[:2], zi), z0_anchor_ref(SIa, zphi[:2], zi, [[0, 1]]),
                               rtol=1e-12)


def test_reciprocity_recovers_impedances():
    """An exactly reciprocal S: voltage-referenced groups tied by reciprocity
    recover the true ratios, up to the common level set by zphi's mean."""
    rng = np.random.default_rng(4)
    z = np.array([50.0, 50.0, 20.0, 80.0])
    S = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    S = S + S.T
    SI = S * np.outer(1 / np.sqrt(z), np.sqrt(z))
    got = voltage_z0(SI, z * 1.001, [[0, 1], [2], [3]], 1e-8)
    np.testing.assert_allclose(got / got[0], z / z[0], rtol=1e-6)
