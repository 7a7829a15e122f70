"""The oracle's Green's-function tables: closed forms (eps_r = 1 images),
path independence, layered == slab, and the reference's own tables
(golden vectors from tests/golden/make_golden.py)."""

import numpy as np
import pytest

from mom25d_amd.layered import Dielectric, Stackup
from oracle.greens_ref import LayeredGreensRef, SlabGreens

R = np.array([1e-5, 1e-4, 5e-4, 2e-3, 1e-2, 0.03, 0.06])


@pytest.mark.parametrize("f", [1e9, 10e9, 30e9])
def test_air_slab_is_source_minus_image(f):
    """Mirrors Engine/tests/test_mom25d.py:13-24."""
    h = 0.5e-3
    g = SlabGreens(f, 1.0, h, rho_max=0.06)
    k = g.k0
    R2 = np.sqrt(R ** 2 + 4 * h ** 2)
    exact = (np.exp(-1j * k * R) / R - np.exp(-1j * k * R2) / R2) / (4 * np.pi)
    scale = 1 / (4 * np.pi * R)
    assert np.max(np.abs(g.gA(R) - exact) / scale) < 1e-6
    assert np.max(np.abs(g.gV(R) - exact) / scale) < 1e-6


def test_lift_independence():
    """Mirrors Engine/tests/test_mom25d.py:27-33."""
    g1 = SlabGreens(10e9, 4.4, 0.8e-3, rho_max=0.05)
    g2 = SlabGreens(10e9, 4.4, 0.8e-3, rho_max=0.02)
    r = np.array([1e-4, 1e-3, 5e-3, 1.5e-2])
    for a, b in ((g1.gA(r), g2.gA(r)), (g1.gV(r), g2.gV(r))):
        assert np.max(np.abs(a - b) * 4 * np.pi * r) < 1e-6


@pytest.mark.parametrize("f", [1e9, 10e9])
def test_layered_one_layer_equals_slab(f):
    """Mirrors Engine/tests/test_multilayer.py:15-22."""
    r = R[:-1]
    g = SlabGreens(f, 3.55, 0.508e-3, 0.002, rho_max=0.04)
    p = LayeredGreensRef(f, Stackup.microstrip(3.55, 0.508e-3, 0.002), rho_max=0.04).pair("top", "top")
    sc = 4 * np.pi * r
    for a, b in ((p.gA(r), g.gA(r)), (p.gV(r), g.gV(r)), (p.gW(r), g.gW(r))):
        assert np.max(np.abs(a - b) * sc) < 1e-12


@pytest.mark.parametrize("f", [1e9, 20e9])
def test_layered_air_stack_matches_images(f):
    """Mirrors Engine/tests/test_multilayer.py:25-37."""
    r = R[:-1]
    st = Stackup([Dielectric(0.3e-3, 1.0), Dielectric(0.5e-3, 1.0)], {"mid": 1, "top": 2})
    L = LayeredGreensRef(f, st, rho_max=0.04)
    for a in ("top", "mid"):
        for b in ("top", "mid"):
            za, zb = st.level_z(a), st.level_z(b)
            R1, R2 = np.hypot(r, za - zb), np.hypot(r, za + zb)
            ex = (np.exp(-1j * L.k0 * R1) / R1 - np.exp(-1j * L.k0 * R2) / R2) / (4 * np.pi)
            p = L.pair(a, b)
            assert np.max(np.abs(p.gA(r) - ex) * 4 * np.pi * R1) < 1e-6
            assert np.max(np.abs(p.gV(r) - ex) * 4 * np.pi * R1) < 1e-6


@pytest.mark.parametrize("f", [1e9, 7.3e9, 20e9])
def test_slab_tables_match_reference(f, golden):
    from mom25d_amd.solver import Substrate

    t = golden("tables.npz")
    key = f"slab_{f / 1e9:g}"
    zs = Substrate(3.55, 0.508e-3, 0.0027, 5.8e7).surface_impedance(f)
    g = SlabGreens(f, 3.55, 0.508e-3, 0.0027, rho_max=0.05, zs_ground=zs)
    np.testing.assert_array_equal(g.rho_tab, t[key + "_rho"])
    for name in ("SA", "SV", "SW"):
        ref = t[f"{key}_{name}"]
        assert np.max(np.abs(getattr(g, f"{name}_tab") - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("f", [2e9, 12e9])
def test_layered_tables_match_reference(f, golden):
    from mom25d_amd import workloads as W

    t = golden("tables.npz")
    key = f"layered_{f / 1e9:g}"
    lay, st, mk, _ = W.via2()
    g = LayeredGreensRef(f, st, rho_max=0.04, zs_ground=st.surface_impedance(f), levels=["inner", "top"])
    r = t[key + "_R"]
    for a in ("inner", "top"):
        for b in ("inner", "top"):
            p = g.pair(a, b)
            np.testing.assert_allclose(np.array([p.cA, p.cV, p.cW, p.dz], complex), t[f"{key}_{a}_{b}_c"],
                                       rtol=1e-13, atol=0)
            for k in ("gA", "gV", "gW"):
                ref = t[f"{key}_{a}_{b}_{k}"]
                scale = np.abs(ref).max()
                assert np.max(np.abs(getattr(p, k)(r) - ref)) < 1e-12 * scale
