"""Lumped components across delta gaps (Engine/mom25d/lumped.py, Woodbury
loading Engine/mom25d/solver.py:1039-1074): the value models and Touchstone
reader on the CPU, and on the GPU the reference's own tests
(Engine/tests/test_lumped.py): circuit limits of a series and a shunt
resistor, loading after the solve equal to loading Z, value overrides and
open gaps, and a series capacitor against the reference's stored openEMS
data (tests/golden/lumped_seriesC_openems.txt, physics tolerance).  The
triangular-mesh variants of those tests are not part of this build."""

import os

import numpy as np
import pytest

from mom25d_amd.lumped import C, L, R, Touchstone, parallel, parse, series

mm = 1e-3
ER, H, W = 3.66, 0.254 * mm, 0.56 * mm
G = 0.5 * mm
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_values(tmp_path):
    """(Engine/tests/test_lumped.py:38-56)"""
    f = 1e9
    assert parse("R 50").z(f) == 50
    assert parse("C 1p").z(f) == pytest.approx(1 / (2j * np.pi * f * 1e-12))
    assert parse("R 10 + L 0.5n").z(f) == pytest.approx(10 + 2j * np.pi * f * 0.5e-9)
    assert parse("C 1p || R 1k").z(f) == pytest.approx(parallel(C(1e-12), R(1e3)).z(f))
    assert series(R(1), L(1e-9)).z(f) == pytest.approx(1 + 2j * np.pi * f * 1e-9)
    p1 = tmp_path / "r.s1p"
    p1.write_text("# GHz S RI R 50\n0.5 -0.333333333333 0\n20 -0.333333333333 0\n")
    assert Touchstone(str(p1)).z(3e9)[0, 0] == pytest.approx(25, rel=1e-9)
    p2 = tmp_path / "t.s2p"
    p2.write_text("# GHz S RI R 50\n1 0.1 0 0.2 0 0.3 0 0.4 0\n10 0.1 0 0.2 0 0.3 0 0.4 0\n")
    t = Touchstone(str(p2))
    s = np.array([[0.1, 0.3], [0.2, 0.4]])  # S21 = 0.2 is row 2, column 1
    assert t.z(5e9) == pytest.approx(50 * (np.eye(2) + s) @ np.linalg.inv(np.eye(2) - s))


def test_touchstone_formats_and_interpolation(tmp_path):
    """MA and DB records give the same S as RI; S is interpolated linearly
    in frequency; a frequency outside the file raises."""
    s11 = 0.3 * np.exp(1j * 0.7)
    ri = tmp_path / "a.s1p"
    ri.write_text(f"# MHz S RI R 50\n1000 {s11.real} {s11.imag}\n3000 0 0\n")
    ma = tmp_path / "b.s1p"
    ma.write_text(f"! comment\n# MHz S MA R 50\n1000 {abs(s11)} {np.degrees(np.angle(s11))}\n3000 0 0\n")
    db = tmp_path / "c.s1p"
    db.write_text(f"# MHz S DB R 50\n1000 {20 * np.log10(abs(s11))} {np.degrees(np.angle(s11))}\n3000 -400 0\n")
    for p in (ri, ma, db):
        t = Touchstone(str(p))
        assert t.s(1e9)[0, 0] == pytest.approx(s11, abs=1e-12)
        assert t.s(2e9)[0, 0] == pytest.approx(0.5 * s11, abs=1e-12)
    with pytest.raises(ValueError, match="outside"):
        Touchstone(str(ri)).z(5e9)


def test_series_part_from_s2p(tmp_path):
    """A series R + L + C written as a 2-port over 1-10 GHz: inside the band
    Z is the file's (interpolated), the model fitted for outside it recovers
    R, L and C and continues Z above and below the band; without
    extrapolation a frequency outside the band raises."""
    from mom25d_amd.lumped import series_from_s2p

    r, l, c = 0.5, 0.15e-9, 0.2e-12
    f = np.linspace(1e9, 10e9, 40)
    z = r + 1j * (2 * np.pi * f * l - 1 / (2 * np.pi * f * c))
    s21 = 2 * 50 / (2 * 50 + z)
    s11 = z / (2 * 50 + z)
    p = tmp_path / "part.s2p"
    p.write_text("# Hz S RI R 50\n" + "".join(f"{fi} {a.real} {a.imag} {b.real} {b.imag} {b.real} {b.imag} "
                                              f"{a.real} {a.imag}\n" for fi, a, b in zip(f, s11, s21)))
    v = parse(str(p))
    r0, r1, ind, cap = v.fit
    assert abs(ind / l - 1) < 1e-9 and abs(cap / c - 1) < 1e-9 and abs(r0 - r) < 1e-9 and abs(r1) < 1e-18
    for fx in (20e9, 0.5e9):
        w = 2 * np.pi * fx
        assert abs(v.z(fx) - (r + 1j * (w * l - 1 / (w * c)))) < 1e-6
    z5 = r + 1j * (2 * np.pi * 5e9 * l - 1 / (2 * np.pi * 5e9 * c))
    assert abs(v.z(5e9) - z5) < 1e-3 * abs(z5)
    with pytest.raises(ValueError, match="outside"):
        series_from_s2p(str(p), extrapolate=False).z(20e9)


def test_parse_grammar_errors():
    """'||' binds weaker than '+', and malformed text raises ValueError."""
    f = 2e9
    v = parse("R 10 + L 1n || C 2p")
    ref = parallel(series(R(10), L(1e-9)), C(2e-12))
    assert v.z(f) == pytest.approx(ref.z(f))
    for bad in ("X 5", "R", "R 5 +", "R 5 C 1p"):
        with pytest.raises(ValueError):
            parse(bad)


def test_layout_rejects_what_this_build_lacks():
    from mom25d_amd.mesh import Layout

    with pytest.raises(ValueError, match="apertures"):
        Layout().add_aperture("L2", (0, 0, 1e-3, 1e-3))
    with pytest.raises(ValueError, match="antipad"):
        Layout().add_via(0, 0, drill=0.3e-3, antipad=0.2e-3)  # smaller than the drill
    lay = Layout().add_rect(0, 0, 1e-3, 1e-3)
    with pytest.raises(ValueError, match="grid"):
        lay.mesh(0.2e-3, kind="tri")


def test_layout_rejects_oblique_and_duplicate_gaps():
    from mom25d_amd.mesh import Layout

    lay = Layout()
    lay.add_lumped("Z1", (0, 0), (1 * mm, 0), W, R(50))
    with pytest.raises(ValueError, match="already used"):
        lay.add_lumped("Z1", (0, 0), (0, 1 * mm), W, R(50))
    with pytest.raises(ValueError, match="oblique"):
        lay.add_lumped("Z2", (0, 0), (1 * mm, 1 * mm), W, R(50))
    with pytest.raises(ValueError, match="no lumped gap"):
        lay.add_component(R(1), ["nope"])


def _series_line(value, second=None):
    """50-ohm line, 5 mm either side of a 0.5 mm gap bridged by `value`;
    optional second gap 2 mm further on (Engine/tests/test_lumped.py:15-31)."""
    from mom25d_amd.mesh import Layout

    lay = Layout()
    lay.add_rect(-5 * mm, -W / 2, -G / 2, W / 2)
    lay.add_port("P1", (-5 * mm, 0), "-x", W, 10 * mm)
    lay.add_lumped("Z1", (-G / 2, 0), (G / 2, 0), W, value)
    if second is None:
        lay.add_rect(G / 2, -W / 2, 5 * mm, W / 2)
    else:
        x2 = 2 * mm
        lay.add_rect(G / 2, -W / 2, x2 - G / 2, W / 2)
        lay.add_lumped("Z2", (x2 - G / 2, 0), (x2 + G / 2, 0), W, second)
        lay.add_rect(x2 + G / 2, -W / 2, 5 * mm, W / 2)
    lay.add_port("P2", (5 * mm, 0), "+x", W, 10 * mm)
    return lay


def _solver(lay, precision="fp64"):
    from mom25d_amd import Solver, Substrate

    return Solver(lay.mesh(0.4 * mm, 0.05 * mm), Substrate(ER, H), verbose=False, precision=precision)


def test_lumped_vectors_are_the_gap_cut():
    """One column per gap: the signed lengths of the edges on the gap's cut
    add up to the bridge width (all current crosses the gap)."""
    lay = _series_line(R(50), second=C(1e-12))
    m = lay.mesh(0.4 * mm, 0.05 * mm)
    Lm = m.lumped_vectors()
    assert Lm.shape == (m.n_unknowns, 2)
    for k in range(2):
        idx = np.nonzero(Lm[:, k])[0]
        assert idx.size >= 2
        np.testing.assert_allclose(np.abs(Lm[idx, k]).sum(), W, rtol=1e-6)  # grid coordinates carry ~1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_series_resistor_matches_circuit(precision):
    """S21 = 2 Z0 / (2 Z0 + R), S11 = R / (2 Z0 + R) for a series R across
    the gap (Engine/tests/test_lumped.py:59-66, grid mesh)."""
    r = _solver(_series_line(R(50)), precision).sweep([1e9, 3e9], interp=None)
    z0 = r.z0_line[:, 0].real
    np.testing.assert_allclose(np.abs(r.S[:, 1, 0]), 2 * z0 / (2 * z0 + 50), atol=1e-3)
    np.testing.assert_allclose(np.abs(r.S[:, 0, 0]), 50 / (2 * z0 + 50), atol=1e-3)


@pytest.mark.gpu
def test_woodbury_equals_direct_loading(tmp_path):
    """Loads applied after the solve (Woodbury) == loads added to Z, for a
    one-port value and a coupled 2-port component (Engine/tests/test_lumped.py:69-91)."""
    import scipy.linalg as sla

    from mom25d_amd.mesh import Layout

    p2 = tmp_path / "c.s2p"
    p2.write_text("# GHz S RI R 50\n1 0.2 0.1 0.5 -0.2 0.5 -0.2 0.1 0.3\n10 0.2 0.1 0.5 -0.2 0.5 -0.2 0.1 0.3\n")
    lay = Layout()
    lay.add_rect(-5 * mm, -W / 2, -G / 2, W / 2)
    lay.add_port("P1", (-5 * mm, 0), "-x", W, 10 * mm)
    lay.add_lumped("A", (-G / 2, 0), (G / 2, 0), W)
    lay.add_rect(G / 2, -W / 2, 1.75 * mm, W / 2)
    lay.add_lumped("B", (1.75 * mm, 0), (2.25 * mm, 0), W)
    lay.add_rect(2.25 * mm, -W / 2, 5 * mm, W / 2)
    lay.add_port("P2", (5 * mm, 0), "+x", W, 10 * mm)
    lay.add_component(Touchstone(str(p2)), ["A", "B"])
    sol = _solver(lay)
    f = 2e9
    Z = sol.impedance(f)
    E = np.stack([pm.excitation for pm in sol.pm], 1)
    x1 = sol.solve_currents(f, Z=Z.copy())
    D = sol.lumped_impedance(f)
    x2 = sla.solve(Z + sol.Lm @ D @ sol.Lm.T, E)
    assert np.abs(x1 - x2).max() < 1e-8 * np.abs(x2).max()
    # the same through the fill (no Z given) and more than 4 right-hand sides
    x3 = sol.solve_currents(f, excitations=np.hstack([E, E, -E]))
    np.testing.assert_allclose(x3[:, :2], x2, atol=1e-8 * np.abs(x2).max())
    np.testing.assert_allclose(x3[:, 4:], -x2, atol=1e-8 * np.abs(x2).max())


@pytest.mark.gpu
def test_value_override_and_open_gap():
    """(Engine/tests/test_lumped.py:94-101)"""
    sol = _solver(_series_line(R(50)))
    f = [2e9]
    a = sol.sweep(f, interp=None, values={"Z1": R(100)}).S
    b = _solver(_series_line(R(100))).sweep(f, interp=None).S
    np.testing.assert_allclose(a, b, atol=1e-10)
    s_open = _solver(_series_line(None)).sweep(f, interp=None).S
    assert abs(s_open[0, 1, 0]) < 0.05 and abs(s_open[0, 0, 0]) > 0.99


@pytest.mark.gpu
def test_shunt_resistor_to_ground():
    """A shunt R from the line to a grounded pad (via): S21 = 2 R / (2 R + Z0)
    (Engine/tests/test_lumped.py:104-115)."""
    from mom25d_amd.mesh import Layout

    lay = Layout()
    lay.add_rect(-5 * mm, -W / 2, 5 * mm, W / 2)
    lay.add_port("P1", (-5 * mm, 0), "-x", W, 10 * mm)
    lay.add_port("P2", (5 * mm, 0), "+x", W, 10 * mm)
    lay.add_lumped("R1", (0, W / 2), (0, W / 2 + G), 0.4 * mm, R(25))
    lay.add_rect(-0.4 * mm, W / 2 + G, 0.4 * mm, W / 2 + G + 0.8 * mm)
    lay.add_via(0, W / 2 + G + 0.4 * mm, drill=0.3 * mm, pad=0.6 * mm)
    r = _solver(lay).sweep([1e9], interp=None)
    z0 = r.z0_line[0, 0].real
    assert abs(r.S[0, 1, 0]) == pytest.approx(2 * 25 / (2 * 25 + z0), abs=0.02)


@pytest.mark.gpu
def test_series_capacitor_vs_openems():
    """1 pF in the 0.5 mm gap against the reference's openEMS lumped
    capacitor data (Engine/tests/test_lumped.py:118-125)."""
    d = np.loadtxt(os.path.join(GOLDEN, "lumped_seriesC_openems.txt"))[[4, 20, 36, 52]]
    f, s11r, s21r = d[:, 0], d[:, 1] + 1j * d[:, 2], d[:, 3] + 1j * d[:, 4]
    S = _solver(_series_line(C(1e-12))).sweep(f, interp=None).renormalize(50)
    np.testing.assert_allclose(np.abs(S[:, 1, 0]), np.abs(s21r), atol=0.03)
    np.testing.assert_allclose(np.abs(S[:, 0, 0]), np.abs(s11r), atol=0.05)
