"""The reference's pour model (Models/gcpw_pour: a 50-ohm line in a
via-fenced copper pour, 4 980 RWG + 88 via unknowns) solved on the device:
against the openEMS simulation the reference holds for it
(Engine/tests/data/gcpw_pour_openems.txt: f, re/im S11, re/im S21), with
the line constants the reference states for it (eeff 2.72-2.76, Z0
47.8-48.5 ohm: Engine/tests/test_gcpw.py:77-79), and mixed precision
against FP64 (|dS| < 1e-4)."""

import os

import numpy as np
import pytest

from mom25d_amd import Solver
from mom25d_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)


@pytest.fixture(scope="module")
def pour_runs():
    d = np.loadtxt(os.path.join(HERE, "golden", "gcpw_pour_openems.txt"))[[10, 42]]  # 3 and 11 GHz
    lay, st, mk, _ = W.gcpw_pour()
    m = lay.mesh(**mk)
    assert (m.n_basis, m.n_vias) == (4980, 88)
    out = {p: Solver(m, st, verbose=False, precision=p).sweep(d[:, 0], interp=None) for p in ("fp64", "mixed")}
    return d, out


def test_pour_line_vs_openems(pour_runs):
    d, out = pour_runs
    r = out["fp64"]
    s11r, s21r = d[:, 1] + 1j * d[:, 2], d[:, 3] + 1j * d[:, 4]
    S = r.renormalize(50.0)
    print("eeff", r.eeff[:, 0], "Z0", r.z0_line[:, 0].real, "|S11|", np.abs(S[:, 0, 0]), "openEMS", np.abs(s11r),
          "dphase21", np.angle(S[:, 1, 0] / s21r, deg=True))
    # the reference's values for this line at 3 and 11 GHz, with its tolerances
    np.testing.assert_allclose(r.eeff[:, 0], [2.725, 2.749], rtol=3e-3)
    np.testing.assert_allclose(r.z0_line[:, 0].real, [47.8, 48.2], rtol=5e-3)
    assert np.all(np.abs(np.angle(S[:, 1, 0] / s21r, deg=True)) < 2.5)
    np.testing.assert_allclose(np.abs(S[:, 0, 0]), np.abs(s11r), atol=0.015)
    # reciprocal and passive; the lossless substrate leaves only radiation
    assert np.allclose(S[:, 0, 1], S[:, 1, 0], atol=1e-5)
    power = np.sum(np.abs(S) ** 2, axis=1)
    assert np.all((power < 1 + 1e-6) & (power > 0.97))


def test_pour_mixed_matches_fp64(pour_runs):
    _, out = pour_runs
    dS = np.max(np.abs(out["mixed"].S - out["fp64"].S))
    print(f"gcpw_pour mixed vs fp64 |dS| = {dS:.2e}")
    assert dS < 1e-4


@pytest.fixture(scope="module")
def feed_run():
    d = np.loadtxt(os.path.join(HERE, "golden", "gcpw_pour_openems.txt"))[[10, 42]]  # 3 and 11 GHz
    lay, st, mk, _ = W.gcpw_feed()
    m = lay.mesh(**mk)
    assert (m.n_basis, m.n_vias) == (4980, 64)
    return d, Solver(m, st, verbose=False).sweep(d[:, 0], interp=None)


def test_gcpw_feeds_vs_openems(feed_run):
    """GCPW feeds (side grounds and fences on the feeds, the pour only
    between the reference planes) reproduce the pour-everywhere line and
    openEMS, with the reference's tolerances (Engine/tests/test_gcpw.py:72-89);
    plain microstrip feeds on the same windowed pour would see eeff 2.86 and
    Z0 50.6 ohm instead."""
    d, r = feed_run
    s11r, s21r = d[:, 1] + 1j * d[:, 2], d[:, 3] + 1j * d[:, 4]
    S = r.renormalize(50.0)
    print("eeff", r.eeff[:, 0], "Z0", r.z0_line[:, 0].real, "|S11|", np.abs(S[:, 0, 0]), "openEMS", np.abs(s11r),
          "dphase21", np.angle(S[:, 1, 0] / s21r, deg=True))
    np.testing.assert_allclose(r.eeff[:, 0], [2.725, 2.749], rtol=3e-3)
    np.testing.assert_allclose(r.z0_line[:, 0].real, [47.8, 48.2], rtol=5e-3)
    assert np.all(np.abs(np.angle(S[:, 1, 0] / s21r, deg=True)) < 2.5)
    np.testing.assert_allclose(np.abs(S[:, 0, 0]), np.abs(s11r), atol=0.015)
    assert np.allclose(S[:, 0, 1], S[:, 1, 0], atol=1e-5)


def test_gcpw_and_microstrip_ports_do_not_share_a_line():
    """A GCPW feed and a microstrip feed of the same width are calibrated
    as two lines (Engine/tests/test_gcpw.py:51-57)."""
    from mom25d_amd import CoplanarGround, Layout, Substrate, mm

    w = 0.5561 * mm
    lay = Layout()
    lay.add_rect(-5 * mm, -w / 2, 5 * mm, w / 2)
    lay.add_port("P1", (-5 * mm, 0), "-x", w, 6 * mm, ground=CoplanarGround(0.2 * mm, 1.2 * mm, via_pitch=1 * mm))
    lay.add_port("P2", (5 * mm, 0), "+x", w, 6 * mm)
    sol = Solver(lay.mesh(0.3 * mm, 0.05 * mm), Substrate(3.66, 0.254 * mm), verbose=False)
    assert len(sol.groups) == 2
    lay2 = Layout()
    lay2.add_rect(-5 * mm, -w / 2, 5 * mm, w / 2)
    for name, x, dirn in (("P1", -5 * mm, "-x"), ("P2", 5 * mm, "+x")):
        lay2.add_port(name, (x, 0), dirn, w, 6 * mm, ground=CoplanarGround(0.2 * mm, 1.2 * mm, via_pitch=1 * mm))
    assert len(Solver(lay2.mesh(0.3 * mm, 0.05 * mm), Substrate(3.66, 0.254 * mm), verbose=False).groups) == 1
