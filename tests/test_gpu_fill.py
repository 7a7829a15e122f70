"""GPU Green's tables and FP64 fill against the oracle (itself pinned to the
reference's golden Z); the bar is 1e-10 relative to max |Z|."""

import numpy as np
import pytest

from mom25d_amd import Solver
from mom25d_amd import workloads as W
from oracle.solve_ref import OracleSolver

pytestmark = pytest.mark.gpu


def _pair_tables_close(gp, op, R, tol):
    for k in ("gA", "gV", "gW"):
        a, b = getattr(gp, k)(R), getattr(op, k)(R)
        scale = np.max(np.abs(b))
        assert np.max(np.abs(a - b)) < tol * scale, k


def test_slab_tables_match_oracle():
    from oracle.greens_ref import SlabGreens

    lay, sub, mk, _ = W.tee()
    s = Solver(lay.mesh(**mk), sub, verbose=False)
    freqs = np.array([1e9, 5.5e9, 10e9, 20e9])
    tabs = s.tables(freqs)
    R = np.geomspace(1e-6, 0.9 * s.rho_max, 300)
    for fi, f in enumerate(freqs):
        st = s.stack
        o = SlabGreens(f, 3.55, 0.508e-3, 0.0027, s.rho_max, zs_ground=st.surface_impedance(f))
        np.testing.assert_array_equal(tabs[fi].pair(0, 0).x, o.rho_tab)
        _pair_tables_close(tabs[fi].pair(0, 0), o.pair(), R, 1e-11)


def test_layered_tables_match_oracle():
    from oracle.greens_ref import LayeredGreensRef

    lay, st, mk, _ = W.via2()
    m = lay.mesh(**mk)
    s = Solver(m, st, verbose=False)
    freqs = np.array([2e9, 12e9])
    tabs = s.tables(freqs)
    R = np.geomspace(1e-6, 0.9 * s.rho_max, 200)
    for fi, f in enumerate(freqs):
        o = LayeredGreensRef(f, st, s.rho_max, zs_ground=st.surface_impedance(f), levels=m.layer_names)
        for i in range(2):
            for j in range(2):
                _pair_tables_close(tabs[fi].pair(i, j), o.pair(i, j), R, 1e-11)


@pytest.mark.parametrize("f", [1.5e9, 9e9])
def test_tiny_fill_matches_reference_golden(f, golden):
    from tests.test_oracle_assemble import golden_tiny

    rows, R = golden_tiny(golden("impedance.npz"), f)
    lay, sub, mk, _ = W.tiny()
    m = lay.mesh(**mk)
    Z = Solver(m, sub, verbose=False).impedance(f)
    nb = m.n_basis
    rwg = rows < nb
    Zr, Rr = Z[rows[rwg]][:, :nb], R[rwg][:, :nb]
    assert np.max(np.abs(Zr - Rr)) < 1e-10 * np.max(np.abs(Rr))
    assert np.max(np.abs(Z[rows] - R)) < 1e-10 * np.max(np.abs(R))
    assert np.array_equal(Z, Z.T)


def blocks_close(Z, R, nb, tol):
    """Relative error per block (RWG-RWG, via-RWG, via-via): the via-via
    block can be orders of magnitude larger than the rest."""
    errs = []
    for rows, cols in ((slice(0, nb), slice(0, nb)), (slice(nb, None), slice(0, nb)), (slice(nb, None), slice(nb, None))):
        r = R[rows, cols]
        if r.size:
            errs.append(np.max(np.abs(Z[rows, cols] - r)) / np.max(np.abs(r)))
    return max(errs)


@pytest.mark.parametrize("name,f", [("tee", 1e9), ("tee", 10e9), ("via2", 5e9), ("tiny", 3e9)])
def test_fill_matches_oracle(name, f):
    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    Z = Solver(m, st, verbose=False).impedance(f)
    R = OracleSolver(m, st, use_c=True).impedance(f)
    err = blocks_close(Z, R, m.n_basis, 1e-10)
    print(f"{name} {f / 1e9:g} GHz N={m.n_unknowns} max blockwise rel err {err:.2e}")
    assert err < 1e-10


@pytest.mark.parametrize("name,f", [("tee", 10e9), ("via2", 5e9)])
def test_staged_and_global_spline_reads_agree(name, f):
    """The fill with each work item's spline intervals staged in LDS (the
    default) and with every spline read from global memory (fill_stage = 0)
    computes the same arithmetic: the matrices agree to the rounding of the
    accumulation order (the LDS and global atomics sum in arrival order),
    and a cap of 4 intervals mixes both paths within one launch."""
    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    s = Solver(m, st, verbose=False)
    Zs = s.impedance(f)
    out = {}
    for cap in (0, 4):
        s.fill_stage = cap
        out[cap] = s.impedance(f)
    for cap, Z in out.items():
        err = np.max(np.abs(Z - Zs)) / np.max(np.abs(Zs))
        print(f"{name} stage cap {cap}: max rel diff {err:.1e}")
        assert err < 1e-14, cap
