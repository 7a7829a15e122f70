"""The reference-side binding of the fill (tests/ref_fill_binding.py) on
the device: an AssemblyGeometry of the reference's shape (the oracle's
restatement) and the Green's-function splines in the layout of
_fast.spline_data go in through plain ctypes; the RWG block of Z and the pad
potentials IVpad come out, against the loop of _fast.chunk_vals + scatter
(Engine/mom25d/solver.py:472-483) run by the oracle's C restatement.  Bar:
1e-10 relative to max |Z| (the fill's), and nothing written outside the
block."""

import ctypes

import numpy as np
import pytest

from mom25d_amd import _lib, mm
from mom25d_amd import workloads as W
from oracle.assemble_ref import EPS0, MU0, _c_chunk_vals, clib
from oracle.solve_ref import OracleSolver
from tests.ref_fill_binding import GpuFill

pytestmark = pytest.mark.gpu


def _reference_loop(geom, greens):
    """Z[:nb, :nb] and IVpad as assemble's loop over blocks adds them."""
    nb, nt = geom.nb, geom.nt
    omega = greens.omega
    jwmu, ijwe = 1j * omega * MU0, 1 / (1j * omega * EPS0)
    Z = np.zeros((nb, nb), complex)
    IVpad = np.zeros((geom.pad_tris.size, nt), complex) if geom.nv else None
    cols = np.ascontiguousarray(geom.b, np.int64)
    for lt, pb, r, groups, wo, Xw in geom.blocks:
        vals, IV = _c_chunk_vals(greens, lt, pb, r, groups, wo, geom.V, geom.cA, geom.cV, nt, jwmu, ijwe)
        rows = np.ascontiguousarray(geom.b[pb], np.int64)
        clib().oracle_scatter(ctypes.c_int(pb.size), ctypes.c_int(nt), Z.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_int64(nb), vals.ctypes.data_as(ctypes.c_void_p),
                              rows.ctypes.data_as(ctypes.c_void_p), cols.ctypes.data_as(ctypes.c_void_p))
        if geom.nv:
            pos = geom.pad_pos[pb]
            sel = pos >= 0
            IVpad[pos[sel]] = IV[sel]
    return Z, IVpad


@pytest.mark.parametrize("name,f", [("tee", 6e9), ("via2", 9e9)])
def test_reference_binding_fill_matches_the_reference_loop(name, f):
    if name == "via2":  # two levels, one via: pads and both level pairs
        lay, st, mk, _ = W.via2(max_cell=0.4 * mm, edge_cell=0.1 * mm)
    else:
        lay, st, mk, _ = W.tee()
    m = lay.mesh(**mk)
    o = OracleSolver(m, st)
    geom, greens = o.geometry(), o.greens(f)
    Zref, IVref = _reference_loop(geom, greens)
    nb = geom.nb
    b = GpuFill(_lib.LIB_PATH, geom)
    try:
        b.set_greens(greens)
        Z = np.zeros((nb + 3, nb + 3), complex)  # the RWG block is a view with a longer row stride
        IV = np.zeros_like(IVref) if IVref is not None else None
        print(f"{name}: nb={nb} nt={geom.nt} pads={0 if IV is None else IV.shape[0]} f={f:.3g}", flush=True)
        b.fill(greens.omega, MU0, EPS0, Z, IV)
        b.fill(greens.omega, MU0, EPS0, Z, IV)  # stores, so a second call gives the same
    finally:
        b.close()
    err = np.max(np.abs(Z[:nb, :nb] - Zref)) / np.max(np.abs(Zref))
    print(f"{name}: RWG block rel err {err:.2e}")
    assert err < 1e-10
    assert not np.any(Z[nb:]) and not np.any(Z[:, nb:])
    if IVref is not None:
        e_iv = np.max(np.abs(IV - IVref)) / np.max(np.abs(IVref))
        print(f"{name}: IVpad rel err {e_iv:.2e}")
        assert e_iv < 1e-10
