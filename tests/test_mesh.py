"""The structured mesher reproduces the reference's meshes exactly (golden
digests / arrays written from the reference by tests/golden/make_golden.py)."""

import hashlib

import numpy as np
import pytest

from mom25d_amd import Layout, mm
from mom25d_amd import workloads as W


def digest(m):
    h = hashlib.sha256()
    for a in (m.nodes, m.tris, m.tri_basis, m.tri_sign, m.tri_layer, m.edge_len):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name", ["tee", "lpf7", "via2", "gcpw40"])
def test_workload_mesh_matches_reference(name, golden):
    g = golden("meshes.npz")
    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    assert m.n_unknowns == int(g[f"{name}_n_unknowns"])
    assert m.n_vias == int(g[f"{name}_n_vias"])
    assert digest(m) == str(g[f"{name}_digest"])


def test_tiny_and_tee_arrays(golden):
    g = golden("meshes.npz")
    for name, build in (("tee", W.tee), ("tiny", W.tiny)):
        lay, st, mk, _ = build()
        m = lay.mesh(**mk)
        np.testing.assert_array_equal(m.nodes, g[f"{name}_nodes"])
        np.testing.assert_array_equal(m.tris, g[f"{name}_tris"])
        np.testing.assert_array_equal(m.tri_basis, g[f"{name}_tri_basis"])


@pytest.mark.parametrize("seed", [0, 5, 11])
def test_random_chain_mesh(seed, golden):
    g = golden("meshes.npz")
    lay, st, mk, _ = W.random_chain(seed)
    m = lay.mesh(**mk)
    assert 500 <= m.n_unknowns <= 3000
    assert m.n_unknowns == int(g[f"random{seed}_n_unknowns"])
    assert digest(m) == str(g[f"random{seed}_digest"])


def test_rwg_topology_invariants():
    lay = Layout()
    lay.add_rect(0, 0, 3 * mm, 1 * mm)
    lay.add_port("P1", (0, 0.5 * mm), "-x", 1 * mm, 5 * mm)
    m = lay.mesh(0.5 * mm, 0.1 * mm)
    # every basis has one + and one - half; areas positive (CCW)
    assert np.all(m.area > 0)
    assert np.all((m.tri_sign[m.tp, m.vp] == 1) & (m.tri_sign[m.tm, m.vm] == -1))
    counts = np.bincount(m.tri_basis[m.tri_basis >= 0], minlength=m.n_basis)
    assert np.all(counts == 2)
    # total area equals the copper area
    assert m.area.sum() == pytest.approx((3 + 5) * mm * 1 * mm, rel=1e-12)
