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


# ------------------------------------------------------- polygons and pours
def same_mesh(a, b):
    assert a.n_unknowns == b.n_unknowns and a.n_vias == b.n_vias
    np.testing.assert_array_equal(a.xg, b.xg)
    np.testing.assert_array_equal(a.yg, b.yg)
    assert digest(a) == digest(b)
    np.testing.assert_array_equal([v.xy for v in a.vias], [v.xy for v in b.vias])


def test_polygon_copper_meshes_like_rectangles():
    """An L-shaped polygon and the two rectangles it is made of give the same
    grid lines, cells and RWG functions."""
    from mom25d_amd import planar

    def build(poly):
        lay = Layout()
        if poly:
            lay.add_polygon(planar.Polygon([[(0, 0), (4 * mm, 0), (4 * mm, 1 * mm), (1 * mm, 1 * mm),
                                             (1 * mm, 3 * mm), (0, 3 * mm)]]))
        else:
            lay.add_rect(0, 0, 4 * mm, 1 * mm)
            lay.add_rect(0, 1 * mm, 1 * mm, 3 * mm)
        lay.add_port("P1", (4 * mm, 0.5 * mm), "+x", 1 * mm, 3 * mm)
        return lay.mesh(0.3 * mm, 0.05 * mm)

    same_mesh(build(True), build(False))


def test_oblique_polygon_is_a_staircase():
    """Cells of a triangle drawn as a polygon: exactly those whose centre lies
    strictly inside its three half-planes (Engine/mom25d/mesh.py:543-552)."""
    lay = Layout()
    lay.add_polygon([(0, 0), (4 * mm, 0), (0, 3 * mm)])
    m = lay.mesh(0.2 * mm)
    xc, yc = 0.5 * (m.xg[1:] + m.xg[:-1]), 0.5 * (m.yg[1:] + m.yg[:-1])
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    want = (X > 0) & (Y > 0) & (3 * X + 4 * Y < 12 * mm)
    assert len(m.tris) == 2 * int(want.sum())
    # the two axis-aligned legs are grid lines, the hypotenuse is not
    assert 0.0 in m.xg and 0.0 in m.yg


def pour_line(pour=True, **kw):
    """Line with feeds in a pour (Engine/tests/test_vias.py:62-80)."""
    w = 0.5561 * mm
    lay = Layout()
    lay.add_rect(-3 * mm, -w / 2, 3 * mm, w / 2)
    lay.add_port("P1", (-3 * mm, 0), "-x", w, 4 * mm)
    lay.add_port("P2", (3 * mm, 0), "+x", w, 4 * mm)
    if pour:
        lay.add_pour((-7 * mm, -2 * mm, 7 * mm, 2 * mm), 0.2 * mm, **kw)
    return lay, w


def test_pour_clearance_and_fence():
    lay, w = pour_line(via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    G = 0.2 * mm
    (geom, polys, vias), = lay.resolved_pours()
    assert len(polys) == 2
    ys = sorted({float(y) for p in polys for y in p.rings[0][:, 1]})
    assert ys == pytest.approx([-2 * mm, -(w / 2 + G), w / 2 + G, 2 * mm], abs=1e-15)
    # fence: each core rectangle (13.2 mm x (2 - 0.4 - w/2 - G - 0.4) mm) carries
    # ceil(perimeter / pitch) vias
    per = 2 * (13.2 * mm + (1.2 * mm - w / 2 - G))
    assert len(vias) == 2 * int(np.ceil(per / (1 * mm)))
    a = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    for v in vias:  # pads inside the pour
        r = 0.99 * v.pad / 2
        assert np.all(geom.contains(v.xy[0] + r * np.cos(a), v.xy[1] + r * np.sin(a)))
        assert v.top == "top" and v.bottom == "gnd" and v.drill == 0.3 * mm
    m = lay.mesh(0.4 * mm)
    assert m.n_vias == len(vias)
    assert np.any(np.isclose(m.yg, w / 2 + G))
    assert lay.layer_names() == ["top"]


def test_pour_fill_equals_rectangles_and_via_grid():
    """A pour with a fill grid meshes exactly like its two rectangles plus
    the vias of the fill formula (Engine/mom25d/mesh.py:293-298: a pitch grid
    over the core's bounds, centred, kept where inside the core)."""
    pitch, drill, pad = 0.9 * mm, 0.3 * mm, 0.5 * mm
    lay, w = pour_line(fill_pitch=pitch, via_drill=drill, via_pad=pad)
    G, off = 0.2 * mm, 0.5 * (pad + drill)
    ref, _ = pour_line(False)
    ref.add_rect(-7 * mm, -2 * mm, 7 * mm, -(w / 2 + G))
    ref.add_rect(-7 * mm, w / 2 + G, 7 * mm, 2 * mm)
    x0, x1 = -7 * mm + off, 7 * mm - off
    y0, y1 = -2 * mm + off, 2 * mm - off
    xs = np.arange(x0, x1 + 1e-12, pitch)
    ys = np.arange(y0, y1 + 1e-12, pitch)
    X, Y = np.meshgrid(xs + 0.5 * (x1 - xs[-1]), ys + 0.5 * (y1 - ys[-1]))
    keep = (X > x0) & (X < x1) & (np.abs(Y) > w / 2 + G + off) & (np.abs(Y) < y1)
    for x, y in zip(X[keep], Y[keep]):
        ref.add_via(float(x), float(y), drill=drill, pad=pad)
    assert len(lay.all_vias()) == int(keep.sum()) > 20
    same_mesh(lay.mesh(0.4 * mm, 0.1 * mm), ref.mesh(0.4 * mm, 0.1 * mm))


def test_pour_clears_via_pads_and_reaches_other_layers():
    """A via pad on the pour layer is cleared by G around its 32-gon; pour
    vias to an inner layer add that layer (Engine/mom25d/mesh.py:508-513)."""
    lay, w = pour_line()
    lay.vias.clear()
    v = lay.add_via(0.0, 1.2 * mm, drill=0.3 * mm, pad=0.6 * mm)
    lay.add_pour((-5 * mm, -1 * mm, 5 * mm, 1 * mm), 0.15 * mm, layer="inner", fill_pitch=2 * mm, via_to="top")
    geoms = lay.resolved_pours()
    g0 = geoms[0][0]
    a = np.linspace(0, 2 * np.pi, 200)
    apothem = 0.3 * mm * np.cos(np.pi / 32) + 0.2 * mm
    assert not np.any(g0.contains(v.xy[0] + 0.99 * apothem * np.cos(a), v.xy[1] + 0.99 * apothem * np.sin(a)))
    r = 1.01 * apothem / np.cos(np.pi / 32)
    assert np.all(g0.contains(v.xy[0] + r * np.cos(a), v.xy[1] + r * np.sin(a)))
    assert lay.layer_names() == ["top", "inner"]
    inner = geoms[1][2]
    assert inner and all(q.top == "inner" and q.bottom == "top" for q in inner)


def test_pour_fence_matches_reference_positions():
    """The 160 fence vias of a 40 mm GCPW pour (line and feeds spanning the
    pour, gap 0.2 mm, 1.2 mm wide, 1 mm pitch, 0.3 / 0.5 mm via) at the
    positions the reference's Pour.resolve wrote for its openEMS model
    (Engine/tests/data/gcpw_pour_openems_vias.txt, micrometres; the case
    'gcpw' of Engine/tests/data/ref_vias_openems.py:26-34)."""
    import os

    ref = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "gcpw_pour_fence_vias.txt"))
    w, G, WP = 0.5561 * mm, 0.2 * mm, 1.2 * mm
    lay = Layout()
    lay.add_rect(-5 * mm, -w / 2, 5 * mm, w / 2)
    lay.add_port("P1", (-5 * mm, 0), "-x", w, 15 * mm)
    lay.add_port("P2", (5 * mm, 0), "+x", w, 15 * mm)
    y1 = w / 2 + G + WP
    lay.add_pour((-20 * mm, -y1, 20 * mm, y1), G, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    got = np.array([v.xy for v in lay.all_vias()]) / 1e-6
    assert got.shape == ref.shape
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)


def gcpw_feed_layout(fence=True):
    """10 mm GCPW line whose pour (with a fence) spans only the reference
    planes; outside them the GCPW feeds carry the side grounds
    (Engine/tests/test_gcpw.py:16-25)."""
    from mom25d_amd import CoplanarGround

    w, G, WP = 0.5561 * mm, 0.2 * mm, 1.2 * mm
    lay = Layout()
    lay.add_rect(-5 * mm, -w / 2, 5 * mm, w / 2)
    gnd = CoplanarGround(G, WP, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm) if fence else None
    lay.add_port("P1", (-5 * mm, 0), "-x", w, 6 * mm, ground=gnd)
    lay.add_port("P2", (5 * mm, 0), "+x", w, 6 * mm, ground=gnd)
    y1 = w / 2 + G + WP
    lay.add_pour((-5 * mm, -y1, 5 * mm, y1), G, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    return lay


def test_gcpw_feed_geometry():
    """Side strips G from the feed, WP wide along the whole 6 mm feed, a
    fence via per millimetre inside each strip, and the pour joined to the
    strips at the reference plane instead of cleared around them
    (Engine/tests/test_gcpw.py:28-48)."""
    from mom25d_amd import CoplanarGround

    w, G, WP = 0.5561 * mm, 0.2 * mm, 1.2 * mm
    lay = gcpw_feed_layout()
    p = lay.ports[0]
    strips = p.ground_strips()
    assert len(strips) == 2
    x0, y0, x1, y1 = p.rect()
    for sx0, sy0, sx1, sy1 in strips:
        assert (sx0, sx1) == pytest.approx((x0, x1), abs=1e-15)
        assert min(abs(sy0 - y1), abs(y0 - sy1)) == pytest.approx(G, rel=1e-9)  # gap to the feed
        assert (sx1 - sx0) * (sy1 - sy0) == pytest.approx(WP * 6 * mm, rel=1e-9)
    vias = p.ground_vias()
    assert len(vias) == 2 * 6
    for v in vias:  # every fence pad inside a strip
        r = 0.99 * v.pad / 2
        assert any(sx0 < v.xy[0] - r and v.xy[0] + r < sx1 and sy0 < v.xy[1] - r and v.xy[1] + r < sy1
                   for sx0, sy0, sx1, sy1 in strips)
        assert (v.drill, v.pad, v.top, v.bottom) == (0.3 * mm, 0.5 * mm, "top", "gnd")
    assert len(lay.vias) == 4 * 6  # both feeds' fences are layout vias
    # the pour meets each strip edge-on at x = -5 mm over the strip's full width
    (geom, _, _), = lay.resolved_pours()
    ys = np.linspace(w / 2 + G, w / 2 + G + WP, 9)[1:-1]
    for sg in (1, -1):
        assert np.all(geom.contains(np.full(7, -5 * mm + 1e-6), sg * ys))
        assert np.all(lay.contains(np.full(7, -5 * mm - 1e-6), sg * ys, "top"))
    m = lay.mesh(0.25 * mm, 0.05 * mm)
    mid = m.edge_mid[:m.n_basis]
    joined = np.isclose(mid[:, 0], -5 * mm) & (np.abs(mid[:, 1]) > w / 2 + G) & (np.abs(mid[:, 1]) < w / 2 + G + WP)
    assert joined.sum() >= 2 * int(WP / (0.25 * mm))  # RWGs across the joint on both sides
    # a one-sided ground: only the +t strip (+t of a +x feed is +y)
    q = Layout().add_port("Q", (0, 0), "+x", w, 6 * mm, ground=CoplanarGround((G, None), (WP, 0)))
    (c,) = q.ground_strips()
    assert c[1] == pytest.approx(w / 2 + G)
    with pytest.raises(ValueError, match="too narrow"):  # pad 0.9 x 0.2 mm < drill 0.3 mm
        Layout().add_port("R", (0, 0), "+x", w, 6 * mm, ground=CoplanarGround(G, 0.2 * mm, via_pitch=1 * mm))


def test_gcpw_fence_vias_are_not_cleared():
    """A pour over the feed too: it clears the feed line but neither the
    side strips nor their fence via pads, which are ground copper, while an
    ordinary via pad on the layer is cleared (Engine/mom25d/mesh.py:528-533)."""
    w, G, WP = 0.5561 * mm, 0.2 * mm, 1.2 * mm
    lay = gcpw_feed_layout()
    lay.pours.clear()
    y1 = w / 2 + G + WP
    lay.add_pour((-11 * mm, -y1, 5 * mm, y1), G)
    lay.add_via(0.0, 1.1 * mm, drill=0.3 * mm, pad=0.5 * mm)
    (geom, _, _), = lay.resolved_pours()
    a = np.linspace(0, 2 * np.pi, 16, endpoint=False)
    for v in lay.ports[0].ground_vias():
        assert np.all(geom.contains(v.xy[0] + 0.3 * mm * np.cos(a), v.xy[1] + 0.3 * mm * np.sin(a)))
    assert not np.any(geom.contains(0.3 * mm * np.cos(a), 1.1 * mm + 0.3 * mm * np.sin(a)))
    assert not np.any(geom.contains(np.array([-8 * mm, -8 * mm]), np.array([w / 2 + 0.5 * G, -w / 2 - 0.5 * G])))
