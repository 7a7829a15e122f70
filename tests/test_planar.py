"""Plane geometry of polygon copper and pours, pinned by brute force: a
mitred offset of Manhattan copper is the Chebyshev (L-infinity) offset, the
outline of a union of rectangles is what the compressed grid of their
coordinates says, a pad's mitred ring lies between its apothem and
circumradius, and the fence is ceil(perimeter / pitch) points evenly
spread (Engine/mom25d/mesh.py:244-306)."""

import numpy as np
import pytest

from mom25d_amd import planar


def random_rects(seed, n=6):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x0, y0 = rng.uniform(0, 8, 2)
        w, h = rng.uniform(0.3, 3, 2)
        out.append((x0, y0, x0 + w, y0 + h))
    return out


def linf_dist(rects, x, y):
    """L-infinity distance from points to a union of rectangles."""
    d = np.full(np.shape(x), np.inf)
    for x0, y0, x1, y1 in rects:
        dx = np.maximum(np.maximum(x0 - x, x - x1), 0)
        dy = np.maximum(np.maximum(y0 - y, y - y1), 0)
        d = np.minimum(d, np.maximum(dx, dy))
    return d


def compressed_grid(rects):
    xs = np.unique([c for r in rects for c in (r[0], r[2])])
    ys = np.unique([c for r in rects for c in (r[1], r[3])])
    XC, YC = np.meshgrid(0.5 * (xs[1:] + xs[:-1]), 0.5 * (ys[1:] + ys[:-1]), indexing="ij")
    F = linf_dist(rects, XC, YC) == 0
    return xs, ys, F


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_outline_of_rectangle_union(seed):
    rects = random_rects(seed)
    shape = planar.Region([planar.Polygon.box(*r) for r in rects])
    polys = planar.outline(shape)
    xs, ys, F = compressed_grid(rects)
    area = float(np.sum(F * np.outer(np.diff(xs), np.diff(ys))))
    assert sum(p.area for p in polys) == pytest.approx(area, rel=1e-12)
    # boundary coordinates: where the occupancy of the compressed grid changes
    Fp = np.pad(F, 1)
    bx = {xs[i] for i in range(len(xs)) if np.any(Fp[i, 1:-1] != Fp[i + 1, 1:-1])}
    by = {ys[j] for j in range(len(ys)) if np.any(Fp[1:-1, j] != Fp[1:-1, j + 1])}
    px, py = planar.axis_coords(polys, 1e-9)
    assert set(px) == bx and set(py) == by
    for p in polys:  # exteriors counter-clockwise, holes clockwise
        assert planar._ring_area(p.rings[0]) > 0
        assert all(planar._ring_area(h) < 0 for h in p.rings[1:])


@pytest.mark.parametrize("seed,d", [(0, 0.05), (1, 0.13), (2, 0.4), (3, 0.9)])
def test_mitre_offset_of_manhattan_copper_is_chebyshev(seed, d):
    rects = random_rects(seed)
    shape = planar.Region([planar.Polygon.box(*r) for r in rects])
    grown = planar.mitre_buffer(shape, d)
    rng = np.random.default_rng(100 + seed)
    x, y = rng.uniform(-2, 13, (2, 20000))
    dist = linf_dist(rects, x, y)
    ok = np.abs(dist - d) > 1e-9
    assert np.array_equal(grown.contains(x, y)[ok], (dist < d)[ok])
    # the grown copper's outline is the union of the grown rectangles
    big = [(r[0] - d, r[1] - d, r[2] + d, r[3] + d) for r in rects]
    xs, ys, F = compressed_grid(big)
    area = float(np.sum(F * np.outer(np.diff(xs), np.diff(ys))))
    assert sum(p.area for p in planar.outline(grown)) == pytest.approx(area, rel=1e-12)


@pytest.mark.parametrize("seed", [4, 5])
def test_pour_core_is_chebyshev_erosion(seed):
    """A pour over a box, cleared by G around Manhattan copper and shrunk by
    off: a point stays iff it is off from the box edge and G + off from the
    copper in the L-infinity sense."""
    rects = random_rects(seed, 4)
    G, off = 0.2, 0.35
    board = (-1.0, -1.0, 12.0, 12.0)
    sig = planar.Region([planar.Polygon.box(*r) for r in rects])
    pour = planar.Region([planar.Polygon.box(*board)], [planar.mitre_buffer(sig, G)])
    core = planar.mitre_buffer(pour, -off)
    rng = np.random.default_rng(seed)
    x, y = rng.uniform(-2, 13, (2, 20000))
    dist = linf_dist(rects, x, y)
    edge = np.minimum.reduce([x - board[0], y - board[1], board[2] - x, board[3] - y])
    want = (edge > off) & (dist > G + off)
    ok = (np.abs(dist - G - off) > 1e-9) & (np.abs(edge - off) > 1e-9)
    assert np.array_equal(core.contains(x, y)[ok], want[ok])


def test_pad_clearance_between_apothem_and_circumradius():
    r, G = 0.25, 0.2
    pad = planar.Polygon.disc((1.0, -2.0), r, 32)
    grown = planar.mitre_buffer(pad, G)
    apothem = r * np.cos(np.pi / 32) + G
    circum = apothem / np.cos(np.pi / 32)
    a = np.linspace(0, 2 * np.pi, 997)
    for rad, want in ((0.999 * apothem, True), (1.001 * circum, False)):
        assert np.all(grown.contains(1.0 + rad * np.cos(a), -2.0 + rad * np.sin(a)) == want)
    # the mitred ring of a regular 32-gon is the regular 32-gon at the grown apothem
    (p,) = planar.outline(grown)
    assert len(p.rings[0]) == 32
    assert p.area == pytest.approx(32 * apothem ** 2 * np.tan(np.pi / 32), rel=1e-12)


def test_offset_of_union_has_no_spike():
    """A line end hidden inside a pad: the union's offset follows the pad,
    while the line's own mitred corner would reach further out."""
    pad = planar.Polygon.disc((0.0, 0.0), 0.25, 32)
    line = planar.Polygon.box(-2.0, -0.2, 0.1, 0.2)
    G = 0.2
    tip = (0.1 + 0.95 * G, 0.2 + 0.95 * G)  # inside the line's own mitre corner
    assert planar.mitre_buffer(line, G).contains(*tip)
    assert np.hypot(*tip) > (0.25 + G) / np.cos(np.pi / 32)  # beyond the grown pad
    assert not planar.mitre_buffer(planar.Region([pad, line]), G).contains(*tip)


def test_polygon_contains_oblique_edges():
    tri = planar.Polygon([[(0, 0), (4, 0), (0, 3)]])
    rng = np.random.default_rng(7)
    x, y = rng.uniform(-1, 5, (2, 5000))
    want = (x > 0) & (y > 0) & (3 * x + 4 * y < 12)
    assert np.array_equal(tri.contains(x, y), want)
    # points on an edge: outside unless closed
    assert not tri.contains(2.0, 1.5) and tri.contains(2.0, 1.5, closed=True)
    hole = planar.Polygon([[(0, 0), (4, 0), (4, 4), (0, 4)], [(1, 1), (3, 1), (3, 3), (1, 3)]])
    assert hole.area == pytest.approx(12.0)
    assert not hole.contains(2.0, 2.0) and hole.contains(0.5, 2.0)


def test_fence_points():
    ring = np.array([(0.0, 0.0), (13.2, 0.0), (13.2, 0.7), (0.0, 0.7)])
    pts = np.array(planar.ring_points(ring, 1.0))
    per = 2 * (13.2 + 0.7)
    assert len(pts) == int(np.ceil(per / 1.0))
    assert tuple(pts[0]) == (0.0, 0.0)
    # consecutive points are per / n apart along the ring
    s = np.where(pts[:, 1] == 0, pts[:, 0], 0)
    assert np.allclose(np.diff(s[:14]), per / len(pts))


def test_as_polygons_forms():
    box = planar.as_polygons((0, 0, 2, 1))[0]
    lst = planar.as_polygons([(0, 0), (2, 0), (2, 1), (0, 1)])[0]
    rings = planar.as_polygons([[(0, 0), (2, 0), (2, 1), (0, 1)]])[0]
    for p in (box, lst, rings):
        assert p.area == pytest.approx(2.0)
