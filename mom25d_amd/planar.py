"""Plane geometry for polygon copper and copper pours.

The reference builds these with shapely (Engine/mom25d/mesh.py:244-306,
488-626); this image has no shapely, so the layout front end works on its
own representation:

  - Polygon: a list of rings, (n, 2) float arrays without the closing
    vertex; the first is the exterior (counter-clockwise), the others holes
    (clockwise), so the polygon lies to the left of every edge;
  - Region: the union of `include` shapes minus the union of `exclude`
    shapes (Polygons or Regions), held implicitly: membership is a point
    test, never a clipped polygon;
  - outline(shape): the boundary of a shape as oriented rings, found from
    its operands' edges: every edge is split where another operand edge
    crosses or touches it, and a piece is kept where the shape differs on
    its two sides (tested a small distance either side of the piece's
    midpoint), oriented with the shape on its left; the kept pieces are then
    linked end to end;
  - mitre_buffer(shape, d): the offset with mitred corners (shapely's
    buffer(d, join_style="mitre"), mitre limit 5) as the shape plus, along
    every boundary edge, the rectangle it sweeps and, at every convex
    corner, the mitre kite (d > 0), or the shape minus the inward
    rectangles and the kites at its reflex corners (d < 0).  Manhattan
    copper stays Manhattan.
"""

import numpy as np

MITRE_LIMIT = 5.0


def _ring_area(r):
    x, y = r[:, 0], r[:, 1]
    return 0.5 * float(np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y))


def _clean_ring(r):
    """(n, 2) float ring without the closing vertex or repeated vertices."""
    r = np.asarray(r, float).reshape(-1, 2)
    if len(r) > 1 and np.all(r[0] == r[-1]):
        r = r[:-1]
    keep = np.ones(len(r), bool)
    keep[1:] = np.any(r[1:] != r[:-1], axis=1)
    r = r[keep]
    if len(r) < 3:
        raise ValueError("a polygon ring needs at least three distinct vertices")
    return r


class Polygon:
    """A polygon with holes: rings[0] the exterior (counter-clockwise), the
    rest holes (clockwise)."""

    def __init__(self, rings, oriented=False):
        rings = [_clean_ring(r) for r in rings]
        if not oriented:
            ext = rings[0] if _ring_area(rings[0]) > 0 else rings[0][::-1]
            rings = [ext] + [h if _ring_area(h) < 0 else h[::-1] for h in rings[1:]]
        self.rings = rings
        v = rings[0]
        self._bb = (float(v[:, 0].min()), float(v[:, 1].min()), float(v[:, 0].max()), float(v[:, 1].max()))

    @classmethod
    def box(cls, x0, y0, x1, y1):
        x0, x1 = min(x0, x1), max(x0, x1)
        y0, y1 = min(y0, y1), max(y0, y1)
        return cls([[(x0, y0), (x1, y0), (x1, y1), (x0, y1)]])

    @classmethod
    def disc(cls, xy, r, segments=32):
        """Regular polygon of `segments` vertices at radius r, one at angle 0
        (a round pad as a buffered point approximates it: 8 segments per
        quarter circle)."""
        a = 2 * np.pi * np.arange(segments) / segments
        return cls([np.stack([xy[0] + r * np.cos(a), xy[1] + r * np.sin(a)], 1)])

    @property
    def bounds(self):
        return self._bb

    @property
    def area(self):
        return sum(_ring_area(r) for r in self.rings)

    def polygons(self):
        return [self]

    def contains(self, x, y, closed=False):
        """Points inside (even-odd over all rings); points on an edge count
        as inside only when `closed`."""
        x, y = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float))
        out = np.zeros(x.shape, bool)
        x0, y0, x1, y1 = self._bb
        sel = (x >= x0) & (x <= x1) & (y >= y0) & (y <= y1)
        if not sel.any():
            return out
        px, py = x[sel], y[sel]
        par = np.zeros(px.shape, bool)
        on = np.zeros(px.shape, bool)
        for r in self.rings:
            ax, ay = r[:, 0:1], r[:, 1:2]
            bx, by = np.roll(r[:, 0], -1)[:, None], np.roll(r[:, 1], -1)[:, None]
            dy = by - ay
            step = max(1, (1 << 22) // len(r))
            for s in range(0, px.size, step):
                qx, qy = px[s:s + step][None, :], py[s:s + step][None, :]
                cross = (ay > qy) != (by > qy)
                xi = ax + (qy - ay) * (bx - ax) / np.where(dy == 0, 1.0, dy)
                par[s:s + step] ^= (np.count_nonzero(cross & (qx < xi), axis=0) & 1).astype(bool)
                cr = (bx - ax) * (qy - ay) - (by - ay) * (qx - ax)
                seg = ((qx >= np.minimum(ax, bx)) & (qx <= np.maximum(ax, bx))
                       & (qy >= np.minimum(ay, by)) & (qy <= np.maximum(ay, by)))
                on[s:s + step] |= np.any((cr == 0) & seg, axis=0)
        out[sel] = (par | on) if closed else (par & ~on)
        return out


class Region:
    """The union of `include` shapes minus the union of `exclude` shapes."""

    def __init__(self, include, exclude=()):
        self.include = list(include)
        self.exclude = list(exclude)

    def polygons(self):
        return [p for s in self.include + self.exclude for p in s.polygons()]

    @property
    def bounds(self):
        """Bounds of the included operands (an outer bound, not the tight one)."""
        b = np.array([s.bounds for s in self.include]).reshape(-1, 4)
        return float(b[:, 0].min()), float(b[:, 1].min()), float(b[:, 2].max()), float(b[:, 3].max())

    def contains(self, x, y, closed=False):
        x, y = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float))
        hit = np.zeros(x.shape, bool)
        for s in self.include:
            hit |= s.contains(x, y, closed)
        for s in self.exclude:
            if not hit.any():
                break
            hit &= ~s.contains(x, y, not closed)
        return hit


def as_polygons(geom):
    """Shapes from a Polygon or Region (returned as is), a rectangle
    (x0, y0, x1, y1), a vertex list, a list of rings (exterior first), or an
    object with `exterior` / `interiors` or `geoms` (shapely-like)."""
    if isinstance(geom, (Polygon, Region)):
        return [geom]
    if hasattr(geom, "geoms"):
        return [p for g in geom.geoms for p in as_polygons(g)]
    if hasattr(geom, "exterior"):
        if getattr(geom, "is_empty", False):
            return []
        return [Polygon([np.asarray(geom.exterior.coords)] + [np.asarray(h.coords) for h in geom.interiors])]
    if isinstance(geom, (tuple, list)) and len(geom) == 4 and all(np.ndim(v) == 0 for v in geom):
        return [Polygon.box(*geom)]
    if isinstance(geom, (tuple, list, np.ndarray)) and len(geom) and all(np.ndim(v) == 1 and len(v) == 2
                                                                        for v in geom):
        return [Polygon([np.asarray(geom, float)])]
    return [Polygon([np.asarray(r, float) for r in geom])]


# ------------------------------------------------------------------ outline
def _edges(polys):
    a = np.concatenate([r for p in polys for r in p.rings]).reshape(-1, 2)
    b = np.concatenate([np.roll(r, -1, axis=0) for p in polys for r in p.rings]).reshape(-1, 2)
    keep = np.any(a != b, axis=1)
    return a[keep], b[keep]


def _candidate_pairs(a, b):
    """Index pairs of edges whose bounding boxes overlap (each pair once)."""
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    order = np.argsort(lo[:, 0], kind="stable")
    slo = lo[order, 0]
    end = np.searchsorted(slo, hi[order, 0], side="right")
    m = len(a)
    cnt = np.maximum(end - np.arange(1, m + 1), 0)
    tot = int(cnt.sum())
    if tot == 0:
        return np.zeros(0, int), np.zeros(0, int)
    p = np.repeat(np.arange(m), cnt)
    q = np.arange(tot) - np.repeat(np.cumsum(cnt) - cnt, cnt) + p + 1
    i, j = order[p], order[q]
    ok = (lo[j, 1] <= hi[i, 1]) & (lo[i, 1] <= hi[j, 1])
    return i[ok], j[ok]


def _split_points(a, b, scale):
    """(edge, t, point) of every point where another edge crosses or
    touches an edge, and of its two ends, sorted by edge and t."""
    d = b - a
    i, j = _candidate_pairs(a, b)
    r, s = d[i], d[j]
    q = a[j] - a[i]
    den = r[:, 0] * s[:, 1] - r[:, 1] * s[:, 0]
    qxr = q[:, 0] * r[:, 1] - q[:, 1] * r[:, 0]
    qxs = q[:, 0] * s[:, 1] - q[:, 1] * s[:, 0]
    lr, ls = np.hypot(r[:, 0], r[:, 1]), np.hypot(s[:, 0], s[:, 1])
    tt = 1e-12
    par = np.abs(den) <= 1e-13 * lr * ls
    n = np.arange(len(a))
    ent = [(n, np.zeros(len(a)), a), (n, np.ones(len(a)), b)]

    # crossing edges: one point for both, coordinates of axis-aligned edges kept exact
    k = ~par
    t = np.where(k, qxs / np.where(k, den, 1.0), -1.0)
    u = np.where(k, qxr / np.where(k, den, 1.0), -1.0)
    hit = k & (t >= -tt) & (t <= 1 + tt) & (u >= -tt) & (u <= 1 + tt)
    ih, jh, th, uh = i[hit], j[hit], t[hit], u[hit]
    P = a[ih] + th[:, None] * d[ih]
    for e in (ih, jh):
        for c in (0, 1):
            P[:, c] = np.where(d[e, c] == 0, a[e, c], P[:, c])
    for e, w in ((ih, th), (jh, uh)):
        P = np.where((np.abs(w) <= tt)[:, None], a[e], P)
        P = np.where((np.abs(w - 1) <= tt)[:, None], b[e], P)
    inner_i = (th > tt) & (th < 1 - tt)
    inner_j = (uh > tt) & (uh < 1 - tt)
    ent.append((ih[inner_i], th[inner_i], P[inner_i]))
    ent.append((jh[inner_j], uh[inner_j], P[inner_j]))

    # collinear overlapping edges: each one's ends split the other
    col = par & (np.abs(qxr) <= 1e-12 * lr * scale)
    ic, jc = i[col], j[col]
    for e, f in ((ic, jc), (jc, ic)):
        de = d[e]
        n2 = np.sum(de * de, axis=1)
        for end in (a[f], b[f]):
            w = np.sum((end - a[e]) * de, axis=1) / n2
            inner = (w > tt) & (w < 1 - tt)
            ent.append((e[inner], w[inner], end[inner]))
    e = np.concatenate([x[0] for x in ent])
    t = np.concatenate([x[1] for x in ent])
    P = np.concatenate([x[2].reshape(-1, 2) for x in ent])
    o = np.lexsort((t, e))
    return e[o], t[o], P[o]


def boundary_pieces(shape):
    """Oriented boundary pieces (P, Q) of a shape, the shape on the left of
    each P -> Q, and the length scale of its operands."""
    polys = shape.polygons()
    if not polys:
        return np.zeros((0, 2)), np.zeros((0, 2)), 1.0
    a, b = _edges(polys)
    v = np.concatenate([a, b])
    scale = float(max(np.ptp(v[:, 0]), np.ptp(v[:, 1]), 1e-30))
    e, t, P = _split_points(a, b, scale)
    same = (e[1:] == e[:-1]) & (t[1:] - t[:-1] > 1e-12) & np.any(P[1:] != P[:-1], axis=1)
    p0, p1 = P[:-1][same], P[1:][same]
    d = p1 - p0
    ln = np.hypot(d[:, 0], d[:, 1])
    nl = np.stack([-d[:, 1], d[:, 0]], 1) / ln[:, None]
    eps = np.minimum(1e-7 * scale, 0.1 * ln)[:, None]
    mid = 0.5 * (p0 + p1)
    probe = np.concatenate([mid + eps * nl, mid - eps * nl])
    side = shape.contains(probe[:, 0], probe[:, 1])
    left, right = side[:len(mid)], side[len(mid):]
    keep = left != right
    flip = right[keep][:, None]
    p, q = p0[keep], p1[keep]
    P, Q = np.where(flip, q, p), np.where(flip, p, q)
    key = np.round(np.concatenate([P, Q], 1) / (1e-9 * scale)).astype(np.int64)
    _, first = np.unique(key, axis=0, return_index=True)
    first = np.sort(first)
    return P[first], Q[first], scale


def _link(P, Q, scale):
    """Rings from oriented pieces linked end to end (at a vertex with several
    ways on, the sharpest left turn), each starting at its lowest-left
    vertex, collinear vertices dropped."""
    quant = 1e-9 * scale
    ks = [tuple(k) for k in np.round(P / quant).astype(np.int64)]
    ke = [tuple(k) for k in np.round(Q / quant).astype(np.int64)]
    out = {}
    for n, k in enumerate(ks):
        out.setdefault(k, []).append(n)
    used = np.zeros(len(P), bool)
    rings = []

    def turn(din, n):
        dout = Q[n] - P[n]
        return np.arctan2(din[0] * dout[1] - din[1] * dout[0], din @ dout)

    for s in range(len(P)):
        if used[s]:
            continue
        chain = [s]
        used[s] = True
        cur = s
        while ke[cur] != ks[s]:
            cand = [n for n in out.get(ke[cur], []) if not used[n]]
            if not cand:
                break
            din = Q[cur] - P[cur]
            cur = max(cand, key=lambda n: turn(din, n))
            used[cur] = True
            chain.append(cur)
        if ke[cur] != ks[s] or len(chain) < 3:
            continue  # an open chain: degenerate slivers, not a boundary
        pts = P[chain]
        e1 = pts - np.roll(pts, 1, axis=0)
        e2 = np.roll(pts, -1, axis=0) - pts
        cr = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
        nrm = np.hypot(e1[:, 0], e1[:, 1]) * np.hypot(e2[:, 0], e2[:, 1])
        pts = pts[~((np.abs(cr) <= 1e-12 * nrm) & (np.sum(e1 * e2, axis=1) > 0))]
        if len(pts) < 3:
            continue
        k = int(np.lexsort((pts[:, 1], pts[:, 0]))[0])
        rings.append(np.roll(pts, -k, axis=0))
    return rings


def outline(shape):
    """The shape as Polygons with oriented rings (exteriors counter-clockwise,
    holes clockwise, each hole in the smallest exterior around it), ordered
    by their lowest-left vertex."""
    P, Q, scale = boundary_pieces(shape)
    rings = _link(P, Q, scale)
    ext = sorted((r for r in rings if _ring_area(r) > 0), key=lambda r: (r[0, 0], r[0, 1]))
    holes = sorted((r for r in rings if _ring_area(r) < 0), key=lambda r: (r[0, 0], r[0, 1]))
    polys = [[r] for r in ext]
    ext_p = [Polygon([r], oriented=True) for r in ext]
    area = [_ring_area(r) for r in ext]
    for h in holes:
        d = h[1] - h[0]
        m = 0.5 * (h[0] + h[1]) + 1e-7 * scale * np.array([-d[1], d[0]]) / np.hypot(*d)
        inside = [n for n, p in enumerate(ext_p) if p.contains(m[0], m[1])]
        if inside:
            polys[min(inside, key=lambda n: area[n])].append(h)
    return [Polygon(r, oriented=True) for r in polys]


def bounds_of(polys):
    """Tight bounds of outlined polygons (None when there are none)."""
    if not polys:
        return None
    b = np.array([p.bounds for p in polys])
    return float(b[:, 0].min()), float(b[:, 1].min()), float(b[:, 2].max()), float(b[:, 3].max())


# ------------------------------------------------------------------ offsets
def _kite(v, n1, n2, e1, e2, d, limit):
    """Mitre corner at v between the offset lines v + d n1 + s e1 and
    v + d n2 + s e2, clipped perpendicular to the bisector at limit |d|
    from v when the mitre is longer."""
    m = (n1 + n2) / (1.0 + n1 @ n2)
    if np.hypot(*m) <= limit:
        return [v, v + d * n1, v + d * m, v + d * n2]
    bis = np.sign(d) * m / np.hypot(*m)
    L = limit * abs(d)
    s1 = (L - d * (n1 @ bis)) / (e1 @ bis)
    s2 = (L - d * (n2 @ bis)) / (e2 @ bis)
    return [v, v + d * n1, v + d * n1 + s1 * e1, v + d * n2 + s2 * e2, v + d * n2]


def offset_parts(polys, d, limit=MITRE_LIMIT):
    """The rectangles swept by every edge of outlined polygons moved by d
    along its outward normal, and the mitre kites at the corners that turn
    away from the move (convex corners for d > 0, reflex ones for d < 0)."""
    parts = []
    for p in polys:
        for r in p.rings:
            e = np.roll(r, -1, axis=0) - r
            eh = e / np.hypot(e[:, 0], e[:, 1])[:, None]
            nr = np.stack([eh[:, 1], -eh[:, 0]], 1)  # outward: the shape is on the left
            for k in range(len(r)):
                a, b = r[k], r[(k + 1) % len(r)]
                parts.append(Polygon([[a, b, b + d * nr[k], a + d * nr[k]]]))
                turn = eh[k - 1, 0] * eh[k, 1] - eh[k - 1, 1] * eh[k, 0]
                if turn * d > 1e-12 * abs(d):
                    parts.append(Polygon([_kite(r[k], nr[k - 1], nr[k], eh[k - 1], -eh[k], d, limit)]))
    return parts


def mitre_buffer(shape, d, polys=None):
    """The shape offset by d with mitred corners (d > 0 grows it); `polys`:
    its outline when already known."""
    polys = outline(shape) if polys is None else polys
    parts = offset_parts(polys, d)
    if d > 0:
        return Region([shape] + parts)
    return Region([shape], parts)


def fence_rings(polys):
    """Rings of outlined polygons in the order and sense the reference's
    fence walks them (shapely output, Engine/mom25d/mesh.py:282-292):
    polygons from the highest top edge down, exteriors clockwise and holes
    counter-clockwise, each ring from its rightmost vertex (the lowest of
    them on a tie).  Fixed by the fence positions the reference committed
    (Engine/tests/data/gcpw_pour_openems_vias.txt) for a pour in two
    rectangles."""
    out = []
    for p in sorted(polys, key=lambda q: -q.bounds[3]):
        for r in p.rings:
            r = r[::-1]
            k = int(np.lexsort((r[:, 1], -r[:, 0]))[0])
            out.append(np.roll(r, -k, axis=0))
    return out


def ring_points(ring, pitch):
    """ceil(perimeter / pitch) points spread evenly along a closed ring from
    its first vertex (Engine/mom25d/mesh.py:288-292: interpolation at
    k * perimeter / n)."""
    seg = np.roll(ring, -1, axis=0) - ring
    ln = np.hypot(seg[:, 0], seg[:, 1])
    per = float(ln.sum())
    n = max(int(np.ceil(per / pitch)), 1)
    t = np.arange(n) * per / n
    cum = np.concatenate([[0.0], np.cumsum(ln)])
    k = np.clip(np.searchsorted(cum, t, side="right") - 1, 0, len(ln) - 1)
    f = (t - cum[k]) / ln[k]
    return [tuple(p) for p in ring[k] + f[:, None] * seg[k]]


def axis_coords(polys, tol):
    """x of the vertical and y of the horizontal edges (within tol) of
    outlined polygons (Engine/mom25d/mesh.py:610-624)."""
    xs, ys = [], []
    for p in polys:
        for r in p.rings:
            d = np.roll(r, -1, axis=0) - r
            xs.extend(r[np.abs(d[:, 0]) < tol, 0].tolist())
            ys.extend(r[np.abs(d[:, 1]) < tol, 1].tolist())
    return xs, ys
