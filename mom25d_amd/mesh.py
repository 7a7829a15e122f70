"""Manhattan layouts, ports, vias and the structured (tensor-grid) RWG mesh.

This is the input side of the per-frequency solve: it must produce exactly
the mesh the reference builds from the same `Layout` calls, because the
impedance matrix is compared entry by entry.  The behaviour follows
Engine/mom25d/mesh.py (Port :62-188, Via :191-241, Layout :309-693,
_grid_lines :696-703, Mesh :706-1116); lumped gaps (add_lumped,
add_component) are axis-aligned bridges as in the reference's grid meshes;
internal S-parameter ports (add_gap_port, add_via_port) as well;
polygon copper and copper pours (add_polygon, add_pour) are meshed as a
staircase on the same grid, their axis-aligned edges becoming grid lines,
with the plane geometry of planar.py in place of shapely; GCPW feeds
(Port.ground) carry side-ground strips and fence vias; solid inner
planes (add_plane) take via ends and split the stack into sections, and
vias pass them through modelled antipads (frills); thick metal, apertures
and triangular meshing are not part of this build.

Conventions kept from the reference because the numbers depend on them:
  * grid lines: every rectangle edge, every port's reference / gap
    coordinate and every lumped bridge's two ends and gap, optional
    refinement lines +-edge_cell, then uniform
    subdivision of each interval into ceil(len / max_cell) parts;
  * polygon and pour edges parallel to an axis are grid lines too, unless
    closer than `snap` (default max_cell / 4) to one already there;
  * a cell is copper if its centre lies strictly inside a rectangle, a
    polygon, a pour or a via pad of the layer; it is split into two
    triangles along a diagonal that alternates in a checkerboard;
  * node numbering: per layer, node (i, j) of the grid has id
    layer * nx * ny + i * ny + j, then the used ids are renumbered in
    ascending order;
  * local edge i of a triangle is the one opposite vertex i; an RWG basis
    lives on every edge shared by two triangles, T+ being the first of the
    two in the stable sort of the (3 Nt) half-edges by node-pair key.
"""

from dataclasses import dataclass, field

import numpy as np

from . import planar

_AXIS_DIRS = {"+x": (1, 0), "-x": (-1, 0), "+y": (0, 1), "-y": (0, -1)}
FRILL_RINGS = 4  # radial potential functions per antipad (antipad.py)


@dataclass
class CoplanarGround:
    """Side grounds of a grounded coplanar (GCPW) feed
    (Engine/mom25d/mesh.py:25-59).

    gap, width : spacing from the line to each ground strip and the strip's
                 width [m]; one number for both sides or a pair (+t, -t),
                 where t is the feed's normal (direction turned +90 deg);
                 a side with gap None or width 0 has no strip
    via_pitch  : fence via spacing along each strip's centre line (None: no
                 fence, the strips must reach ground through other copper)
    via_drill, via_pad : fence via drill and pad diameters [m]; the pad
                 defaults to min(2 drill, 0.9 strip width)
    via_to     : level the fence vias end on ("gnd": the ground plane)
    via_connect : inner planes the fence vias also connect to
    """

    gap: object
    width: object
    via_pitch: float = None
    via_drill: float = 0.3e-3
    via_pad: float = None
    via_to: str = "gnd"
    via_connect: tuple = ()

    def sides(self):
        """[(sign of t, gap, width), ...] of the sides that carry a strip."""
        gaps = tuple(self.gap) if isinstance(self.gap, (tuple, list)) else (self.gap,) * 2
        widths = tuple(self.width) if isinstance(self.width, (tuple, list)) else (self.width,) * 2
        out = []
        for sg, g, w in zip((1, -1), gaps, widths):
            if g is not None and w:
                out.append((sg, float(g), float(w)))
        return out

    def key(self, scale):
        """Hashable form with lengths in units of `scale` (ports with equal
        keys share one line calibration)."""
        return tuple((sg, round(g / scale, 6), round(w / scale, 6)) for sg, g, w in self.sides())


@dataclass
class Port:
    """Microstrip feed with a delta-gap source (Engine/mom25d/mesh.py:62-188).

    ref        : (x, y) of the line centre at the reference plane [m]
    direction  : '+x', '-x', '+y' or '-y', pointing away from the circuit
    width      : line width [m]
    length     : feed length from the reference plane to its open end [m]
    gap_from_end : delta-gap position measured from the open end (default width)
    margin     : distance kept free of near fields in the wave fit (default
                 max(1.5 width, 4 h), applied by the solver)
    layer      : metal layer of the feed
    ground     : CoplanarGround for a GCPW feed (side ground strips along the
                 whole feed, from the reference plane out, with fence vias)
    z0_from    : "voltage" or "reciprocity" (how the line impedance is fixed)
    """

    name: str
    ref: tuple
    direction: str
    width: float
    length: float
    gap_from_end: float = None
    margin: float = None
    layer: str = "top"
    ground: object = None
    z0_from: str = "voltage"

    def __post_init__(self):
        if self.z0_from not in ("voltage", "reciprocity"):
            raise ValueError("z0_from must be 'voltage' or 'reciprocity'")
        if self.ground is not None and not isinstance(self.ground, CoplanarGround):
            raise ValueError("port ground must be a CoplanarGround")
        if not isinstance(self.direction, str) or self.direction not in _AXIS_DIRS:
            raise ValueError(f"direction must be one of {list(_AXIS_DIRS)} (axis-aligned feeds only)")
        self.u = np.array(_AXIS_DIRS[self.direction], float)
        self.n = np.array([-self.u[1], self.u[0]])
        self.ref = (float(self.ref[0]), float(self.ref[1]))
        if self.gap_from_end is None:
            self.gap_from_end = self.width
        self.aligned = True
        self.axis = 0 if self.direction[1] == "x" else 1
        self.sgn = _AXIS_DIRS[self.direction][self.axis]

    @property
    def s_gap(self):
        return self.length - self.gap_from_end

    def coord(self, s):
        """Grid coordinate along the feed axis at outward distance s."""
        return self.ref[self.axis] + self.sgn * s

    def point(self, s, t=0.0):
        """(x, y) at outward distance s along the feed and t across it."""
        s, t = np.broadcast_arrays(np.asarray(s, float), np.asarray(t, float))
        return np.asarray(self.ref) + s[..., None] * self.u + t[..., None] * self.n

    def local(self, xy):
        """Feed coordinates (s, t) of points (..., 2)."""
        d = np.asarray(xy, float) - np.asarray(self.ref)
        return d @ self.u, d @ self.n

    def _span(self, t0, t1):
        """Axis-aligned rectangle over the whole feed length between the
        feed offsets t0 and t1."""
        a = self.axis
        c0, c1 = sorted((self.ref[a], self.coord(self.length)))
        q0, q1 = sorted(float(self.ref[1 - a] + self.n[1 - a] * t) for t in (t0, t1))
        return (c0, q0, c1, q1) if a == 0 else (q0, c0, q1, c1)

    def rect(self):
        return self._span(-0.5 * self.width, 0.5 * self.width)

    def ground_strips(self):
        """Rectangles (x0, y0, x1, y1) of the GCPW side-ground strips, one
        per grounded side (none for a microstrip feed)
        (Engine/mom25d/mesh.py:151-160)."""
        if self.ground is None:
            return []
        W2 = 0.5 * self.width
        return [self._span(sg * (W2 + g), sg * (W2 + g + w)) for sg, g, w in self.ground.sides()]

    def ground_vias(self):
        """Fence vias down the centre line of each side-ground strip, at
        (i + 1/2) via_pitch from the reference plane, leaving out any whose
        pad would cross the feed's open end (Engine/mom25d/mesh.py:162-178)."""
        G = self.ground
        if G is None or not G.via_pitch:
            return []
        s = (np.arange(int(np.floor(self.length / G.via_pitch))) + 0.5) * G.via_pitch
        out = []
        for sg, g, w in G.sides():
            pad = min(G.via_pad or 2 * G.via_drill, 0.9 * w)
            if pad < G.via_drill:
                raise ValueError(f"port {self.name}: ground strip {w * 1e3:.3g} mm too narrow "
                                 f"for {G.via_drill * 1e3:.3g} mm fence vias")
            t = sg * (0.5 * self.width + g + 0.5 * w)
            for si in s[s <= self.length - 0.5 * pad]:
                out.append(Via(tuple(self.point(si, t)), G.via_drill, pad, top=self.layer, bottom=G.via_to,
                               connect=tuple(G.via_connect)))
        return out


@dataclass
class Via:
    """Plated via between two levels (Engine/mom25d/mesh.py:191-241).

    xy : centre [m]; drill : hole diameter [m]; pad : pad diameter (default
    2 drill), added on `top` and on `bottom` unless that is "gnd";
    connect : inner planes between top and bottom it connects to; antipad :
    clearance diameter [m] in the planes it passes (None: no modelled hole,
    the plane solid around the via), or {plane: diameter}."""

    xy: tuple
    drill: float
    pad: float = None
    name: str = None
    top: str = "top"
    bottom: str = "gnd"
    connect: tuple = ()
    antipad: float = None

    def __post_init__(self):
        self.xy = (float(self.xy[0]), float(self.xy[1]))
        self.connect = tuple(self.connect)
        sizes = self.antipad.values() if isinstance(self.antipad, dict) else [self.antipad]
        if any(d is not None and d <= self.drill for d in sizes):
            raise ValueError("via antipad must be larger than its drill")
        if self.pad is None:
            self.pad = 2 * self.drill
        if self.pad < self.drill:
            raise ValueError("via pad smaller than its drill")
        if self.top == self.bottom:
            raise ValueError("via top and bottom layer are the same")

    def pad_layers(self):
        return [n for n in (self.top, self.bottom) if n != "gnd"]

    def antipad_at(self, plane):
        """Clearance diameter [m] in `plane` (None: no modelled hole)."""
        if isinstance(self.antipad, dict):
            return self.antipad.get(plane)
        return self.antipad


@dataclass
class Pour:
    """Copper pour (Engine/mom25d/mesh.py:244-306): `region` minus every
    other copper feature of its layer grown by `clearance` (mitred, so
    Manhattan copper stays Manhattan), optionally stitched to `via_to` by a
    via fence along all of its edges (`via_pitch`) and a square grid of vias
    inside (`fill_pitch`), fence via centres `via_offset` (default pad radius
    + drill / 2) inside the pour edge.  region: a list of planar shapes."""

    region: list
    clearance: float
    via_pitch: float = None
    via_drill: float = 0.3e-3
    via_pad: float = None
    via_offset: float = None
    fill_pitch: float = None
    layer: str = "top"
    via_to: str = "gnd"

    def resolve(self, signal):
        """(pour shape, its outline, [Via, ...]) against the other copper
        `signal` (a planar shape or None)."""
        geom = planar.Region(self.region)
        if signal is not None:
            geom = planar.Region(self.region, [planar.mitre_buffer(signal, self.clearance)])
        polys = planar.outline(geom)
        vias = []
        if (self.via_pitch or self.fill_pitch) and polys:
            pad = self.via_pad or 2 * self.via_drill
            off = self.via_offset if self.via_offset is not None else 0.5 * (pad + self.via_drill)
            core = planar.mitre_buffer(geom, -off, polys)
            rings = planar.outline(core)
            pts = []
            if self.via_pitch:
                for r in planar.fence_rings(rings):
                    pts += planar.ring_points(r, self.via_pitch)
            if self.fill_pitch and rings:
                x0, y0, x1, y1 = planar.bounds_of(rings)
                xs = np.arange(x0, x1 + 1e-12, self.fill_pitch)
                ys = np.arange(y0, y1 + 1e-12, self.fill_pitch)
                X, Y = np.meshgrid(xs + 0.5 * (x1 - xs[-1]), ys + 0.5 * (y1 - ys[-1]))
                inside = core.contains(X.ravel(), Y.ravel())
                pts += list(zip(X.ravel()[inside], Y.ravel()[inside]))
            # near-duplicates (fence corners, fill beside the fence) are dropped
            dmin = 0.5 * min(p for p in (self.via_pitch, self.fill_pitch) if p)
            kept = []
            for q in pts:
                if all(np.hypot(q[0] - k[0], q[1] - k[1]) >= dmin for k in kept):
                    kept.append(q)
            vias = [Via(q, self.via_drill, pad, top=self.layer, bottom=self.via_to) for q in kept]
        return geom, polys, vias


@dataclass
class Layout:
    """Copper rectangles, polygons and pours, feed ports and vias on named
    metal layers (Engine/mom25d/mesh.py:309-693)."""

    rects: list = field(default_factory=list)
    ports: list = field(default_factory=list)
    vias: list = field(default_factory=list)
    rect_layers: list = field(default_factory=list)
    planes: list = field(default_factory=list)
    lumped: list = field(default_factory=list)
    components: list = field(default_factory=list)
    internal_ports: list = field(default_factory=list)
    polys: list = field(default_factory=list)
    poly_layers: list = field(default_factory=list)
    pours: list = field(default_factory=list)
    ground_rects: set = field(default_factory=set)  # rectangles that are GCPW feed grounds

    def add_rect(self, x0, y0, x1, y1, layer="top"):
        self.rects.append((min(x0, x1), min(y0, y1), max(x0, x1), max(y0, y1)))
        self.rect_layers.append(layer)
        return self

    def add_port(self, name, ref, direction, width, length, **kw):
        p = Port(name, tuple(ref), direction, width, length, **kw)
        self.ports.append(p)
        self.add_rect(*p.rect(), layer=p.layer)
        # a GCPW feed's side strips are ground copper: pours join them rather
        # than clearing around them (Engine/mom25d/mesh.py:570-598)
        for r in p.ground_strips():
            self.ground_rects.add(len(self.rects))
            self.add_rect(*r, layer=p.layer)
        self.vias += p.ground_vias()
        return p

    def add_lumped(self, name, a, b, width, value=None, layer="top"):
        """Lumped component from pad point a to pad point b (x, y) [m]: a
        copper bridge of `width` with a delta gap in its middle, loaded with
        `value` (lumped.R / L / C / series / parallel, a number [ohm], a
        callable f -> Z, a string such as 'C 1p' or 'R 10 + L 0.5n'; None: an
        open gap, e.g. one port of a multi-port component).  The voltage is
        taken a -> b.  Returns the bridge's Port geometry
        (Engine/mom25d/mesh.py:353-378; bridges along x or y only)."""
        from .lumped import as_value, parse

        a, b = np.asarray(a, float), np.asarray(b, float)
        d = b - a
        length = float(np.hypot(*d))
        if length <= 0:
            raise ValueError("lumped component: a and b coincide")
        if any(q.name == name for q in self.lumped) or any(q.name == name for q in self.ports):
            raise ValueError(f"name {name!r} already used")
        if abs(d[1]) <= 1e-12 * length:
            direction = "+x" if d[0] > 0 else "-x"
        elif abs(d[0]) <= 1e-12 * length:
            direction = "+y" if d[1] > 0 else "-y"
        else:
            raise ValueError(f"lumped {name}: oblique bridges need a triangular mesh (not part of this build)")
        g = Port(name, tuple(a), direction, width, length, gap_from_end=0.5 * length, layer=layer)
        g.value = None if value is None else (parse(value) if isinstance(value, str) else as_value(value))
        self.lumped.append(g)
        self.add_rect(*g.rect(), layer=layer)
        return g

    def add_gap_port(self, name, a, b, width, layer="top", z_ref=50.0):
        """Internal S-parameter port (reference impedance z_ref [ohm]) across
        a delta gap: a copper bridge from pad point a to pad point b like
        add_lumped, voltage a -> b (Engine/mom25d/mesh.py:406-413)."""
        g = self.add_lumped(name, a, b, width, None, layer)
        self.internal_ports.append(dict(kind="gap", name=name, z_ref=float(z_ref)))
        return g

    def add_via_port(self, name, via, z_ref=50.0):
        """Internal S-parameter port at the bottom of a via that ends on the
        ground plane: a delta gap between the barrel and the plane, voltage
        barrel -> ground, port current flowing down the via.  `via`: index into
        Layout.vias or the Via (Engine/mom25d/mesh.py:415-431)."""
        k = via if isinstance(via, (int, np.integer)) else next(i for i, v in enumerate(self.vias) if v is via)
        if self.vias[k].bottom != "gnd":
            raise ValueError(f"via port {name!r}: the via must end on the ground plane (bottom='gnd')")
        used = [q.name for q in self.lumped] + [q.name for q in self.ports] + [q["name"] for q in self.internal_ports]
        if name in used:
            raise ValueError(f"name {name!r} already used")
        self.internal_ports.append(dict(kind="via", name=name, via=int(k), z_ref=float(z_ref)))
        return k

    def add_component(self, model, gaps):
        """Multi-port component (lumped.Touchstone or a .sNp path) across the
        named lumped gaps, in the model's port order (add the gaps with
        value=None) (Engine/mom25d/mesh.py:433-446)."""
        from .lumped import Touchstone

        if isinstance(model, str):
            model = Touchstone(model)
        names = [q.name for q in self.lumped]
        for n in gaps:
            if n not in names:
                raise ValueError(f"no lumped gap {n!r}")
        if getattr(model, "n", len(gaps)) != len(gaps):
            raise ValueError(f"{model} has {model.n} ports, {len(gaps)} gaps given")
        self.components.append((model, list(gaps)))
        return self

    def add_via(self, x, y, drill=0.3e-3, pad=None, name=None, top="top", bottom="gnd", connect=(),
                antipad=None):
        v = Via((x, y), drill, pad, name, top, bottom, connect, antipad)
        self.vias.append(v)
        return v

    def add_polygon(self, geom, layer="top"):
        """Polygon copper (Engine/mom25d/mesh.py:543-552): a planar.Polygon,
        a vertex list, a list of rings (exterior first) or a shapely-like
        polygon / multipolygon [m].  Axis-aligned edges become grid lines;
        other edges are meshed as a staircase at the grid resolution."""
        for g in planar.as_polygons(geom):
            self.polys.append(g)
            self.poly_layers.append(layer)
        self._pours = None
        return self

    def add_pour(self, region, clearance, via_pitch=None, via_drill=0.3e-3, via_pad=None, via_offset=None,
                 fill_pitch=None, layer="top", via_to="gnd"):
        """Copper pour (normally ground) over `region` (a rectangle
        (x0, y0, x1, y1) or any shape add_polygon takes, metres) on `layer`
        (Engine/mom25d/mesh.py:488-506).  All other copper of that layer --
        rectangles, polygons, port feed lines and via pads -- is cut out of
        it with `clearance` when the layout is meshed, so ports added later
        are respected.  With `via_pitch` a via fence follows every pour edge
        (with `fill_pitch` a grid of vias fills the inside), from the pour to
        `via_to`; without vias the pour floats."""
        self.pours.append(Pour(planar.as_polygons(region), clearance, via_pitch, via_drill, via_pad, via_offset,
                               fill_pitch, layer, via_to))
        self._pours = None
        return self

    def resolved_pours(self):
        """[(shape, outline, vias), ...] of every pour against the current
        copper of its layer, except GCPW feed grounds and their fence vias
        (Engine/mom25d/mesh.py:516-537); cached until copper is added."""
        key = (len(self.rects), len(self.polys), len(self.vias), len(self.pours))
        cache = getattr(self, "_pours", None)
        if cache is not None and cache[0] == key:
            return cache[1]
        out = []
        fence = {v.xy for pt in self.ports for v in pt.ground_vias()}
        for p in self.pours:
            sig = [planar.Polygon.box(*r) for k, (r, n) in enumerate(zip(self.rects, self.rect_layers))
                   if n == p.layer and k not in self.ground_rects]
            sig += [g for g, n in zip(self.polys, self.poly_layers) if n == p.layer]
            sig += [planar.Polygon.disc(v.xy, 0.5 * v.pad, 32) for v in self.vias
                    if p.layer in v.pad_layers() and v.xy not in fence]
            out.append(p.resolve(planar.Region(sig) if sig else None))
        self._pours = (key, out)
        return out

    def _shapes(self, layer=None):
        """Polygons and resolved pours of one layer (None: all layers)."""
        shapes = [g for g, n in zip(self.polys, self.poly_layers) if layer is None or n == layer]
        shapes += [g for (g, _, _), p in zip(self.resolved_pours(), self.pours) if layer is None or p.layer == layer]
        return shapes

    def polygon_grid_coords(self, tol):
        """x of the vertical and y of the horizontal edges of the union of all
        polygons and pours (Engine/mom25d/mesh.py:610-624)."""
        shapes = self._shapes()
        if not shapes:
            return [], []
        return planar.axis_coords(planar.outline(planar.Region(shapes)), tol)

    def add_aperture(self, plane, region, max_cell=None, edge_cell=None, size=None):
        """Openings in inner planes (Engine/mom25d/mesh.py:458-486) carry
        magnetic-current unknowns of their own (aperture.py), which are not
        part of this build."""
        raise ValueError("apertures in inner planes are not supported by this build (solid planes only)")

    def add_plane(self, name):
        if name not in self.planes:
            self.planes.append(name)
        return self

    def layer_names(self):
        """Metal layers in order of first appearance (rectangles, polygons,
        pours, via pads, then the layers pour vias go to)."""
        names = list(self.rect_layers) + list(self.poly_layers) + [p.layer for p in self.pours]
        names += [n for v in self.vias for n in v.pad_layers()]
        names += [p.via_to for p in self.pours if p.via_to != "gnd" and (p.via_pitch or p.fill_pitch)]
        return [n for n in dict.fromkeys(names) if n not in self.planes]

    def all_vias(self):
        """Explicit vias, then the stitching vias of every pour."""
        return list(self.vias) + [v for _, _, vs in self.resolved_pours() for v in vs]

    def bbox(self):
        boxes = [tuple(r) for r in self.rects]
        outl = [q for p in self.polys for q in planar.outline(p)] if self.polys else []
        outl += [q for _, o, _ in self.resolved_pours() for q in o]
        if outl:
            boxes.append(planar.bounds_of(outl))
        for v in self.all_vias():
            r = 0.5 * v.pad
            boxes.append((v.xy[0] - r, v.xy[1] - r, v.xy[0] + r, v.xy[1] + r))
        b = np.array(boxes)
        return b[:, 0].min(), b[:, 1].min(), b[:, 2].max(), b[:, 3].max()

    def contains(self, x, y, layer=None):
        """Points strictly inside the copper of `layer` (None: any layer)."""
        hit = np.zeros(np.broadcast(x, y).shape, bool)
        for (x0, y0, x1, y1), n in zip(self.rects, self.rect_layers):
            if layer is None or n == layer:
                hit |= (x > x0) & (x < x1) & (y > y0) & (y < y1)
        for g in self._shapes(layer):
            hit |= g.contains(x, y)
        for v in self.all_vias():  # via pads: round copper that adds no grid lines
            if layer is None or layer in v.pad_layers():
                hit |= (x - v.xy[0]) ** 2 + (y - v.xy[1]) ** 2 < (0.5 * v.pad) ** 2
        return hit

    def mesh(self, max_cell, edge_cell=None, snap=None, kind="grid"):
        if kind != "grid":
            raise ValueError("only kind='grid' meshes are supported by this build")
        return Mesh.from_layout(self, max_cell, edge_cell, snap)


def _grid_lines(fixed, refine, lo, hi, max_cell, tol):
    """Sorted unique grid coordinates (snapped to tol) subdivided so that no
    interval exceeds max_cell (Engine/mom25d/mesh.py:696-703)."""
    c = np.unique(np.round(np.concatenate([np.asarray(fixed, float), np.asarray(refine, float)]) / tol) * tol)
    c = c[(c >= lo - tol) & (c <= hi + tol)]
    lines = [c[0]]
    for a, b in zip(c[:-1], c[1:]):
        n = int(np.ceil((b - a) / max_cell - 1e-9))
        lines.extend(np.linspace(a, b, n + 1)[1:])
    return np.array(lines)


class Mesh:
    """Triangles of every metal layer on one shared tensor grid; each layer
    has its own nodes, so RWG edges never join layers."""

    def __init__(self, nodes, tris, layout=None, xg=None, yg=None, tri_layer=None, layer_names=("top",)):
        self.nodes = np.asarray(nodes, float)
        self.tris = np.asarray(tris, int)
        self.layout = layout
        self.xg, self.yg = xg, yg
        self.layer_names = list(layer_names)
        self.tri_layer = np.zeros(len(self.tris), int) if tri_layer is None else np.asarray(tri_layer, int)
        self.twin_layer = -np.ones(len(self.layer_names), int)
        self.n_frills = 0
        self.n_ap = 0
        self._orient()
        self._edges()
        self.edge_layer = self.tri_layer[self.tp]
        self.planes = list(getattr(layout, "planes", []))
        self._attach_vias(layout.all_vias() if layout is not None else [])

    @classmethod
    def from_layout(cls, layout, max_cell, edge_cell=None, snap=None):
        """Structured mesh of a layout (Engine/mom25d/mesh.py:728-781)."""
        x0, y0, x1, y1 = layout.bbox()
        tol = 1e-9 * max(x1 - x0, y1 - y0)
        fx = [c for r in layout.rects for c in (r[0], r[2])]
        fy = [c for r in layout.rects for c in (r[1], r[3])]
        for p in layout.ports:
            (fx if p.axis == 0 else fy).extend([p.coord(0), p.coord(p.s_gap)])
        for p in getattr(layout, "lumped", []):  # a lumped gap is a grid line, and so are the bridge's ends
            (fx if p.axis == 0 else fy).extend([p.coord(0), p.coord(p.s_gap), p.coord(p.length)])
        # polygon and pour edges: merged into an existing line closer than snap
        snap = 0.25 * max_cell if snap is None else snap
        px, py = layout.polygon_grid_coords(1e3 * tol) if hasattr(layout, "polygon_grid_coords") else ([], [])
        for f, extra in ((fx, px), (fy, py)):
            for c in sorted(extra):
                if not f or np.min(np.abs(np.asarray(f) - c)) > snap:
                    f.append(c)
        rx, ry = [], []
        if edge_cell:
            ax, ay = np.array(fx), np.array(fy)
            rx = [c + s for c in fx for s in (-edge_cell, edge_cell)]
            ry = [c + s for c in fy for s in (-edge_cell, edge_cell)]
            rx = [c for c in rx if np.min(np.abs(ax - c)) > 0.5 * edge_cell]
            ry = [c for c in ry if np.min(np.abs(ay - c)) > 0.5 * edge_cell]
        xg = _grid_lines(fx, rx, x0, x1, max_cell, tol)
        yg = _grid_lines(fy, ry, y0, y1, max_cell, tol)
        nx, ny = len(xg), len(yg)
        XC, YC = np.meshgrid(0.5 * (xg[1:] + xg[:-1]), 0.5 * (yg[1:] + yg[:-1]), indexing="ij")
        names = layout.layer_names() or ["top"]
        first, second, lay = [], [], []
        for li, name in enumerate(names):
            i, j = np.nonzero(layout.contains(XC, YC, name))
            base = li * nx * ny
            n00 = base + i * ny + j
            n10 = n00 + ny
            n11 = n10 + 1
            n01 = n00 + 1
            odd = ((i + j) % 2 == 1)[:, None]
            first_l = np.where(odd, np.stack([n00, n10, n01], 1), np.stack([n00, n10, n11], 1))
            second_l = np.where(odd, np.stack([n10, n11, n01], 1), np.stack([n00, n11, n01], 1))
            first.append(first_l)
            second.append(second_l)
            lay.append(np.full(2 * i.size, li))
        tris = np.concatenate([t for pair in zip(first, second) for t in pair])
        used, inv = np.unique(tris, return_inverse=True)
        X, Y = np.meshgrid(xg, yg, indexing="ij")
        gx = np.tile(X.ravel(), len(names))[used]
        gy = np.tile(Y.ravel(), len(names))[used]
        return cls(np.stack([gx, gy], 1), inv.reshape(tris.shape), layout, xg, yg, np.concatenate(lay), names)

    # ------------------------------------------------------------ geometry
    def _orient(self):
        """Counter-clockwise vertex order, areas and centroids."""
        v = self.nodes[self.tris]
        e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
        cross = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
        cw = cross < 0
        self.tris[cw] = self.tris[cw][:, [0, 2, 1]]
        self.verts = self.nodes[self.tris]
        self.area = 0.5 * np.abs(cross)
        self.centroid = self.verts.mean(1)

    def _edges(self):
        """RWG bases on shared edges (Engine/mom25d/mesh.py:850-898)."""
        nt = len(self.tris)
        opposite = [(1, 2), (2, 0), (0, 1)]
        half = np.concatenate([np.sort(self.tris[:, list(e)], 1) for e in opposite])
        owner = np.tile(np.arange(nt), 3)
        local = np.repeat(np.arange(3), nt)
        key = half[:, 0] * len(self.nodes) + half[:, 1]
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = ks[1:] == ks[:-1]
        if np.any(dup[1:] & dup[:-1]):
            raise ValueError("non-manifold edge (more than two triangles)")
        a, b = order[:-1][dup], order[1:][dup]
        self.edge_nodes = half[a]
        self.tp, self.vp = owner[a], local[a]
        self.tm, self.vm = owner[b], local[b]
        d = self.nodes[self.edge_nodes[:, 1]] - self.nodes[self.edge_nodes[:, 0]]
        self.edge_len = np.hypot(d[:, 0], d[:, 1])
        self.edge_mid = 0.5 * (self.nodes[self.edge_nodes[:, 0]] + self.nodes[self.edge_nodes[:, 1]])
        self.n_basis = len(self.tp)
        self.wall = np.zeros(self.n_basis, bool)
        self.n_walls = 0
        self.tri_basis = -np.ones((nt, 3), int)
        self.tri_sign = np.zeros((nt, 3))
        idx = np.arange(self.n_basis)
        self.tri_basis[self.tp, self.vp] = idx
        self.tri_sign[self.tp, self.vp] = 1.0
        self.tri_basis[self.tm, self.vm] = idx
        self.tri_sign[self.tm, self.vm] = -1.0

    # ---------------------------------------------------------------- vias
    def _attach_vias(self, vias):
        """Triangles inside each via hole per layer (Engine/mom25d/mesh.py:900-930)."""
        self.vias = list(vias)
        self.via_attach = []
        for v in self.vias:
            att = {}
            for li, name in enumerate(self.layer_names):
                t = self._hole_tris(v, li)
                if t.size:
                    att[name] = t
            for n in v.pad_layers():
                if n in self.layer_names and n not in att:
                    raise ValueError(f"via at ({v.xy[0] * 1e3:.4g}, {v.xy[1] * 1e3:.4g}) mm: "
                                     f"no meshed copper on layer {n}")
            self.via_attach.append(att)
        self.assign_levels(self.layer_names + self.planes)

    def _hole_tris(self, v, li):
        on = self.tri_layer == li
        c = self.centroid
        d2 = (c[:, 0] - v.xy[0]) ** 2 + (c[:, 1] - v.xy[1]) ** 2
        t = np.nonzero(on & (d2 < (0.5 * v.drill) ** 2))[0]
        if t.size == 0:
            t = self._cell_tris(*v.xy)
            t = t[on[t]]
        return t

    def _cell_tris(self, x, y):
        """Triangles of the grid cell(s) holding (x, y); every touching cell
        when the point lies on a grid line (Engine/mom25d/mesh.py:1011-1032)."""
        def span(g, c):
            tol = 1e-6 * (g[-1] - g[0])
            k = np.nonzero(np.abs(g - c) < tol)[0]
            if k.size:
                return g[max(k[0] - 1, 0)], g[min(k[0] + 1, len(g) - 1)]
            i = np.searchsorted(g, c) - 1
            return (g[i], g[i + 1]) if 0 <= i < len(g) - 1 else (c, c)

        (xa, xb), (ya, yb) = span(self.xg, x), span(self.yg, y)
        c = self.centroid
        return np.nonzero((c[:, 0] > xa) & (c[:, 0] < xb) & (c[:, 1] > ya) & (c[:, 1] < yb))[0]

    def assign_levels(self, order):
        """Split vias into vertical segments for a stack order, bottom to top
        (Engine/mom25d/mesh.py:932-976): at attached copper, at connected
        planes and at planes passed through a modelled antipad, where a
        frill node (FRILL_RINGS ring unknowns, antipad.py) joins the segments
        above and below.  Unknown n_basis + k is the downward current of
        segment k; the frill rings follow the segments."""
        rank = {n: i for i, n in enumerate(order)}
        rank["gnd"] = -1
        self.via_segments = []
        self.frills = []  # antipads: FRILL_RINGS unknowns each, after the via segments
        for k, (v, att) in enumerate(zip(self.vias, self.via_attach)):
            for n in (v.top, v.bottom) + v.connect:
                if n not in rank:
                    raise ValueError(f"via layer {n!r} is not in the stack order {order}")
            for n in v.connect:
                if n not in self.planes:
                    raise ValueError(f"via connects to {n!r}, which is not a declared plane")
            hi, lo = sorted((rank[v.top], rank[v.bottom]), reverse=True)
            between = [n for n in order if lo < rank[n] < hi]
            # planes passed through a modelled antipad split the via too (no charge there)
            holes = {n for n in between if n in self.planes and n not in v.connect and v.antipad_at(n) is not None}
            mid = [n for n in between if n in att or n in v.connect or n in holes]
            chain = [order[hi]] + sorted(mid, key=lambda n: -rank[n]) + [order[lo] if lo >= 0 else "gnd"]
            for a, b in zip(chain[:-1], chain[1:]):
                if a in holes:
                    self.frills[-1]["seg_low"] = len(self.via_segments)
                self.via_segments.append(dict(via=k, top=a, bottom=b, top_tris=att.get(a),
                                              bottom_tris=att.get(b)))
                if b in holes:
                    self.frills.append(dict(via=k, plane=b, seg_up=len(self.via_segments) - 1, seg_low=None,
                                            a=0.5 * v.drill, b=0.5 * v.antipad_at(b), xy=v.xy, nr=FRILL_RINGS))
        self.level_order = list(order)
        self.n_vias = len(self.via_segments)
        self.n_frills = sum(fr["nr"] for fr in self.frills)
        self.via_tris = [g["top_tris"] for g in self.via_segments]
        self.n_unknowns = self.n_basis + self.n_vias + self.n_frills

    # --------------------------------------------------------------- ports
    def flow_direction(self):
        """Unit vector of positive RWG flow (T+ -> T-) across each edge."""
        d = self.centroid[self.tm] - self.centroid[self.tp]
        n = np.linalg.norm(d, axis=1)
        return d / np.where(n > 0, n, 1.0)[:, None]

    def cut_edges(self, port, s):
        """RWG edges on the transverse cut at outward distance s, with signs
        making positive current flow toward the circuit."""
        es, _ = port.local(self.nodes[self.edge_nodes])
        _, tm = port.local(self.edge_mid)
        tol = 1e-6 * port.width
        on = (np.abs(es[:, 0] - s) < tol) & (np.abs(es[:, 1] - s) < tol)
        on &= np.abs(tm) < 0.5 * port.width
        li = self.layer_names.index(port.layer)
        on &= self.edge_layer == li
        on &= ~self.wall
        idx = np.nonzero(on)[0]
        sign = np.sign(self.flow_direction()[idx] @ -port.u)
        return idx, sign

    def lumped_vectors(self):
        """(N_unknowns x n_lumped) signed edge lengths l of each lumped gap's
        cut (the delta-gap excitation; the gap current a -> b is l . x)
        (Engine/mom25d/mesh.py:1077-1087)."""
        lps = list(getattr(self.layout, "lumped", []))
        Lm = np.zeros((self.n_unknowns, len(lps)))
        for k, p in enumerate(lps):
            idx, sgn = self.cut_edges(p, p.s_gap)
            if idx.size == 0:
                raise ValueError(f"lumped {p.name}: no mesh edges on its gap")
            Lm[idx, k] = -sgn * self.edge_len[idx]
        return Lm

    def internal_port_vectors(self):
        """(N_unknowns x n_internal) delta-gap vectors of the internal ports:
        a gap port's gap cut (as lumped_vectors, current a -> b), a via port's
        bottom segment (current flowing down into the ground plane)
        (Engine/mom25d/mesh.py:1058-1075)."""
        ips = list(getattr(self.layout, "internal_ports", []))
        names = [q.name for q in getattr(self.layout, "lumped", [])]
        Lm = self.lumped_vectors() if names else None
        out = np.zeros((self.n_unknowns, len(ips)))
        for k, ip in enumerate(ips):
            if ip["kind"] == "gap":
                out[:, k] = Lm[:, names.index(ip["name"])]
            else:
                seg = [i for i, g in enumerate(self.via_segments) if g["via"] == ip["via"] and g["bottom"] == "gnd"]
                if len(seg) != 1:
                    raise ValueError(f"via port {ip['name']}: no via segment ending on the ground plane")
                out[self.n_basis + seg[0], k] = 1.0
        return out

    def cut_positions(self, port):
        g = self.xg if port.axis == 0 else self.yg
        s = port.sgn * (g - port.ref[port.axis])
        s = s[(s > -1e-12) & (s <= port.length + 1e-12)]
        return np.sort(s)
