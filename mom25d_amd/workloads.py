"""The five benchmark / parity workloads, built through the public Layout API.

Each builder takes an optional namespace `ns` exposing `Layout`, `Substrate`,
`Stackup`, `Dielectric` and `analytic` (default: this package), so that the
same geometry can be built by any implementation of that API.  Returns
(layout, substrate-or-stackup, mesh kwargs, frequencies [Hz]).

  tee        single-layer microstrip open-stub tee, ~600 RWG, 37 f in 1-10 GHz
  lpf7       seven-section stepped-impedance lowpass (seed 0), ~4k RWG, 101 f in 0.5-12 GHz
  random16   16 seeded random chains of rectangles, 500-3000 RWG each, 64 f in 1-20 GHz
  via2       2-layer top -> via -> inner transition, ~6k RWG, 81 f in 1-15 GHz
  gcpw40     40 x 20 mm via-stitched grounded-coplanar pour, ~12k RWG, 51 f in 1-20 GHz
  gcpw_pour  the reference's pour model: a line in a via-fenced pour, ~5k RWG + 88 vias, 15 f in 1-15 GHz
  plane_transition  top -> via through an inner plane's antipad -> stripline, ~2k unknowns, 59 f in 0.5-15 GHz
"""

import types

import numpy as np

mm = 1e-3
GHz = 1e9


def _default_ns():
    from . import analytic
    from .layered import Dielectric, Stackup
    from .mesh import CoplanarGround, Layout
    from .solver import Substrate

    return types.SimpleNamespace(Layout=Layout, Substrate=Substrate, Stackup=Stackup, Dielectric=Dielectric,
                                 CoplanarGround=CoplanarGround, analytic=analytic)


RO4003 = dict(eps_r=3.55, h=0.508 * mm, tand=0.0027, sigma=5.8e7)


def _ro4003(ns):
    return ns.Substrate(RO4003["eps_r"], RO4003["h"], tand=RO4003["tand"], sigma=RO4003["sigma"])


def tee(ns=None):
    """README quick-start tee (Engine README.md:46-55)."""
    ns = ns or _default_ns()
    W = 1.137 * mm
    lay = ns.Layout()
    lay.add_rect(-W / 2 - 2 * mm, -W / 2, W / 2 + 2 * mm, W / 2)
    lay.add_rect(-W / 2, W / 2, W / 2, W / 2 + 6.6 * mm)
    lay.add_port("P1", (-W / 2 - 2 * mm, 0), "-x", W, 15 * mm)
    lay.add_port("P2", (W / 2 + 2 * mm, 0), "+x", W, 15 * mm)
    return lay, _ro4003(ns), dict(max_cell=0.5 * mm, edge_cell=0.08 * mm), np.linspace(1, 10, 37) * GHz


def tiny(ns=None):
    """Short line with a side stub and one grounding via (~300 RWG) on a
    0.254 mm RO4003-like slab: the smallest parity case (its full impedance
    matrix stays around a megabyte)."""
    ns = ns or _default_ns()
    W = 0.6 * mm
    lay = ns.Layout()
    lay.add_rect(-2 * mm, -W / 2, 2 * mm, W / 2)
    lay.add_rect(-W / 2, W / 2, W / 2, W / 2 + 2 * mm)
    lay.add_port("P1", (-2 * mm, 0), "-x", W, 5 * mm)
    lay.add_port("P2", (2 * mm, 0), "+x", W, 5 * mm)
    lay.add_via(0, W / 2 + 1.5 * mm, drill=0.3 * mm, pad=0.6 * mm)
    sub = ns.Substrate(3.55, 0.254 * mm, tand=0.0027, sigma=5.8e7)
    return lay, sub, dict(max_cell=0.5 * mm, edge_cell=0.15 * mm), np.array([1.5e9, 9e9])


def lpf7(ns=None, max_cell=0.25 * mm, edge_cell=0.04 * mm):
    """Seven sections alternating 0.3 / 4.0 mm wide (narrow first), lengths
    U[2, 6] mm from default_rng(0), 50-ohm feeds of 16 mm."""
    ns = ns or _default_ns()
    lengths = np.random.default_rng(0).uniform(2, 6, 7) * mm
    widths = [0.3 * mm if k % 2 == 0 else 4.0 * mm for k in range(7)]
    lay = ns.Layout()
    x = 0.0
    for w, L in zip(widths, lengths):
        lay.add_rect(x, -w / 2, x + L, w / 2)
        x += L
    W0 = ns.analytic.width_for_z0(50.0, RO4003["h"], RO4003["eps_r"])
    lay.add_port("P1", (0, 0), "-x", W0, 16 * mm)
    lay.add_port("P2", (x, 0), "+x", W0, 16 * mm)
    return lay, _ro4003(ns), dict(max_cell=max_cell, edge_cell=edge_cell), np.linspace(0.5, 12, 101) * GHz


def random_chain(seed, ns=None, n_range=(500, 3000), max_cell=0.5 * mm, edge_cell=0.08 * mm):
    """1-6 rectangles chained along x (widths U[0.2, 3] mm, lengths U[1, 10] mm)
    from default_rng(seed), a port on each end; redrawn from the same
    generator until the mesh has n_range[0] <= N <= n_range[1] unknowns."""
    ns = ns or _default_ns()
    rng = np.random.default_rng(seed)
    h = RO4003["h"]
    while True:
        n = int(rng.integers(1, 7))
        widths = rng.uniform(0.2, 3.0, n) * mm
        lengths = rng.uniform(1.0, 10.0, n) * mm
        lay = ns.Layout()
        x = 0.0
        for w, L in zip(widths, lengths):
            lay.add_rect(x, -w / 2, x + L, w / 2)
            x += L
        for name, w, ref, d in (("P1", widths[0], (0.0, 0.0), "-x"), ("P2", widths[-1], (x, 0.0), "+x")):
            margin = max(1.5 * w, 4 * h)
            lay.add_port(name, ref, d, w, 2 * margin + w + 4 * mm)
        mk = dict(max_cell=max_cell, edge_cell=edge_cell)
        m = lay.mesh(**mk)
        if n_range[0] <= m.n_unknowns <= n_range[1]:
            return lay, _ro4003(ns), mk, np.linspace(1, 20, 64) * GHz


def random16(ns=None):
    return [random_chain(seed, ns) for seed in range(16)]


def via2(ns=None, max_cell=0.16 * mm, edge_cell=0.04 * mm):
    """Top microstrip -> via (drill 0.3 / pad 0.6 mm) -> inner line on two
    0.254 mm layers (Models/via_transition/geometry/via_transition.py:13-25)."""
    ns = ns or _default_ns()
    WT, WI, REF, FEED = 1.1122 * mm, 0.45 * mm, 3 * mm, 12 * mm
    lay = ns.Layout()
    lay.add_rect(-REF, -WT / 2, 0, WT / 2, layer="top")
    lay.add_rect(0, -WI / 2, REF, WI / 2, layer="inner")
    lay.add_port("P1", (-REF, 0), "-x", WT, FEED, layer="top")
    lay.add_port("P2", (REF, 0), "+x", WI, FEED, layer="inner")
    lay.add_via(0, 0, drill=0.3 * mm, pad=0.6 * mm, top="top", bottom="inner")
    st = ns.Stackup([ns.Dielectric(0.254 * mm, 3.66, 0.0037), ns.Dielectric(0.254 * mm, 3.66, 0.0037)],
                    {"inner": 1, "top": 2}, sigma=5.8e7)
    return lay, st, dict(max_cell=max_cell, edge_cell=edge_cell), np.linspace(1, 15, 81) * GHz


def gcpw40(ns=None, max_cell=0.5 * mm, edge_cell=0.1 * mm, pitch=1.5 * mm):
    """Grounded coplanar waveguide in a via-stitched pour on a 40 x 20 mm
    board (0.508 mm RO4003-like substrate, gap 0.2 mm).  The pour is two
    rectangles (the gap cut out of the board), stitched to ground by a
    1.5 mm-pitch grid of vias (drill 0.3 / pad 0.5 mm) kept one pad clear of
    every pour edge."""
    ns = ns or _default_ns()
    h = RO4003["h"]
    W = ns.analytic.width_for_z0(50.0, h, RO4003["eps_r"])
    G, X, Y, FEED = 0.2 * mm, 20 * mm, 10 * mm, 9 * mm
    lay = ns.Layout()
    lay.add_rect(-(X - FEED), -W / 2, X - FEED, W / 2)
    lay.add_port("P1", (-(X - FEED), 0), "-x", W, FEED)
    lay.add_port("P2", (X - FEED, 0), "+x", W, FEED)
    y0 = W / 2 + G
    lay.add_rect(-X, y0, X, Y)
    lay.add_rect(-X, -Y, X, -y0)
    drill, pad = 0.3 * mm, 0.5 * mm
    keep = 0.5 * pad + 0.5 * drill
    xs = np.arange(-X + keep, X - keep + 1e-12, pitch)
    xs += 0.5 * ((X - keep) - xs[-1])
    ys = np.arange(y0 + keep, Y - keep + 1e-12, pitch)
    for y in ys:
        for x in xs:
            for sy in (1, -1):
                lay.add_via(float(x), float(sy * y), drill=drill, pad=pad)
    return lay, _ro4003(ns), dict(max_cell=max_cell, edge_cell=edge_cell), np.linspace(1, 20, 51) * GHz


def gcpw_pour(ns=None, max_cell=0.25 * mm, edge_cell=0.05 * mm):
    """50-ohm microstrip (eps_r 3.66, h 0.254 mm) in a copper pour on both
    sides (gap 0.2 mm, 1.2 mm wide) stitched by via fences (1 mm pitch, 0.3 /
    0.5 mm vias), reference planes 5 mm either side of the centre
    (Models/gcpw_pour/geometry/gcpw_pour.py:8-24, meshed as in
    Models/gcpw_pour/analysis/vs_openems.py:23-25: 15 f in 1-15 GHz)."""
    ns = ns or _default_ns()
    W, G, WP, REF, FEED = 0.5561 * mm, 0.2 * mm, 1.2 * mm, 5 * mm, 6 * mm
    lay = ns.Layout()
    lay.add_rect(-REF, -W / 2, REF, W / 2)
    lay.add_port("P1", (-REF, 0), "-x", W, FEED)
    lay.add_port("P2", (REF, 0), "+x", W, FEED)
    E, y1 = REF + FEED, W / 2 + G + WP
    lay.add_pour((-E, -y1, E, y1), G, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    return lay, ns.Substrate(3.66, 0.254 * mm), dict(max_cell=max_cell, edge_cell=edge_cell), \
        np.linspace(1, 15, 15) * GHz


def gcpw_feed(ns=None, max_cell=0.25 * mm, edge_cell=0.05 * mm):
    """The gcpw_pour line with its pour only between the reference planes
    and GCPW feeds outside them: side-ground strips (gap 0.2 mm, 1.2 mm
    wide) with 1 mm-pitch fence vias along each 6 mm feed
    (Engine/tests/test_gcpw.py:16-25, grid mesh in place of its triangles;
    N = 4 980 RWG + 64 via unknowns)."""
    ns = ns or _default_ns()
    W, G, WP, REF, FEED = 0.5561 * mm, 0.2 * mm, 1.2 * mm, 5 * mm, 6 * mm
    lay = ns.Layout()
    lay.add_rect(-REF, -W / 2, REF, W / 2)
    gnd = ns.CoplanarGround(G, WP, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    lay.add_port("P1", (-REF, 0), "-x", W, FEED, ground=gnd)
    lay.add_port("P2", (REF, 0), "+x", W, FEED, ground=gnd)
    y1 = W / 2 + G + WP
    lay.add_pour((-REF, -y1, REF, y1), G, via_pitch=1 * mm, via_drill=0.3 * mm, via_pad=0.5 * mm)
    return lay, ns.Substrate(3.66, 0.254 * mm), dict(max_cell=max_cell, edge_cell=edge_cell), \
        np.linspace(1, 15, 15) * GHz


def plane_transition(ns=None, antipad=0.7 * mm, stitch=True, max_cell=0.3 * mm, edge_cell=0.05 * mm):
    """Top microstrip (W 0.56 mm) over the inner plane L2 -> via (0.3 / 0.6
    mm) through an antipad -> stripline (W 0.24 mm) between L2 and ground,
    four stitching vias L2 -> ground; three 0.254 mm layers, eps_r 3.66
    (Models/plane_transition/geometry/plane_transition.py:12-29; meshed as in
    Engine/tests/test_antipad.py:21-35).  antipad None: the via passes the
    plane without a modelled hole."""
    ns = ns or _default_ns()
    H, WT, WI = 0.254 * mm, 0.56 * mm, 0.24 * mm
    lay = ns.Layout()
    lay.add_plane("L2")
    lay.add_rect(-3 * mm, -WT / 2, 0, WT / 2, layer="top")
    lay.add_rect(0, -WI / 2, 3 * mm, WI / 2, layer="inner")
    lay.add_port("P1", (-3 * mm, 0), "-x", WT, 12 * mm, layer="top")
    lay.add_port("P2", (3 * mm, 0), "+x", WI, 12 * mm, layer="inner")
    lay.add_via(0, 0, drill=0.3 * mm, pad=0.6 * mm, top="top", bottom="inner", antipad=antipad)
    for a in np.pi / 4 + np.arange(4 if stitch else 0) * np.pi / 2:
        lay.add_via(1 * mm * np.cos(a), 1 * mm * np.sin(a), drill=0.3 * mm, pad=0.6 * mm, top="L2", bottom="gnd")
    st = ns.Stackup([ns.Dielectric(H, 3.66)] * 3, {"inner": 1, "top": 3}, planes={"L2": 2})
    return lay, st, dict(max_cell=max_cell, edge_cell=edge_cell), np.linspace(0.5, 15, 59) * GHz


ALL = {"tiny": tiny, "tee": tee, "lpf7": lpf7, "via2": via2, "gcpw40": gcpw40, "gcpw_pour": gcpw_pour,
       "gcpw_feed": gcpw_feed, "plane_transition": plane_transition}
