"""Write golden vectors of the reference implementation (run in the build
container only, where the reference source tree is mounted read-only at
/root/reference; the GPU box never runs this).

The reference is imported from its source tree and used only to compute
inputs / outputs that no closed form pins: meshes of the workloads, the
Sommerfeld tables, impedance matrices (a small mesh in full, selected rows
of the tee and of the via transition), and S / gamma / z0_line of sweeps with
method="dense", interp=None.  Results are written as small .npz fixtures
next to this script; nothing of the reference itself is stored.

    python tests/golden/make_golden.py [set ...]     (sets: mesh tables z sweep)
"""

import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, "/root/reference/Engine")

import mom25d as ref  # noqa: E402

from mom25d_amd import workloads as W  # noqa: E402

NS = types.SimpleNamespace(Layout=ref.Layout, Substrate=ref.Substrate, Stackup=ref.Stackup,
                           Dielectric=ref.Dielectric, analytic=ref.analytic)
mm = 1e-3


def mesh_digest(m):
    import hashlib

    h = hashlib.sha256()
    for a in (m.nodes, m.tris, m.tri_basis, m.tri_sign, m.tri_layer, m.edge_len):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def do_mesh():
    out = {}
    builders = dict(tee=W.tee, lpf7=W.lpf7, via2=W.via2, gcpw40=W.gcpw40, tiny=W.tiny)
    for name, fn in builders.items():
        lay, st, mk, _ = fn(NS)
        m = lay.mesh(**mk)
        out[f"{name}_n_unknowns"] = m.n_unknowns
        out[f"{name}_n_vias"] = m.n_vias
        out[f"{name}_digest"] = mesh_digest(m)
        if name in ("tee", "tiny"):
            out[f"{name}_nodes"] = m.nodes
            out[f"{name}_tris"] = m.tris
            out[f"{name}_tri_basis"] = m.tri_basis
    for seed in (0, 5, 11):
        lay, st, mk, _ = W.random_chain(seed, NS)
        m = lay.mesh(**mk)
        out[f"random{seed}_n_unknowns"] = m.n_unknowns
        out[f"random{seed}_digest"] = mesh_digest(m)
    np.savez_compressed(os.path.join(HERE, "meshes.npz"), **out)


def do_tables():
    out = {}
    for f in (1e9, 7.3e9, 20e9):
        g = ref.GroundedSlabGreens(f, 3.55, 0.508 * mm, 0.0027, rho_max=0.05,
                                   zs_ground=ref.Substrate(3.55, 0.508 * mm, 0.0027, 5.8e7).surface_impedance(f))
        key = f"slab_{f / 1e9:g}"
        out[key + "_rho"] = g.rho_tab
        out[key + "_SA"], out[key + "_SV"], out[key + "_SW"] = g.SA_tab, g.SV_tab, g.SW_tab
    lay, st, mk, _ = W.via2(NS)
    for f in (2e9, 12e9):
        zg = st.surface_impedance(f)
        g = ref.LayeredGreens(f, st, rho_max=0.04, zs_ground=zg, levels=["inner", "top"])
        R = np.geomspace(1e-6, 0.039, 60)
        key = f"layered_{f / 1e9:g}"
        out[key + "_R"] = R
        for a in ("inner", "top"):
            for b in ("inner", "top"):
                p = g.pair(a, b)
                out[f"{key}_{a}_{b}_gA"] = p.gA(R)
                out[f"{key}_{a}_{b}_gV"] = p.gV(R)
                out[f"{key}_{a}_{b}_gW"] = p.gW(R)
                out[f"{key}_{a}_{b}_c"] = np.array([p.cA, p.cV, p.cW, p.dz], complex)
    np.savez_compressed(os.path.join(HERE, "tables.npz"), **out)


def do_z():
    out = {}
    lay, sub, mk, _ = W.tiny(NS)
    m = lay.mesh(**mk)
    s = ref.Solver(m, sub, verbose=False, method="dense")
    # 9 GHz: the upper triangle of the (symmetric) matrix; 1.5 GHz: every 5th row
    Z = s.impedance(9e9)
    iu = np.triu_indices(Z.shape[0])
    out["tiny_n"] = Z.shape[0]
    out["tiny_Zupper_9"] = Z[iu]
    rows = np.arange(0, m.n_unknowns, 5)
    out["tiny_rows"] = rows
    out["tiny_Zrows_1.5"] = s.impedance(1.5e9)[rows]
    lay, sub, mk, _ = W.tee(NS)
    m = lay.mesh(**mk)
    s = ref.Solver(m, sub, verbose=False, method="dense")
    rows = np.unique(np.linspace(0, m.n_unknowns - 1, 12).astype(int))
    out["tee_rows"] = rows
    for f in (1e9, 10e9):
        out[f"tee_Zrows_{f / 1e9:g}"] = s.impedance(f)[rows]
    lay, st, mk, _ = W.via2(NS)
    m = lay.mesh(**mk)
    s = ref.Solver(m, st, verbose=False, method="dense")
    n = m.n_unknowns
    rows = np.unique(np.r_[np.linspace(0, n - 2, 8).astype(int), n - 1])
    out["via2_rows"] = rows
    t = time.time()
    out["via2_Zrows_5"] = s.impedance(5e9)[rows]
    print(f"via2 Z {time.time() - t:.0f}s", flush=True)
    np.savez_compressed(os.path.join(HERE, "impedance.npz"), **out)


def do_sweep():
    out = {}
    lay, sub, mk, f = W.tee(NS)
    m = lay.mesh(**mk)
    fs = f[[0, 9, 18, 27, 36]]
    r = ref.Solver(m, sub, verbose=False, method="dense").sweep(fs, interp=None)
    for k in ("S", "z0", "gamma", "z0_line", "fit_residual"):
        out[f"tee_{k}"] = getattr(r, k)
    out["tee_freqs"] = fs
    lay, sub, mk, _ = W.tiny(NS)
    m = lay.mesh(**mk)
    fs = np.array([1.5e9, 9e9])
    r = ref.Solver(m, sub, verbose=False, method="dense").sweep(fs, interp=None)
    for k in ("S", "z0", "gamma", "z0_line"):
        out[f"tiny_{k}"] = getattr(r, k)
    out["tiny_freqs"] = fs
    lay, st, mk, f = W.via2(NS)
    m = lay.mesh(**mk)
    fs = np.array([3e9, 11e9])
    t = time.time()
    r = ref.Solver(m, st, verbose=False, method="dense").sweep(fs, interp=None)
    print(f"via2 sweep {time.time() - t:.0f}s", flush=True)
    for k in ("S", "z0", "gamma", "z0_line"):
        out[f"via2_{k}"] = getattr(r, k)
    out["via2_freqs"] = fs
    np.savez_compressed(os.path.join(HERE, "sweeps.npz"), **out)


def _dense_sweep(name, fn, fs, rows_at=None, rows=None, currents_at=None):
    """The reference's plain per-frequency path (method="dense", interp=None)
    on one workload: S / z0 / gamma / z0_line / fit_residual at fs, with
    optionally the impedance rows `rows` at rows_at and the port currents
    at currents_at (one of fs)."""
    lay, sub, mk, _ = fn(NS)
    m = lay.mesh(**mk)
    out = {f"{name}_digest": mesh_digest(m), f"{name}_n_unknowns": m.n_unknowns, f"{name}_freqs": np.asarray(fs)}
    s = ref.Solver(m, sub, verbose=False, method="dense")
    t = time.time()
    r = s.sweep(fs, interp=None, keep_currents=currents_at is not None)
    print(f"{name} sweep N={m.n_unknowns}: {time.time() - t:.0f}s", flush=True)
    for k in ("S", "z0", "gamma", "z0_line", "fit_residual"):
        out[f"{name}_{k}"] = getattr(r, k)
    if currents_at is not None:
        k = int(np.argmin(np.abs(np.asarray(fs) - currents_at)))
        out[f"{name}_currents"] = np.asarray(r.currents[k])
        out[f"{name}_currents_freq"] = fs[k]
    if rows_at is not None:
        t = time.time()
        out[f"{name}_rows"] = np.asarray(rows)
        out[f"{name}_Zrows"] = s.impedance(rows_at)[rows]
        out[f"{name}_Zrows_freq"] = rows_at
        print(f"{name} Z rows: {time.time() - t:.0f}s", flush=True)
    return out


def do_gcpw40():
    """The 12 428-unknown via-stitched pour: S at 1, 10 and 20 GHz, currents
    at 10 GHz, three RWG rows and two via rows of Z at 10 GHz."""
    lay, sub, mk, _ = W.gcpw40(NS)
    m = lay.mesh(**mk)
    nb, n = m.n_basis, m.n_unknowns
    out = _dense_sweep("gcpw40", W.gcpw40, np.array([1e9, 10e9, 20e9]), rows_at=10e9,
                       rows=[0, nb // 2, nb - 1, nb, n - 1], currents_at=10e9)
    np.savez_compressed(os.path.join(HERE, "config_gcpw40.npz"), **out)


def do_lpf7():
    """The stepped-impedance lowpass (8 837 unknowns): S at the sweep's first
    point and its middle, three rows of Z at the middle."""
    lay, sub, mk, f = W.lpf7(NS)
    n = lay.mesh(**mk).n_unknowns
    out = _dense_sweep("lpf7", W.lpf7, f[[0, 50]], rows_at=f[50], rows=[0, n // 2, n - 1])
    np.savez_compressed(os.path.join(HERE, "config_lpf7.npz"), **out)


def do_random():
    """Two of the 16 random chains (seeds 0 and 15) at two of their 64
    frequencies each, with two rows of Z at the second."""
    out = {}
    for seed in (0, 15):
        lay, sub, mk, f = W.random_chain(seed, NS)
        n = lay.mesh(**mk).n_unknowns
        out.update(_dense_sweep(f"random{seed}", lambda ns, s=seed: W.random_chain(s, ns), f[[0, 40]],
                                rows_at=f[40], rows=[0, n - 1]))
    np.savez_compressed(os.path.join(HERE, "config_random.npz"), **out)


def do_planes():
    """Stacks with an inner plane (the plane transition of
    Models/plane_transition): S of the reference's plain path for the via
    passing the plane without a modelled hole (with and without stitching
    vias) and through antipads of 0.4 / 0.7 / 1.2 mm; for the 0.7 mm antipad
    the antipad (frill) blocks C and F of the impedance matrix at 2 and
    15 GHz (antipad.FrillModel.blocks) and the via segments."""
    from mom25d.antipad import FrillModel

    out = {}
    cases = (("open", dict(antipad=None, stitch=False), [3.5e9, 11.5e9]),
             ("stitch", dict(antipad=None, stitch=True), [3.5e9, 11.5e9]),
             ("anti400", dict(antipad=0.4 * mm), [9.5e9, 15e9]),
             ("anti700", dict(antipad=0.7 * mm), [9.5e9, 15e9]),
             ("anti1200", dict(antipad=1.2 * mm), [9.5e9, 15e9]))
    for name, kw, fs in cases:
        out.update(_dense_sweep(f"plane_{name}", lambda ns, kw=kw: W.plane_transition(ns, **kw), np.array(fs)))
    lay, st, mk, _ = W.plane_transition(NS, antipad=0.7 * mm)
    m = lay.mesh(**mk)
    s = ref.Solver(m, st, verbose=False, method="dense")
    out["frill_segments"] = np.array([(g["top"], g["bottom"]) for g in m.via_segments])
    out["frill_meta"] = np.array([(fr["seg_up"], fr["seg_low"], fr["nr"]) for fr in m.frills])
    fm = FrillModel(s)
    for f in (2e9, 15e9):
        g = s.greens(f)
        C, F = fm.blocks(g, g.omega)
        out[f"frill_C_{f / 1e9:g}"] = C
        out[f"frill_F_{f / 1e9:g}"] = F
    np.savez_compressed(os.path.join(HERE, "config_planes.npz"), **out)


if __name__ == "__main__":
    sets = sys.argv[1:] or ["mesh", "tables", "z", "sweep", "random", "lpf7", "gcpw40"]
    for s in sets:
        t = time.time()
        globals()[f"do_{s}"]()
        print(f"{s}: {time.time() - t:.1f}s", flush=True)
