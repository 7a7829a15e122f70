"""Host-side pieces of the product that run without a GPU: the spline
solver, the table grid, the fill plan's tiers against the oracle's, and the
C-ABI library (loads, exports every declared symbol, struct layouts)."""

import os
import re

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from mom25d_amd import workloads as W
from mom25d_amd.greens import BUCKET_DTYPE, PAIR_DTYPE, bucket_table, ppoly_eval, rho_grid, spline_coeffs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_spline_matches_scipy_not_a_knot():
    rng = np.random.default_rng(1)
    x = np.cumsum(rng.uniform(0.1, 1.0, 300))
    Y = rng.standard_normal((300, 3)) + 1j * rng.standard_normal((300, 3))
    c = spline_coeffs(x, Y)
    ref = CubicSpline(x, Y).c
    assert np.max(np.abs(c - ref)) < 1e-9 * np.max(np.abs(ref))
    xs = np.linspace(x[0] - 0.5, x[-1] + 0.5, 1000)
    assert np.max(np.abs(ppoly_eval(x, c, xs) - CubicSpline(x, Y)(xs))) < 1e-9 * np.max(np.abs(Y))


def test_spline_reproduces_cubics_exactly():
    """A not-a-knot spline is exact on cubic polynomials."""
    x = np.sort(np.random.default_rng(2).uniform(0, 3, 40))
    p = lambda t: 1.5 - 2 * t + 0.7 * t ** 2 - 0.3 * t ** 3  # noqa: E731
    Y = (p(x) * (1 + 0.5j))[:, None]
    c = spline_coeffs(x, Y)
    xs = np.linspace(0, 3, 200)
    np.testing.assert_allclose(ppoly_eval(x, c, xs)[:, 0], p(xs) * (1 + 0.5j), atol=1e-11)


def test_rho_grid_matches_reference(golden):
    t = golden("tables.npz")
    from oracle.greens_ref import SlabGreens

    for f in (1e9, 7.3e9, 20e9):
        g = SlabGreens(f, 3.55, 0.508e-3, 0.0027, rho_max=0.05)
        np.testing.assert_array_equal(rho_grid(2 * np.pi / abs(g.k1), 0.508e-3, 0.05), t[f"slab_{f / 1e9:g}_rho"])


@pytest.mark.parametrize("grid", ["slab1", "slab20", "offset", "aligned"])
def test_bucket_table_reproduces_scipy_intervals(grid):
    """The fill kernels' interval lookup -- b = (int)(R * inv_dx) clamped,
    then i0 + (R >= split) of entry b, emulated here in the same double
    arithmetic -- equals scipy's clamped searchsorted(rho, R, 'right') - 1 at
    random distances, at every node and every bucket edge and their
    neighbouring doubles, and beyond both ends of the grid."""
    rng = np.random.default_rng(3)
    if grid == "offset":  # uneven grid that starts above 0
        rho = np.cumsum(rng.uniform(0.5, 3.0, 400)) * 1e-5 + 3e-6
    elif grid == "aligned":  # nodes within rounding of bucket edges
        rho = np.arange(400) * (1e-4 / 3)
    else:
        f = 1e9 if grid == "slab1" else 20e9
        rho = rho_grid(2 * np.pi / (2 * np.pi * f / 3e8 * np.sqrt(3.55)), 0.508e-3, 0.05)
    dxb = 0.5 * np.diff(rho).min()
    t = bucket_table(rho, dxb)
    assert t.dtype == BUCKET_DTYPE and BUCKET_DTYPE.itemsize == 16
    inv = 1.0 / dxb
    edges = np.arange(t.size) * dxb
    pts = [rng.uniform(0, 1.3 * rho[-1], 20000), rho, edges, [0.0, 2 * rho[-1]]]
    R = np.concatenate(pts)
    R = np.concatenate([R, np.nextafter(R, np.inf), np.nextafter(R, 0)])
    b = np.clip((R * inv).astype(np.int64), 0, t.size - 1)
    got = t["i0"][b] + (R >= t["split"][b])
    want = np.clip(np.searchsorted(rho, R, side="right") - 1, 0, rho.size - 2)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", ["tee", "via2"])
def test_plan_tiers_match_oracle_groups(name):
    from mom25d_amd.plan import FillPlan
    from oracle.assemble_ref import Geometry

    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    zl = np.array([(st.to_stackup() if hasattr(st, "to_stackup") else st).level_z(n) for n in m.layer_names])
    plan = FillPlan(m, zl=zl)
    geo = Geometry(m, zl=zl)
    assert len(geo.blocks) == plan.n_blocks
    nl = len(m.layer_names)
    for bi, (lt, pb, r, groups, wo, Xw) in enumerate(geo.blocks):
        np.testing.assert_array_equal(plan.blk_tris[bi, :pb.size], pb)
        oracle_sets = {}
        for ls, sel, pts, w, static in groups:
            tier = 0 if static is not None else (1 if pts.shape[1] == 3 else 2)
            oracle_sets[(ls, tier)] = set(sel.tolist())
        mine = {}
        for tier, items in plan.items.items():
            for b, pp, off, cnt, _coff, _nc in items[items[:, 0] == bi]:
                mine.setdefault((pp % nl, tier), set()).update(plan.src[off:off + cnt].tolist())
        assert mine == oracle_sets


@pytest.mark.parametrize("name", ["tee", "via2"])
def test_plan_scatter_maps_point_at_the_basis_functions(name):
    """Every block row and item column the fill kernel sums into maps back to
    the basis function of the triangle edge it stands for, and every
    accumulator fits the LDS budget."""
    from mom25d_amd.plan import FillPlan

    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    zl = np.array([(st.to_stackup() if hasattr(st, "to_stackup") else st).level_z(n) for n in m.layer_names])
    plan = FillPlan(m, zl=zl)
    tb = plan.tb
    for bi in range(plan.n_blocks):
        pb = plan.blk_tris[bi]
        pb = pb[pb >= 0]
        loc = plan.blk_rloc[bi, :pb.size]
        rows = plan.blk_rows[plan.blk_row_off[bi]:plan.blk_row_off[bi + 1]]
        assert np.all(np.diff(rows) > 0)
        np.testing.assert_array_equal(np.where(loc >= 0, rows[np.maximum(loc, 0)], -1), tb[pb])
    for tier, items in plan.items.items():
        for bi, pp, off, cnt, coff, nc in items:
            cols = plan.item_cols[coff:coff + nc]
            loc = plan.src_cloc[off:off + cnt]
            np.testing.assert_array_equal(np.where(loc >= 0, cols[np.maximum(loc, 0)], -1), tb[plan.src[off:off + cnt]])
            nr = plan.blk_row_off[bi + 1] - plan.blk_row_off[bi]
            assert 16 * nr * nc <= plan.acc_bytes[tier] <= plan.ACC_MAX_BYTES


@pytest.mark.parametrize("name", ["tee", "via2", "gcpw_pour"])
def test_plan_owned_entries_have_one_writer(name):
    """An entry the fill adds with a plain load and store (owned block row
    and owned item column) is reached by no other work item of its tier's
    launch: brute force over every (row, column) pair each item flushes.
    The flags are also not vacuous: a good share of the entries is owned."""
    from collections import Counter

    from mom25d_amd.plan import FillPlan

    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    zl = np.array([(st.to_stackup() if hasattr(st, "to_stackup") else st).level_z(n) for n in m.layer_names])
    plan = FillPlan(m, zl=zl)
    assert plan.row_own.shape == plan.blk_rows.shape and plan.col_own.shape == plan.item_cols.shape
    owned_total = flushed = 0
    for tier, items in plan.items.items():
        writers = Counter()
        owned = []
        for bi, pp, off, cnt, coff, nc in items:
            r = slice(plan.blk_row_off[bi], plan.blk_row_off[bi + 1])
            rows, cols = plan.blk_rows[r], plan.item_cols[coff:coff + nc]
            R, C = np.meshgrid(rows, cols, indexing="ij")
            keys = (R.astype(np.int64) * m.n_basis + C).ravel()
            writers.update(keys.tolist())
            own = (plan.row_own[r][:, None] & plan.col_own[coff:coff + nc][None, :]).ravel()
            owned.append(keys[own])
            flushed += keys.size
        owned = np.concatenate(owned) if owned else np.zeros(0, np.int64)
        assert all(writers[k] == 1 for k in owned.tolist()), tier
        owned_total += owned.size
    print(f"{name}: {owned_total / flushed:.2f} of the flushed entries owned")
    assert owned_total > 0.2 * flushed


@pytest.mark.parametrize("name", ["tee", "via2"])
def test_plan_item_bounds_hold_every_distance(name):
    """The distance bounds of every work item (from which the fill kernel
    stages spline intervals) contain the in-plane distance of every (test
    point, source point) pair of the item, by brute force over the pairs: a
    bound too tight would read the wrong interval."""
    from mom25d_amd.plan import FAR, MID, NEAR, FillPlan

    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    zl = np.array([(st.to_stackup() if hasattr(st, "to_stackup") else st).level_z(n) for n in m.layer_names])
    plan = FillPlan(m, zl=zl)
    src_pts = {NEAR: plan.qp7, MID: plan.qp3, FAR: plan.cen[:, None, :]}
    for tier, items in plan.items.items():
        bounds = plan.item_r[tier]
        assert bounds.shape == (len(items), 2)
        for (bi, pp, off, cnt, coff, nc), (lo, hi) in zip(items, bounds):
            pb = plan.blk_tris[bi]
            test = plan.qp7[pb[pb >= 0]].reshape(-1, 2)
            srcp = src_pts[tier][plan.src[off:off + cnt]].reshape(-1, 2)
            R = np.hypot(*(test[:, None, :] - srcp[None, :, :]).transpose(2, 0, 1))
            assert 0.0 <= lo <= R.min() and R.max() <= hi, (tier, bi, off)


def test_complex64_viability_prediction():
    """The mixed sweep's prediction that complex64 factors will not reach the
    target below a converged complex64 point: with a reference at 13.16 GHz
    that took 7 steps (contraction 10^(-15/7) per step, times (f_r / f)^2),
    1 GHz diverges and 2.14 GHz needs about 26 steps (> lu64_steps = 20),
    while 5 GHz needs about 11; no reference above f, or no lu64_steps,
    predicts nothing."""
    from mom25d_amd import Solver

    lay, st, mk, _ = W.tee()
    s = Solver(lay.mesh(**mk), st, verbose=False, precision="mixed", device="cpu", lu64_steps=20)
    assert not s._c64_predicted_poor(1e9)
    s._c64_refs = [(13.16e9, 7)]
    assert s._c64_predicted_poor(1e9) and s._c64_predicted_poor(2.14e9)
    assert not s._c64_predicted_poor(5e9) and not s._c64_predicted_poor(14e9)
    rho = 1e-15 ** (1 / 7) * (13.16 / 2.14) ** 2
    assert 25 < np.log(1e-15) / np.log(rho) < 27
    s._c64_refs.append((3e9, 30))  # the nearest reference above f is the one used
    assert s._c64_predicted_poor(2.5e9)
    s.lu64_steps = None
    assert not s._c64_predicted_poor(1e9)


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "mom25d.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int64_t|int)\s+(mom_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes

    from mom25d_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.SIGNATURES, n


def test_struct_layouts_match():
    import ctypes

    from mom25d_amd import _lib
    from mom25d_amd.plan import check_struct_sizes

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    L = _lib.lib()
    check_struct_sizes()
    assert L.mom_fill_args_size() > 0
    assert L.mom_pair_tab_size() == PAIR_DTYPE.itemsize
    assert L.mom_abi_version() == 1
    assert L.mom_getrf_work_bytes(64, 1000, 64) > 0


def test_error_path_reports_message():
    """A bad argument is rejected on the host with a message (no device call)."""
    from mom25d_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    with pytest.raises(RuntimeError, match="prec must be"):
        _lib.call("mom_cgemm_sub", 16, 4, 4, 4, None, 4, None, 4, None, 4, None)
    with pytest.raises(RuntimeError, match="ncol must be"):
        _lib.call("mom_sommerfeld", 1, None, None, None, None, None, None, None, None, None, None, 5, None, None)
    with pytest.raises(RuntimeError, match="tile must be"):
        _lib.call("mom_set_gemm_tile", 64, 32, 64)
    with pytest.raises(RuntimeError, match="block width must be"):
        _lib.call("mom_set_lu_block", 64, 300)
    with pytest.raises(RuntimeError, match="nslots must be"):
        _lib.call("mom_mlp_train_step", 100, 1, 1, 1, 1.0, 1, 1, 1, 1, None)


def test_gmres_givens_solves_hessenberg_least_squares():
    """The rotations of the GMRES host loop reduce a complex upper-Hessenberg
    least-squares problem min |beta e1 - H y| column by column; the
    triangular solve must match numpy's lstsq, and |g[k]| its residual."""
    from mom25d_amd.solver import givens_column

    rng = np.random.default_rng(3)
    m, nrhs = 12, 2
    H = rng.standard_normal((m + 1, m, nrhs)) + 1j * rng.standard_normal((m + 1, m, nrhs))
    for c in range(m):
        H[c + 2:, c] = 0
    beta = np.array([1.7, 0.4])
    g = np.zeros((m + 1, nrhs), complex)
    g[0] = beta
    cs = np.zeros((m, nrhs), complex)
    sn = np.zeros((m, nrhs), complex)
    R = np.zeros_like(H)
    for j in range(m):
        h = H[:j + 2, j].copy()
        givens_column(h, cs, sn, g, j)
        R[:j + 2, j] = h
    for c in range(nrhs):
        y = np.linalg.solve(np.triu(R[:m, :m, c]), g[:m, c])
        e1 = np.zeros(m + 1, complex)
        e1[0] = beta[c]
        ref, res, *_ = np.linalg.lstsq(H[:, :, c], e1, rcond=None)
        np.testing.assert_allclose(y, ref, rtol=1e-10, atol=1e-12)
        assert abs(abs(g[m, c]) - np.linalg.norm(e1 - H[:, :, c] @ ref)) < 1e-12


def test_fit_pool_does_not_rerun_unguarded_scripts(tmp_path):
    """The wave-fit workers are spawned processes; a caller script without a
    __main__ guard must run once, not once per worker."""
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = tmp_path / "unguarded.py"
    script.write_text(
        "import sys\n"
        f"sys.path.insert(0, {root!r})\n"
        "print('MAIN RUN', flush=True)\n"
        "import numpy as np\n"
        "from mom25d_amd.solver import fit_pool\n"
        "p = fit_pool(2)\n"
        "print('sum', p.submit(np.add, 1, 2).result() + p.submit(np.add, 3, 4).result(), flush=True)\n")
    out = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=300).stdout
    assert out.count("MAIN RUN") == 1 and "sum 10" in out


@pytest.mark.parametrize("count,lanes", [(0, 2), (1, 3), (10, 1), (51, 2), (64, 3), (7, 4), (20, 4), (100, 5), (3, 8)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_lane_cursor_works_outward_from_segment_middles(count, lanes, seed):
    """Every index once, under any interleaving of the lanes.  A lane starts
    at its segment's middle (the lowest segment: at its first index) and
    takes its own indices in order of distance from there, so that the
    points reusing its first factorisation stay close; once its segment is
    used up it takes other segments' outer indices, working inward one
    index at a time."""
    import random

    from mom25d_amd.solver import LaneCursor

    c = LaneCursor(count, lanes)
    nseg = min(lanes, count)
    cuts = np.round(np.arange(nseg + 1) / max(nseg, 1) * count).astype(int)
    rng = random.Random(seed * 1000 + count * 31 + lanes)
    seen, per_lane, active = [], {k: [] for k in range(lanes)}, list(range(lanes))
    while active:
        lane = rng.choice(active)
        i = c.take(lane)
        if i is None:
            active.remove(lane)
            continue
        seen.append(i)
        per_lane[lane].append(i)
    assert sorted(seen) == list(range(count))
    for lane, idx in per_lane.items():
        if lane >= nseg:
            continue
        lo, hi = cuts[lane], cuts[lane + 1] - 1
        mid = (lo + hi) // 2 if lane > 0 else lo  # the lowest segment runs upward from its first index
        own = [i for i in idx if lo <= i <= hi]
        assert mid not in own or own[0] == mid  # (others may reach the middle before a late lane starts)
        d = [abs(i - mid) + (i < mid) * 0.5 for i in own]  # above before below at equal distance
        assert d == sorted(d), (lane, idx)
        stolen = [i for i in idx if not lo <= i <= hi]
        runs = sum(int(abs(b - a) != 1) for a, b in zip(stolen, stolen[1:]))
        assert runs <= lanes * max(1, int(np.ceil(np.log2(count + 2))))
    # taken round robin, as the sweeps do, every lane opens at its segment's middle
    c = LaneCursor(count, lanes)
    first = [c.take(lane) for lane in range(lanes)]
    for lane in range(nseg):
        assert first[lane] == ((cuts[lane] + cuts[lane + 1] - 1) // 2 if lane > 0 else cuts[0])


@pytest.mark.parametrize("count,lanes", [(0, 2), (1, 3), (10, 1), (51, 2), (64, 3), (7, 4), (100, 5)])
def test_run_cursor_hands_out_runs_of_neighbours(count, lanes):
    """The cursor of multi-mesh batches: every index once; a lane's indices
    are consecutive (up or down) except where it splits another lane's run,
    so neighbouring frequencies share a lane."""
    import random

    from mom25d_amd.solver import RunCursor

    c = RunCursor(count, lanes)
    rng = random.Random(count * 31 + lanes)
    seen, per_lane, active = [], {k: [] for k in range(lanes)}, list(range(lanes))
    while active:
        lane = rng.choice(active)
        i = c.take(lane)
        if i is None:
            active.remove(lane)
            continue
        seen.append(i)
        per_lane[lane].append(i)
    assert sorted(seen) == list(range(count))
    # a lane jumps only when it splits another run (each split halves one)
    jumps = sum(int(abs(b - a) != 1) for idx in per_lane.values() for a, b in zip(idx, idx[1:]))
    assert jumps <= lanes * max(1, int(np.ceil(np.log2(count + 2))))
    if lanes <= 2:
        assert jumps <= 1


def test_given_matrix_symmetry_flag():
    """A given impedance matrix is marked symmetric only when it equals its
    transpose exactly (the mixed path then keeps its one-triangle GMRES
    product; otherwise the solve takes the complex128 LU)."""
    from mom25d_amd.solver import _GivenZ

    rng = np.random.default_rng(0)
    A = rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
    assert _GivenZ(A + A.T).symmetric
    assert not _GivenZ(A).symmetric
    B = A + A.T
    B[1, 3] += 1e-15
    assert not _GivenZ(B).symmetric


def _host_arrays(handle, names):
    import ctypes

    from mom25d_amd import _lib

    out = {}
    for name, dtype in names:
        p, n = ctypes.c_void_p(), ctypes.c_int64()
        _lib.call("mom_plan_host_array", handle, name.encode(), ctypes.byref(p), ctypes.byref(n))
        buf = (ctypes.c_char * max(n.value * np.dtype(dtype).itemsize, 1)).from_address(p.value)
        out[name] = np.frombuffer(buf, dtype)[:n.value].copy()
    return out


@pytest.mark.parametrize("name", ["tee", "via2"])
def test_reference_binding_builds_the_product_plan(name):
    """The reference-side binding (tests/ref_fill_binding.py), handed an
    AssemblyGeometry of the reference's shape (the oracle's restatement),
    makes the library build exactly the plan the product builds from its
    own set-up: same work items, distance bounds and scatter maps."""
    from mom25d_amd import _lib
    from mom25d_amd.plan import FillPlan
    from oracle.assemble_ref import Geometry
    from tests.ref_fill_binding import GpuFill

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    lay, st, mk, _ = getattr(W, name)()
    m = lay.mesh(**mk)
    zl = np.array([(st.to_stackup() if hasattr(st, "to_stackup") else st).level_z(n) for n in m.layer_names])
    plan = FillPlan(m, zl=zl)
    b = GpuFill(_lib.LIB_PATH, Geometry(m, zl=zl), on_device=False)
    try:
        names = [(f"items{t}", np.int32) for t in range(3)] + [(f"item_r{t}", np.float64) for t in range(3)]
        names += [(k, np.int32) for k in ("src", "blk_tris", "blk_rows", "blk_rloc", "item_cols", "src_cloc")]
        got = _host_arrays(b.plan, names + [("item_soff", np.int64)])
    finally:
        b.close()
    for t in range(3):
        np.testing.assert_array_equal(got[f"items{t}"].reshape(-1, 6), plan.items[t])
        np.testing.assert_array_equal(got[f"item_r{t}"].reshape(-1, 2), plan.item_r[t])
    for k in ("src", "blk_tris", "blk_rows", "blk_rloc", "item_cols", "src_cloc", "item_soff"):
        np.testing.assert_array_equal(got[k], getattr(plan, k).ravel())


def test_plan_rejects_bad_geometry():
    """Malformed AssemblyGeometry arrays are refused on the host with a
    message, before anything reaches a device."""
    import ctypes

    from mom25d_amd import _lib
    from mom25d_amd.plan import geometry_struct

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    lay, st, mk, _ = W.tiny()
    from mom25d_amd.plan import FillPlan

    good = FillPlan(lay.mesh(**mk)).geometry_arrays
    cases = {"triangle is in two test blocks": dict(blk_tris=np.zeros_like(good["blk_tris"])),
             "tri_basis entry outside": dict(tri_basis=good["tri_basis"] + good["nb"] + 5),
             "group block, layer or tier out of range": dict(grp_tier=good["grp_tier"] + 3),
             "source triangle outside": dict(grp_src=good["grp_src"] + good["nt"])}
    for msg, change in cases.items():
        a = dict(good, **change)
        g = geometry_struct(a)
        h = ctypes.c_void_p()
        with pytest.raises(RuntimeError, match=msg):
            _lib.call("mom_plan_create", ctypes.byref(g), 0, None, ctypes.byref(h))


def test_plan_own_tables_from_reference_splines():
    """Splines handed over pair by pair in the layout of _fast.spline_data
    (nodes x, PPoly coefficients (4, n-1, 2) complex) become the fill
    kernel's tables: the same nodes, each interval's coefficients as
    [power][kernel][re, im], the interval lookup of greens.bucket_table at
    half the smallest node spacing, and the pair records."""
    from mom25d_amd import _lib
    from mom25d_amd.greens import BUCKET_DTYPE, PAIR_DTYPE, bucket_table
    from oracle.solve_ref import OracleSolver
    from tests.ref_fill_binding import GpuFill

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libmom25d.so not built")
    lay, st, mk, _ = W.via2(max_cell=0.5e-3, edge_cell=0.15e-3)
    m = lay.mesh(**mk)
    o = OracleSolver(m, st)
    greens = o.greens(7e9)
    b = GpuFill(_lib.LIB_PATH, o.geometry(), on_device=False)
    try:
        b.set_greens(greens)
        got = _host_arrays(b.plan, [("tab_x", np.float64), ("tab_c", np.float64), ("tab_b", BUCKET_DTYPE),
                                    ("pairs", PAIR_DTYPE)])
    finally:
        b.close()
    nl = len(m.layer_names)
    xs, cs, bs = [], [], []
    for q, rec in enumerate(got["pairs"]):
        pr = greens.pair(q // nl, q % nl)
        x, c = pr._spl.x, pr._spl.c
        assert rec["x_off"] == sum(a.size for a in xs) and rec["c_off"] == sum(a.size for a in cs)
        assert rec["b_off"] == sum(a.size for a in bs)
        assert rec["n"] == x.size and rec["dz"] == pr.dz and rec["zero"] == 0
        assert (rec["cA_re"], rec["cA_im"], rec["cV_re"], rec["cV_im"]) == (complex(pr.cA).real, complex(pr.cA).imag,
                                                                          complex(pr.cV).real, complex(pr.cV).imag)
        dxb = 0.5 * np.diff(x).min()
        assert rec["inv_dx"] == 1.0 / dxb
        xs.append(x)
        cs.append(np.ascontiguousarray(c.transpose(1, 0, 2)).view(np.float64).reshape(-1))
        bs.append(bucket_table(x, dxb))
    np.testing.assert_array_equal(got["tab_x"], np.concatenate(xs))
    np.testing.assert_array_equal(got["tab_c"], np.concatenate(cs))
    want_b = np.concatenate(bs)
    np.testing.assert_array_equal(got["tab_b"]["i0"], want_b["i0"])
    np.testing.assert_array_equal(got["tab_b"]["split"], want_b["split"])
