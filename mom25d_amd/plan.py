"""Frequency-independent fill plan of one mesh (host set-up, device arrays).

Mirrors the set-up of Engine/mom25d/solver.py:294-426 (`_morton_order`,
`AssemblyGeometry`): quadrature points and area-scaled weights, RWG
coefficients cA = s l / (2A) and cV = s l / A, Morton-ordered test blocks of
32 triangles per metal layer, and the near / mid / far tier of every
(test block, source triangle) pair:

    rel = min over the block of |c_p - c_s|_3D / max(size_p, size_s)
    near: rel < 3,   mid: 3 <= rel < 10,   far: rel >= 10

The tier decides which source rule enters Z, so it is computed here with the
same floating-point expressions as the reference (the 3D distance is the
planar norm, squared again and combined with dz) and never re-derived on the
device.  The plan is then flattened into work items of <= 64 source
triangles for the GPU fill, and uploaded once.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .rules import TRI3_BARY, TRI3_W, TRI7_BARY, TRI7_W, quad_points

BLOCK = 32
CHUNK = 64
NEAR, MID, FAR = 0, 1, 2

_PTR_FIELDS = ("V", "qp7", "qw7", "qp3", "qw3", "cen", "area", "cA", "cV", "tb", "pad_pos", "blk_tris", "items",
               "src")


class FillArgs(ctypes.Structure):
    """Mirror of the FillArgs block of csrc/fill.hip (size checked at first use)."""

    _fields_ = [(n, ctypes.c_void_p) for n in _PTR_FIELDS] + [
        ("nt", ctypes.c_int32), ("pairs", ctypes.c_void_p), ("tab_x", ctypes.c_void_p), ("tab_c", ctypes.c_void_p),
        ("tab_b", ctypes.c_void_p), ("jwmu_re", ctypes.c_double), ("jwmu_im", ctypes.c_double),
        ("ijwe_re", ctypes.c_double), ("ijwe_im", ctypes.c_double), ("Z", ctypes.c_void_p), ("ldz", ctypes.c_int64),
        ("IVpad", ctypes.c_void_p), ("blk_row_off", ctypes.c_void_p), ("blk_rows", ctypes.c_void_p),
        ("blk_rloc", ctypes.c_void_p), ("item_cols", ctypes.c_void_p), ("src_cloc", ctypes.c_void_p),
        ("acc_bytes", ctypes.c_int64), ("item_r", ctypes.c_void_p), ("stage_max", ctypes.c_int32),
        ("pad2_", ctypes.c_int32), ("near_static", ctypes.c_void_p), ("item_soff", ctypes.c_void_p)]


def morton_order(xy):
    """Stable argsort of the Z-order (Morton) code of 16-bit quantised points."""
    span = max(np.ptp(xy, 0).max(), 1e-30)
    q = ((xy - xy.min(0)) / span * 65535).astype(np.uint64)

    def spread(v):
        for shift, mask in ((8, 0x00FF00FF), (4, 0x0F0F0F0F), (2, 0x33333333), (1, 0x55555555)):
            v = (v | (v << np.uint64(shift))) & np.uint64(mask)
        return v

    return np.argsort(spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)), kind="stable")


class FillPlan:
    """Host arrays of the plan; `upload()` makes the device copies."""

    def __init__(self, mesh, zl=None, near_factor=3.0, far_factor=10.0, block=BLOCK):
        V, A = mesh.verts, mesh.area
        nt = len(A)
        self.mesh = mesh
        self.nt, self.nb = nt, mesh.n_basis
        self.V = V
        self.area = A
        self.cen = mesh.centroid
        self.qp7 = quad_points(V, TRI7_BARY)
        self.qw7 = TRI7_W[None, :] * A[:, None]
        self.qp3 = quad_points(V, TRI3_BARY)
        self.qw3 = TRI3_W[None, :] * A[:, None]
        size = np.max(np.linalg.norm(V - np.roll(V, 1, axis=1), axis=-1), axis=1)
        self.size = size
        tl = mesh.tri_layer
        nl = len(mesh.layer_names)
        self.zl = np.zeros(nl) if zl is None else np.asarray(zl, float)
        b = mesh.tri_basis
        self.tb = b
        lab = np.where(b >= 0, mesh.edge_len[np.maximum(b, 0)], 0.0)
        self.cA = mesh.tri_sign * lab / (2 * A[:, None])
        self.cV = mesh.tri_sign * lab / A[:, None]

        blocks, blk_layer = [], []
        src_lists = []
        per_tier = {NEAR: [], MID: [], FAR: []}
        n_src = 0
        self.tier_pairs = np.zeros(3, np.int64)  # triangle pairs per tier
        cen = self.cen
        # Morton rank of every triangle within its layer: the source lists are
        # chunked in this order, so that a chunk is a compact patch whose
        # edges (the columns it adds into) are few and close together.
        mrank = np.zeros(nt, np.int64)
        for ls in range(nl):
            on_s = np.nonzero(tl == ls)[0]
            if on_s.size:
                mrank[on_s[morton_order(cen[on_s])]] = np.arange(on_s.size)
        for lt in range(nl):
            on_t = np.nonzero(tl == lt)[0]
            if not on_t.size:
                continue
            order = on_t[morton_order(cen[on_t])]
            for p0 in range(0, order.size, block):
                pb = order[p0:p0 + block]
                bi = len(blocks)
                blocks.append(pb)
                blk_layer.append(lt)
                for ls in range(nl):
                    src = np.nonzero(tl == ls)[0]
                    if not src.size:
                        continue
                    dz = abs(self.zl[lt] - self.zl[ls])
                    dist = np.linalg.norm(cen[pb][:, None, :] - cen[None, src, :], axis=-1)
                    dist = np.sqrt(dist ** 2 + dz ** 2)
                    rel = np.min(dist / np.maximum(size[pb][:, None], size[None, src]), axis=0)
                    tiers = (src[rel < near_factor], src[(rel >= near_factor) & (rel < far_factor)],
                             src[rel >= far_factor])
                    for tier, sel in zip((NEAR, MID, FAR), tiers):
                        if not sel.size:
                            continue
                        sel = sel[np.argsort(mrank[sel], kind="stable")]
                        self.tier_pairs[tier] += sel.size * pb.size
                        src_lists.append(sel.astype(np.int32))
                        for c0 in range(0, sel.size, CHUNK):
                            cnt = min(CHUNK, sel.size - c0)
                            per_tier[tier].append((bi, lt * nl + ls, n_src + c0, cnt))
                        n_src += sel.size
        self.n_blocks = len(blocks)
        bt = -np.ones((len(blocks), block), np.int32)
        for i, pb in enumerate(blocks):
            bt[i, :pb.size] = pb
        self.blk_tris = bt
        self.blk_layer = np.array(blk_layer, np.int32)
        self.blocks = blocks
        self.src = np.concatenate(src_lists) if src_lists else np.zeros(0, np.int32)
        self.n_levels = nl
        self._scatter_maps(blocks, per_tier, block)
        self._item_bounds(blocks)
        # near work items: offsets of their closed-form statics (test triangles x 7 points x 3 values x sources)
        it = self.items[NEAR].astype(np.int64).reshape(-1, 6)
        ntest = np.array([pb.size for pb in blocks], np.int64)[it[:, 0]]
        size = 21 * ntest * it[:, 3]
        self.item_soff = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.int64)
        self.near_static_len = int(size.sum())

        # conductor-loss Gram matrix per triangle, times cA_i cA_j
        X = self.qp7[:, None, :, :] - V[:, :, None, :]
        G = np.einsum("tk,tikd,tjkd->tij", self.qw7, X, X)
        self.gram = (self.cA[:, :, None] * self.cA[:, None, :] * G).reshape(nt, 9)

        # via segments: pad triangles and the drain matrix D (pad x segment)
        segs = getattr(mesh, "via_segments", [])
        self.nv = nv = len(segs)
        self.pad_pos = -np.ones(nt, np.int32)
        if nv:
            self.vias = [mesh.vias[g["via"]] for g in segs]
            tris = [g[e] for g in segs for e in ("top_tris", "bottom_tris") if g[e] is not None]
            pad = np.unique(np.concatenate(tris)) if tris else np.zeros(0, int)
            self.pad_tris = pad
            self.pad_pos[pad] = np.arange(pad.size)
            D = np.zeros((pad.size, nv))
            self.via_ends = []
            for k, g in enumerate(segs):
                ends = []
                t = g["top_tris"]
                if t is not None:
                    D[self.pad_pos[t], k] += 1 / A[t].sum()
                    ends.append((tl[t[0]], 1.0))
                t = g["bottom_tris"]
                if t is not None:
                    D[self.pad_pos[t], k] -= 1 / A[t].sum()
                    ends.append((tl[t[0]], -1.0))
                self.via_ends.append(ends)
            self.Dvia = D
            # CSR by segment
            off, rows, vals = [0], [], []
            for k in range(nv):
                nz = np.nonzero(D[:, k])[0]
                rows.extend(nz.tolist())
                vals.extend(D[nz, k].tolist())
                off.append(len(rows))
            self.d_off = np.array(off, np.int32)
            self.d_pad = np.array(rows, np.int32)
            self.d_val = np.array(vals, float)
        else:
            self.pad_tris = np.zeros(0, int)
        self.dev = None

    # LDS budget of one fill workgroup's accumulator (complex128 entries)
    ACC_MAX_BYTES = 136 * 1024

    def _scatter_maps(self, blocks, per_tier, block):
        """Distinct rows (test basis functions) of every block and columns
        (source basis functions) of every work item, with each triangle edge's
        position in them: the fill kernel sums a work item's contributions per
        (row, column) in LDS and adds every entry to Z once.  A work item whose
        accumulator would exceed ACC_MAX_BYTES is split."""
        tb = self.tb
        nb = len(blocks)
        row_off = np.zeros(nb + 1, np.int64)
        rows, nrow = [], np.zeros(nb, np.int64)
        rloc = -np.ones((nb, block, 3), np.int32)
        for bi, pb in enumerate(blocks):
            e = tb[pb]
            u = np.unique(e[e >= 0])
            loc = np.searchsorted(u, np.maximum(e, 0)).astype(np.int32)
            loc[e < 0] = -1
            rloc[bi, :pb.size] = loc
            rows.append(u)
            nrow[bi] = u.size
            row_off[bi + 1] = row_off[bi] + u.size
        self.blk_row_off = row_off.astype(np.int32)
        self.blk_rows = np.concatenate(rows).astype(np.int32) if rows else np.zeros(0, np.int32)
        self.blk_rloc = rloc
        src_cloc = -np.ones((self.src.size, 3), np.int32)
        cols, col_off = [], 0
        self.items, self.acc_bytes = {}, {}
        for tier, lst in per_tier.items():
            out, acc = [], 0
            todo = list(lst)
            while todo:
                bi, pp, so, cnt = todo.pop(0)
                e = tb[self.src[so:so + cnt]]
                u = np.unique(e[e >= 0])
                nbytes = 16 * int(nrow[bi]) * u.size
                if nbytes > self.ACC_MAX_BYTES and cnt > 1:
                    h = cnt // 2
                    todo[:0] = [(bi, pp, so, h), (bi, pp, so + h, cnt - h)]
                    continue
                loc = np.searchsorted(u, np.maximum(e, 0)).astype(np.int32)
                loc[e < 0] = -1
                src_cloc[so:so + cnt] = loc
                out.append((bi, pp, so, cnt, col_off, u.size))
                cols.append(u)
                col_off += u.size
                acc = max(acc, nbytes)
            self.items[tier] = np.array(out, np.int32).reshape(-1, 6)
            self.acc_bytes[tier] = acc
        self.item_cols = np.concatenate(cols).astype(np.int32) if cols else np.zeros(0, np.int32)
        self.src_cloc = src_cloc

    def _item_bounds(self, blocks):
        """Per work item, bounds [lo, hi] on the in-plane distance R between
        any test point of its block (7-point rule) and any source point of
        its tier (7-point rule near, 3-point mid, centroid far): from the
        bounding boxes' centres and half-diagonals, padded by 1e-9 relative so
        that the kernel's own rounding of R stays inside.  The fill kernel
        stages the spline intervals of [lo, hi] in LDS."""
        def circles(pts, counts):
            # pts: (sum(counts), 2) grouped by item; centre and radius of each group's bounding box
            starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
            lo = np.minimum.reduceat(pts, starts, axis=0)
            hi = np.maximum.reduceat(pts, starts, axis=0)
            return 0.5 * (lo + hi), 0.5 * np.hypot(*(hi - lo).T)

        tp = [self.qp7[pb].reshape(-1, 2) for pb in blocks]
        bc, br = circles(np.concatenate(tp), [t.shape[0] for t in tp]) if tp else (np.zeros((0, 2)), np.zeros(0))
        src_pts = {NEAR: self.qp7, MID: self.qp3, FAR: self.cen[:, None, :]}
        self.item_r = {}
        for tier, it in self.items.items():
            if not it.size:
                self.item_r[tier] = np.zeros((0, 2))
                continue
            so, cnt = it[:, 2].astype(np.int64), it[:, 3].astype(np.int64)
            idx = np.repeat(so - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt) + np.arange(cnt.sum())
            pts = src_pts[tier][self.src[idx]].reshape(-1, 2)
            sc, sr = circles(pts, cnt * src_pts[tier].shape[1])
            D = np.hypot(*(sc - bc[it[:, 0]]).T)
            r = br[it[:, 0]] + sr
            lo = np.maximum(0.0, (D - r) * (1 - 1e-9) - 1e-12)
            self.item_r[tier] = np.stack([np.where(D - r > 0, lo, 0.0), (D + r) * (1 + 1e-9) + 1e-12], 1)

    def upload(self, device="cuda"):
        """Device copies of every array the kernels read (kept on the plan)."""
        if self.dev is not None:
            return self.dev
        dev = torch.device(device)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)  # noqa: E731
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)  # noqa: E731
        m = self.mesh
        d = dict(
            V=f64(self.V), qp7=f64(self.qp7), qw7=f64(self.qw7), qp3=f64(self.qp3), qw3=f64(self.qw3),
            cen=f64(self.cen), area=f64(self.area), cA=f64(self.cA), cV=f64(self.cV), tb=i32(self.tb),
            pad_pos=i32(self.pad_pos), blk_tris=i32(self.blk_tris), src=i32(self.src if self.src.size else [0]),
            gram=f64(self.gram), tp=i32(m.tp), vp=i32(m.vp), tm=i32(m.tm), vm=i32(m.vm), tri_layer=i32(m.tri_layer),
        )
        for t, it in self.items.items():
            d[f"items{t}"] = i32(it if it.size else np.zeros((1, 6)))
            d[f"item_r{t}"] = f64(self.item_r[t] if self.item_r[t].size else np.zeros((1, 2)))
        d.update(blk_row_off=i32(self.blk_row_off), blk_rows=i32(self.blk_rows if self.blk_rows.size else [0]),
                 blk_rloc=i32(self.blk_rloc), item_cols=i32(self.item_cols if self.item_cols.size else [0]),
                 src_cloc=i32(self.src_cloc if self.src_cloc.size else np.zeros((1, 3))))
        if self.nv:
            d.update(d_off=i32(self.d_off), d_pad=i32(self.d_pad), d_val=f64(self.d_val),
                     pad_tri=i32(self.pad_tris))
        d["item_soff"] = torch.from_numpy(self.item_soff if self.item_soff.size else np.zeros(1, np.int64)).to(dev)
        d["near_static"] = torch.empty(max(self.near_static_len, 1), dtype=torch.float64, device=dev)
        n_near = int(self.items[NEAR].shape[0])
        if n_near:
            a = self.args(d)
            a.items = d["items0"].data_ptr()
            pair_dz = f64(np.abs(self.zl[:, None] - self.zl[None, :]).ravel())
            with torch.cuda.device(dev):
                _lib.call("mom_near_statics", n_near, ctypes.byref(a), pair_dz.data_ptr(), d["near_static"].data_ptr(),
                          _lib.stream_ptr())
                torch.cuda.current_stream().synchronize()
        self.dev = d
        return d

    def args(self, d):
        """A FillArgs block with the plan's device arrays `d` (upload()) set;
        the caller adds the frequency data, Z and the work items of a tier."""
        a = FillArgs()
        for name in ("V", "qp7", "qw7", "qp3", "qw3", "cen", "area", "cA", "cV", "tb", "pad_pos", "blk_tris", "src",
                     "blk_row_off", "blk_rows", "blk_rloc", "item_cols", "src_cloc", "near_static", "item_soff"):
            setattr(a, name, d[name].data_ptr())
        a.nt = self.nt
        return a

    # Algorithmic FP64 flops per triangle pair, counted from the pair
    # kernel's loop body with sqrt and divide as one flop each: per test point
    # 115 (outer 3x3 accumulation) plus per source point 59 (distance,
    # spline, 1/R term, inner sums; 49 without the 1/R term for near pairs),
    # plus 180 for the final scaling; near pairs add their closed-form static
    # part (precomputed once per mesh in the reference), counted as in the
    # reference's per-frequency work.
    FLOPS_PER_PAIR = {NEAR: 7 * (7 * 49 + 115) + 180, MID: 7 * (3 * 59 + 115) + 180, FAR: 7 * (59 + 115) + 180}

    def algorithmic_flops(self):
        """Algorithmic FP64 flops of one fill (tier pair counts x flops per pair)."""
        return float(sum(self.FLOPS_PER_PAIR[t] * self.tier_pairs[t] for t in (NEAR, MID, FAR)))
