"""Frequency-independent fill plan of one mesh.

Mirrors the set-up of Engine/mom25d/solver.py:294-426 (`_morton_order`,
`AssemblyGeometry`): quadrature points and area-scaled weights, RWG
coefficients cA = s l / (2A) and cV = s l / A, Morton-ordered test blocks of
32 triangles per metal layer, and the near / mid / far tier of every
(test block, source triangle) pair of levels in one stack section (a solid
plane between two levels leaves their pairs out):

    rel = min over the block of |c_p - c_s|_3D / max(size_p, size_s)
    near: rel < 3,   mid: 3 <= rel < 10,   far: rel >= 10

The tier decides which source rule enters Z, so it is computed here with the
same floating-point expressions as the reference (the 3D distance is the
planar norm, squared again and combined with dz) and never re-derived on the
device.  These arrays -- the reference's AssemblyGeometry, flattened -- go
through the library's plan entry point (mom_plan_create), which turns them
into work items of <= 64 source triangles for the GPU fill, uploads them
and computes the near tier's closed forms; the same entry point serves a
caller holding the reference's own AssemblyGeometry.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .rules import TRI3_BARY, TRI3_W, TRI7_BARY, TRI7_W, quad_points

BLOCK = 32
CHUNK = 64
NEAR, MID, FAR = 0, 1, 2

class MomAssemblyGeometry(ctypes.Structure):
    """Mirror of MomAssemblyGeometry (include/mom25d.h)."""

    _fields_ = [(n, ctypes.c_int32) for n in ("nt", "nb", "n_levels", "n_blocks", "n_groups", "pad_")] + [
        (n, ctypes.c_void_p) for n in ("V", "qpts", "qw", "mpts", "mw", "fpts", "fw", "cA", "cV", "tri_basis", "zl",
                                       "blk_off", "blk_tris", "blk_layer", "grp_block", "grp_layer", "grp_tier",
                                       "grp_off", "grp_src", "pad_pos")]


class MomFillCall(ctypes.Structure):
    """Mirror of MomFillCall (include/mom25d.h)."""

    _fields_ = [(n, ctypes.c_double) for n in ("jwmu_re", "jwmu_im", "ijwe_re", "ijwe_im")] + [
        (n, ctypes.c_void_p) for n in ("pairs", "tab_x", "tab_c", "tab_b", "Z")] + [
        ("ldz", ctypes.c_int64), ("IVpad", ctypes.c_void_p), ("stage_max", ctypes.c_int32),
        ("tier_mask", ctypes.c_int32)]


class MomPlanInfo(ctypes.Structure):
    """Mirror of MomPlanInfo (include/mom25d.h)."""

    _fields_ = [("n_items", ctypes.c_int32 * 3), ("n_pad", ctypes.c_int32), ("acc_bytes", ctypes.c_int64 * 3),
                ("tier_pairs", ctypes.c_int64 * 3), ("near_static_len", ctypes.c_int64), ("n_src", ctypes.c_int64)]


def check_struct_sizes():
    """The three plan blocks have the library's sizes."""
    out = (ctypes.c_int32 * 3)()
    _lib.call("mom_plan_struct_sizes", out)
    want = [ctypes.sizeof(t) for t in (MomAssemblyGeometry, MomFillCall, MomPlanInfo)]
    if list(out) != want:
        raise RuntimeError(f"libmom25d plan structures differ from the Python mirrors: {list(out)} != {want}")


def geometry_struct(arrays):
    """A MomAssemblyGeometry over the dict of contiguous numpy arrays
    `arrays` (the caller keeps them alive while the struct is used)."""
    g = MomAssemblyGeometry()
    for k in ("nt", "nb", "n_levels", "n_blocks", "n_groups"):
        setattr(g, k, int(arrays[k]))
    for k, _ in MomAssemblyGeometry._fields_[6:]:
        a = arrays.get(k)
        setattr(g, k, None if a is None else a.ctypes.data)
    return g


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
    """Host arrays of the plan; `upload()` makes the device plan."""

    # LDS budget of one fill workgroup's accumulator (complex128 entries),
    # as the library splits work items
    ACC_MAX_BYTES = 136 * 1024

    def __init__(self, mesh, zl=None, near_factor=3.0, far_factor=10.0, block=BLOCK, section=None):
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

        # test blocks and, per block and source layer, the tiers' source lists
        blocks, blk_layer = [], []
        grp_block, grp_layer, grp_tier, grp_src = [], [], [], []
        cen = self.cen
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
                    if not src.size or (section is not None and section[ls] != section[lt]):
                        continue
                    dz = abs(self.zl[lt] - self.zl[ls])
                    dist = np.linalg.norm(cen[pb][:, None, :] - cen[None, src, :], axis=-1)
                    dist = np.sqrt(dist ** 2 + dz ** 2)
                    rel = np.min(dist / np.maximum(size[pb][:, None], size[None, src]), axis=0)
                    tiers = (src[rel < near_factor], src[(rel >= near_factor) & (rel < far_factor)],
                             src[rel >= far_factor])
                    for tier, sel in zip((NEAR, MID, FAR), tiers):
                        if sel.size:
                            grp_block.append(bi)
                            grp_layer.append(ls)
                            grp_tier.append(tier)
                            grp_src.append(sel)
        self.n_blocks = len(blocks)
        self.blocks = blocks
        self.n_levels = nl

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

        # the reference's AssemblyGeometry, flattened for mom_plan_create
        c64 = lambda a, t=np.float64: np.ascontiguousarray(a, t)  # noqa: E731
        grp_off = np.concatenate([[0], np.cumsum([g.size for g in grp_src])]).astype(np.int64)
        self.geometry_arrays = dict(
            nt=nt, nb=self.nb, n_levels=nl, n_blocks=len(blocks), n_groups=len(grp_src),
            V=c64(V), qpts=c64(self.qp7), qw=c64(self.qw7), mpts=c64(self.qp3), mw=c64(self.qw3), fpts=c64(cen),
            fw=c64(A), cA=c64(self.cA), cV=c64(self.cV), tri_basis=c64(b, np.int64), zl=c64(self.zl),
            blk_off=np.concatenate([[0], np.cumsum([pb.size for pb in blocks])]).astype(np.int64),
            blk_tris=c64(np.concatenate(blocks), np.int64), blk_layer=c64(blk_layer, np.int32),
            grp_block=c64(grp_block, np.int32), grp_layer=c64(grp_layer, np.int32),
            grp_tier=c64(grp_tier, np.int32), grp_off=grp_off,
            grp_src=c64(np.concatenate(grp_src) if grp_src else np.zeros(1), np.int64),
            pad_pos=c64(self.pad_pos, np.int64) if nv else None)
        self._read_host_plan()

        # conductor-loss Gram matrix per triangle, times cA_i cA_j
        X = self.qp7[:, None, :, :] - V[:, :, None, :]
        G = np.einsum("tk,tikd,tjkd->tij", self.qw7, X, X)
        self.gram = (self.cA[:, :, None] * self.cA[:, None, :] * G).reshape(nt, 9)
        self.dev = None
        self.handle = None

    def _create(self, on_device, stream=None):
        check_struct_sizes()
        g = geometry_struct(self.geometry_arrays)
        h = ctypes.c_void_p()
        _lib.call("mom_plan_create", ctypes.byref(g), int(on_device), stream, ctypes.byref(h))
        return h

    def _read_host_plan(self):
        """The library's work items, scatter maps and bounds, read from a
        host-only plan (also what the device plan holds)."""
        h = self._create(False)
        try:
            info = MomPlanInfo()
            _lib.call("mom_plan_info", h, ctypes.byref(info))

            def arr(name, dtype, shape=(-1,)):
                p, n = ctypes.c_void_p(), ctypes.c_int64()
                _lib.call("mom_plan_host_array", h, name.encode(), ctypes.byref(p), ctypes.byref(n))
                if n.value == 0:
                    return np.zeros((0,) + tuple(shape[1:]), dtype)
                buf = (ctypes.c_char * (n.value * np.dtype(dtype).itemsize)).from_address(p.value)
                return np.frombuffer(buf, dtype).copy().reshape(shape)

            self.items = {t: arr(f"items{t}", np.int32, (-1, 6)) for t in (NEAR, MID, FAR)}
            self.item_r = {t: arr(f"item_r{t}", np.float64, (-1, 2)) for t in (NEAR, MID, FAR)}
            self.src = arr("src", np.int32)
            self.blk_tris = arr("blk_tris", np.int32, (-1, BLOCK))
            self.blk_row_off = arr("blk_row_off", np.int32)
            self.blk_rows = arr("blk_rows", np.int32)
            self.blk_rloc = arr("blk_rloc", np.int32, (-1, BLOCK, 3))
            self.item_cols = arr("item_cols", np.int32)
            self.src_cloc = arr("src_cloc", np.int32, (-1, 3))
            self.item_soff = arr("item_soff", np.int64)
            self.row_own = arr("row_own", np.uint8).astype(bool)
            self.col_own = arr("col_own", np.uint8).astype(bool)
            self.acc_bytes = {t: int(info.acc_bytes[t]) for t in (NEAR, MID, FAR)}
            self.tier_pairs = np.array(list(info.tier_pairs), np.int64)
            self.near_static_len = int(info.near_static_len)
        finally:
            _lib.call("mom_plan_destroy", h)

    def upload(self, device="cuda"):
        """The device plan (mom_plan_create) and device copies of the arrays
        the other kernels read (Gram matrix, via drains, line voltages)."""
        if self.dev is not None:
            return self.dev
        dev = torch.device(device)
        f64 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(dev)  # noqa: E731
        i32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)  # noqa: E731
        m = self.mesh
        d = dict(V=f64(self.V), qp7=f64(self.qp7), qw7=f64(self.qw7), tb=i32(self.tb), cV=f64(self.cV),
                 gram=f64(self.gram), tp=i32(m.tp), vp=i32(m.vp), tm=i32(m.tm), vm=i32(m.vm),
                 tri_layer=i32(m.tri_layer))
        if self.nv:
            d.update(d_off=i32(self.d_off), d_pad=i32(self.d_pad), d_val=f64(self.d_val), pad_tri=i32(self.pad_tris))
        with torch.cuda.device(dev):
            self.handle = self._create(True, _lib.stream_ptr())
        self.dev = d
        return d

    def fill(self, call, stream, tier_mask=0):
        """mom_plan_fill of the device plan with a MomFillCall."""
        call.tier_mask = tier_mask
        _lib.call("mom_plan_fill", self.handle, ctypes.byref(call), stream)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.call("mom_plan_destroy", h)
            except Exception:
                pass
            self.handle = None

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
