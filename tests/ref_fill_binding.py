"""The binding a maintainer of the reference would add to hand its fill to
libmom25d.so: plain ctypes over the reference's own objects, importing
nothing from this package.

  - GpuFill(lib_path, geom) flattens an AssemblyGeometry
    (Engine/mom25d/solver.py:307-426: V, qpts, qw, mpts, mw, fpts, fw, cA,
    cV, b, zl, blocks, pad_pos) into the plan entry point mom_plan_create;
  - set_greens(greens) passes every level pair's spline as _fast.spline_data
    gives it (Engine/mom25d/_fast.py:35-45: nodes x, PPoly coefficients c
    (4, n-1, 2) complex of [S_A, S_V]) with the pair's cA, cV and dz;
  - fill(omega, Z, IVpad) then stores what the loop of _fast.chunk_vals +
    scatter of assemble adds into the zeroed Z[:nb, :nb] and IVpad
    (Engine/mom25d/solver.py:472-483), MU0 / EPS0 being the reference's
    constants.

In the reference, assemble would call it where it branches on
_fast.ENABLED (solver.py:472) and keep the rest (Gram matrix, walls, via
blocks from IVpad, symmetrisation) as it is."""

import ctypes

import numpy as np

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


class _Geometry(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("nt", "nb", "n_levels", "n_blocks", "n_groups", "pad_")] + [
        (n, ctypes.c_void_p) for n in ("V", "qpts", "qw", "mpts", "mw", "fpts", "fw", "cA", "cV", "tri_basis", "zl",
                                       "blk_off", "blk_tris", "blk_layer", "grp_block", "grp_layer", "grp_tier",
                                       "grp_off", "grp_src", "pad_pos")]


class _FillCall(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("jwmu_re", "jwmu_im", "ijwe_re", "ijwe_im")] + [
        (n, ctypes.c_void_p) for n in ("pairs", "tab_x", "tab_c", "tab_b", "Z")] + [
        ("ldz", ctypes.c_int64), ("IVpad", ctypes.c_void_p), ("stage_max", ctypes.c_int32),
        ("tier_mask", ctypes.c_int32)]


_SIGS = {"mom_last_error": (ctypes.c_char_p, []), "mom_plan_struct_sizes": (_I, [_P]),
         "mom_plan_create": (_I, [_P, _I, _P, _P]), "mom_plan_destroy": (_I, [_P]),
         "mom_plan_set_pair": (_I, [_P, _I, _I, _I, _P, _P, _D, _D, _D, _D, _D, _I]),
         "mom_plan_fill_host": (_I, [_P, _P])}


class GpuFill:
    """Fill plan of one AssemblyGeometry (on_device=False: host part only)."""

    def __init__(self, lib_path, geom, on_device=True):
        L = ctypes.CDLL(lib_path)
        for name, (res, args) in _SIGS.items():
            getattr(L, name).restype, getattr(L, name).argtypes = res, args
        self.L = L
        sizes = (ctypes.c_int32 * 3)()
        self._call("mom_plan_struct_sizes", sizes)
        if sizes[0] != ctypes.sizeof(_Geometry) or sizes[1] != ctypes.sizeof(_FillCall):
            raise RuntimeError("libmom25d plan structures differ from this binding")
        self.nl = len(geom.zl)
        self.nb, self.nt = geom.nb, geom.nt
        c = np.ascontiguousarray
        blk_tris, blk_layer, gb, gl, gt, gs = [], [], [], [], [], []
        for k, (lt, pb, r, groups, wo, Xw) in enumerate(geom.blocks):
            blk_tris.append(pb)
            blk_layer.append(lt)
            for ls, sel, pts, w, static in groups:
                gb.append(k)
                gl.append(ls)
                gt.append(0 if static is not None else (1 if pts.shape[1] == 3 else 2))
                gs.append(sel)
        a = dict(V=c(geom.V, float), qpts=c(geom.qpts, float), qw=c(geom.qw, float), mpts=c(geom.mpts, float),
                 mw=c(geom.mw, float), fpts=c(geom.fpts, float), fw=c(geom.fw, float), cA=c(geom.cA, float),
                 cV=c(geom.cV, float), tri_basis=c(geom.b, np.int64), zl=c(geom.zl, float),
                 blk_off=c(np.cumsum([0] + [p.size for p in blk_tris]), np.int64),
                 blk_tris=c(np.concatenate(blk_tris), np.int64), blk_layer=c(blk_layer, np.int32),
                 grp_block=c(gb, np.int32), grp_layer=c(gl, np.int32), grp_tier=c(gt, np.int32),
                 grp_off=c(np.cumsum([0] + [s.size for s in gs]), np.int64),
                 grp_src=c(np.concatenate(gs) if gs else np.zeros(1), np.int64),
                 pad_pos=c(geom.pad_pos, np.int64) if getattr(geom, "nv", 0) else None)
        self._arrays = a
        g = _Geometry(nt=geom.nt, nb=geom.nb, n_levels=self.nl, n_blocks=len(blk_tris), n_groups=len(gs))
        for name, _ in _Geometry._fields_[6:]:
            setattr(g, name, None if a[name] is None else a[name].ctypes.data)
        self.n_pad = int(np.max(geom.pad_pos)) + 1 if a["pad_pos"] is not None else 0
        self.plan = ctypes.c_void_p()
        self._call("mom_plan_create", ctypes.byref(g), int(on_device), None, ctypes.byref(self.plan))

    def _call(self, name, *args):
        if getattr(self.L, name)(*args) != 0:
            raise RuntimeError(f"{name}: {self.L.mom_last_error().decode()}")

    def set_greens(self, greens):
        """Every level pair's spline (the layout of _fast.spline_data)."""
        for lt in range(self.nl):
            for ls in range(self.nl):
                pr = greens.pair(lt, ls)
                if getattr(pr, "zero", False):
                    self._call("mom_plan_set_pair", self.plan, lt, ls, 0, None, None, 0.0, 0.0, 0.0, 0.0, 0.0, 1)
                    continue
                x = np.ascontiguousarray(pr._spl.x, float)
                cf = np.ascontiguousarray(pr._spl.c, complex)
                assert cf.shape == (4, x.size - 1, 2)
                cA, cV = complex(pr.cA), complex(pr.cV)
                self._call("mom_plan_set_pair", self.plan, lt, ls, x.size, x.ctypes.data, cf.ctypes.data,
                           float(pr.dz), cA.real, cA.imag, cV.real, cV.imag, 0)

    def fill(self, omega, mu0, eps0, Z, IVpad=None):
        """Z[:nb, :nb] (a row-major complex128 array of row stride Z.shape[1])
        and IVpad (n_pad, nt) receive the fill at omega."""
        assert Z.dtype == np.complex128 and Z.flags.c_contiguous
        jwmu, ijwe = 1j * omega * mu0, 1 / (1j * omega * eps0)
        if self.n_pad:
            assert IVpad is not None and IVpad.shape == (self.n_pad, self.nt) and IVpad.flags.c_contiguous
        call = _FillCall(jwmu_re=jwmu.real, jwmu_im=jwmu.imag, ijwe_re=ijwe.real, ijwe_im=ijwe.imag, Z=Z.ctypes.data,
                         ldz=Z.shape[1], IVpad=IVpad.ctypes.data if self.n_pad else None)
        self._call("mom_plan_fill_host", self.plan, ctypes.byref(call))

    def close(self):
        if self.plan:
            self._call("mom_plan_destroy", self.plan)
            self.plan = ctypes.c_void_p()
