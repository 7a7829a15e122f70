"""Oracle: frequency-independent assembly set-up and the MPIE impedance
matrix, numpy path and C path (Engine/mom25d/solver.py:256-566,
Engine/mom25d/_fast.py:190-211).  Test infrastructure only."""

import ctypes
import os

import numpy as np
import scipy.sparse as sp

from .integrals_ref import TRI3_BARY, TRI3_W, TRI7_BARY, TRI7_W, points, wilton

MU0 = 4e-7 * np.pi
C0 = 299792458.0
EPS0 = 1.0 / (MU0 * C0 ** 2)

_HERE = os.path.dirname(os.path.abspath(__file__))
_CLIB = None


def clib():
    """The C restatement (oracle/liboracle_fill.so, built by oracle/Makefile)."""
    global _CLIB
    if _CLIB is None:
        path = os.path.join(_HERE, "liboracle_fill.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        _CLIB = ctypes.CDLL(path)
    return _CLIB


def morton(xy):
    """Stable argsort of Z-order codes (solver.py:294-304)."""
    q = ((xy - xy.min(0)) / max(np.ptp(xy, 0).max(), 1e-30) * 65535).astype(np.uint64)
    v = [q[:, 0], q[:, 1]]
    for k in range(2):
        x = v[k]
        for sh, mask in ((8, 0x00FF00FF), (4, 0x0F0F0F0F), (2, 0x33333333), (1, 0x55555555)):
            x = (x | (x << np.uint64(sh))) & np.uint64(mask)
        v[k] = x
    return np.argsort(v[0] | (v[1] << np.uint64(1)), kind="stable")


class Geometry:
    """Set-up of assemble for one mesh (solver.py:307-426)."""

    def __init__(self, mesh, block=32, near_factor=3.0, far_factor=10.0, zl=None, via_z=None, section=None):
        V, A = mesh.verts, mesh.area
        self.nt, self.nb = nt, nb = len(A), mesh.n_basis
        self.V = V
        self.qpts = qp = points(V, TRI7_BARY)
        self.qw = qw = TRI7_W[None, :] * A[:, None]
        mp, mw = points(V, TRI3_BARY), TRI3_W[None, :] * A[:, None]
        fp, fw = mesh.centroid[:, None, :], A[:, None]
        self.mpts, self.mw, self.fpts, self.fw = mp, mw, fp, fw  # attribute names of solver.py:345
        size = np.max(np.linalg.norm(V - np.roll(V, 1, axis=1), axis=-1), axis=1)
        cen, tl = mesh.centroid, mesh.tri_layer
        nl = len(mesh.layer_names)
        zl = np.zeros(nl) if zl is None else np.asarray(zl, float)
        self.zl = zl
        self.b = b = mesh.tri_basis
        lab = np.where(b >= 0, mesh.edge_len[np.maximum(b, 0)], 0.0)
        self.cA = cA = mesh.tri_sign * lab / (2 * A[:, None])
        self.cV = mesh.tri_sign * lab / A[:, None]
        ok = (b >= 0).ravel()
        self.PT = sp.csr_matrix((np.ones(ok.sum()), (np.nonzero(ok)[0], b.ravel()[ok])), shape=(3 * nt, nb)).T.tocsr()
        self.blocks = []
        for lt in range(nl):
            on = np.nonzero(tl == lt)[0]
            if not on.size:
                continue
            order = on[morton(cen[on])]
            for p0 in range(0, order.size, block):
                pb = order[p0:p0 + block]
                r = qp[pb].reshape(-1, 2)
                groups = []
                for ls in range(nl):
                    src = np.nonzero(tl == ls)[0]
                    if not src.size or (section is not None and section[ls] != section[lt]):
                        continue
                    dz = abs(zl[lt] - zl[ls])
                    dist = np.linalg.norm(cen[pb][:, None, :] - cen[None, src, :], axis=-1)
                    dist = np.sqrt(dist ** 2 + dz ** 2)
                    rel = np.min(dist / np.maximum(size[pb][:, None], size[None, src]), axis=0)
                    near = src[rel < near_factor]
                    mid = src[(rel >= near_factor) & (rel < far_factor)]
                    far = src[rel >= far_factor]
                    if near.size:
                        I0, I1 = wilton(V[near], r, dz)
                        static = (I0 / (4 * np.pi), (r[None] * I0[..., None] + I1) / (4 * np.pi))
                        groups.append((ls, near, qp[near], qw[near], static))
                    for sel, pts, w in ((mid, mp, mw), (far, fp, fw)):
                        if sel.size:
                            groups.append((ls, sel, pts[sel], w[sel], None))
                wo = qw[pb]
                Xw = (qp[pb][:, None, :, :] - V[pb][:, :, None, :]) * wo[:, None, :, None]
                self.blocks.append((lt, pb, r, groups, wo, Xw))
        X = qp[:, None, :, :] - V[:, :, None, :]
        G = np.einsum("tk,tikd,tjkd->tij", qw, X, X)
        ii = np.broadcast_to(b[:, :, None], G.shape)
        jj = np.broadcast_to(b[:, None, :], G.shape)
        ok = (ii >= 0) & (jj >= 0)
        self.gram = (ii[ok], jj[ok], (cA[:, :, None] * cA[:, None, :] * G)[ok])
        segs = getattr(mesh, "via_segments", [])
        self.nv = nv = len(segs)
        if nv:
            self.vias = [mesh.vias[g["via"]] for g in segs]
            tris = [g[e] for g in segs for e in ("top_tris", "bottom_tris") if g[e] is not None]
            self.pad_tris = pad = np.unique(np.concatenate(tris))
            self.pad_pos = -np.ones(nt, int)
            self.pad_pos[pad] = np.arange(pad.size)
            self.Dvia = np.zeros((pad.size, nv))
            self.via_ends = []
            for k, g in enumerate(segs):
                ends = []
                for key, sgn in (("top_tris", 1.0), ("bottom_tris", -1.0)):
                    t = g[key]
                    if t is not None:
                        self.Dvia[self.pad_pos[t], k] += sgn / A[t].sum()
                        ends.append((tl[t[0]], sgn))
                self.via_ends.append(ends)
            self.via_spans = [(via_z[g["bottom"]], via_z[g["top"]]) for g in segs] if via_z is not None else None
            ok = b >= 0
            self.Cdiv = sp.csr_matrix((self.cV[ok], (np.nonzero(ok)[0], b[ok])), shape=(nt, nb))


def _cmatmul(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return np.matmul(a.real, b) + 1j * np.matmul(a.imag, b)


def _potentials(pair, r, pts, w, static):
    """K0A, K1A, K0V of the sources at observers r (solver.py:265-291)."""
    d = r[None, :, None, :] - pts[:, None, :, :]
    R = np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2)
    SA, SV = pair.smooth(R)
    if static is not None:
        K0A, K1A, K0V = pair.cA * static[0], pair.cA * static[1], pair.cV * static[0]
    else:
        inv = 1 / (4 * np.pi * (R if not pair.dz else np.sqrt(R * R + pair.dz ** 2)))
        SA = SA + pair.cA * inv
        SV = SV + pair.cV * inv
        K0A = K1A = K0V = 0.0
    SAw = SA * w[:, None, :]
    return K0A + SAw.sum(-1), K1A + _cmatmul(SAw, pts), K0V + np.einsum("qml,ql->qm", SV, w)


def _c_chunk_vals(greens, lt, pb, r, groups, wo, V, cA, cV, nt, jwmu, ijwe):
    """vals (npb, 3, nt, 3) and IV (npb, nt) by the C restatement (_fast.py:190-211)."""
    L = clib()
    npb = pb.size
    vals = np.zeros((npb, 3, nt, 3), complex)
    IV = np.zeros((npb, nt), complex)
    Vt = np.ascontiguousarray(V[pb])
    cAt, cVt = np.ascontiguousarray(cA[pb]), np.ascontiguousarray(cV[pb])
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    for ls, sel, pts, w, static in groups:
        pr = greens.pair(lt, ls)
        if getattr(pr, "zero", False):
            continue
        spl = pr._spl
        x, c = np.ascontiguousarray(spl.x, float), np.ascontiguousarray(spl.c, complex)
        sel64 = np.ascontiguousarray(sel, np.int64)
        pts_, w_ = np.ascontiguousarray(pts, float), np.ascontiguousarray(w, float)
        Vs, cAs, cVs = (np.ascontiguousarray(a[sel]) for a in (V, cA, cV))
        if static is None:
            S0 = S1 = np.zeros(1, complex)
            st = 0
        else:
            S0, S1 = np.ascontiguousarray(static[0], complex), np.ascontiguousarray(static[1], complex)
            st = 1
        kA, kV = complex(pr.cA), complex(pr.cV)
        L.oracle_group_vals(ctypes.c_int(npb), ctypes.c_int(7), ctypes.c_int(nt), ctypes.c_int(sel.size),
                            ctypes.c_int(pts_.shape[1]), P(vals), P(IV), P(sel64), P(pts_), P(w_), P(Vs), P(cAs),
                            P(cVs), ctypes.c_int(st), P(S0), P(S1), P(np.ascontiguousarray(r)),
                            P(np.ascontiguousarray(wo)), P(Vt), P(cAt), P(cVt), P(x), P(c), ctypes.c_int(x.size),
                            ctypes.c_double(kA.real), ctypes.c_double(kA.imag), ctypes.c_double(kV.real),
                            ctypes.c_double(kV.imag), ctypes.c_double(float(pr.dz)), ctypes.c_double(jwmu.real),
                            ctypes.c_double(jwmu.imag), ctypes.c_double(ijwe.real), ctypes.c_double(ijwe.imag))
    return vals, IV


def assemble(mesh, greens, zs=0.0, geom=None, zs_via=None, stack=None, use_c=False):
    """Galerkin MPIE impedance matrix (solver.py:429-524); use_c selects the C
    restatement of the pair kernel and scatter instead of the numpy path."""
    geom = geom or Geometry(mesh)
    omega = greens.omega
    nt, nb, V, b = geom.nt, geom.nb, geom.V, geom.b
    cA, cV = geom.cA, geom.cV
    K = 7
    jwmu = 1j * omega * MU0
    ijwe = 1 / (1j * omega * EPS0)
    nv = geom.nv
    Z = np.zeros((nb + nv, nb + nv), complex)
    Zb = Z[:nb, :nb]
    if nv:
        IVpad = np.zeros((geom.pad_tris.size, nt), complex)
    for lt, pb, r, groups, wo, Xw in geom.blocks:
        npb = pb.size
        if use_c:
            vals, IV = _c_chunk_vals(greens, lt, pb, r, groups, wo, V, cA, cV, nt, jwmu, ijwe)
            Zc = np.ascontiguousarray(Zb) if not Zb.flags.c_contiguous else Zb
            rows = np.ascontiguousarray(b[pb], np.int64)
            cols = np.ascontiguousarray(b, np.int64)
            clib().oracle_scatter(ctypes.c_int(npb), ctypes.c_int(nt), Z.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.c_int64(Z.shape[1]), vals.ctypes.data_as(ctypes.c_void_p),
                                  rows.ctypes.data_as(ctypes.c_void_p), cols.ctypes.data_as(ctypes.c_void_p))
            del Zc
        else:
            K0A = np.zeros((nt, npb * K), complex)
            K1A = np.zeros((nt, npb * K, 2), complex)
            K0V = np.zeros((nt, npb * K), complex)
            for ls, sel, pts, w, static in groups:
                K0A[sel], K1A[sel], K0V[sel] = _potentials(greens.pair(lt, ls), r, pts, w, static)
            K0A = K0A.reshape(nt, npb, K).transpose(1, 0, 2)
            K1A = K1A.reshape(nt, npb, K, 2).transpose(1, 0, 2, 3)
            T1 = _cmatmul(K1A.reshape(npb, nt, 2 * K), Xw.reshape(npb, 3, 2 * K).transpose(0, 2, 1))
            T2 = _cmatmul(K0A, Xw.transpose(0, 2, 1, 3).reshape(npb, K, 6)).reshape(npb, nt, 3, 2)
            IA = T1[..., None] - np.einsum("pqid,qjd->pqij", T2, V, optimize=True)
            IV = np.einsum("qpk,pk->pq", K0V.reshape(nt, npb, K), wo)
            vals = (jwmu * cA[pb][:, None, :, None] * cA[None, :, None, :] * IA
                    + ijwe * cV[pb][:, None, :, None] * cV[None, :, None, :] * IV[:, :, None, None])
            H = vals.transpose(0, 2, 1, 3).reshape(npb * 3, nt * 3)
            HB = (geom.PT @ H.T).T
            rows = b[pb].ravel()
            ok = rows >= 0
            np.add.at(Zb, rows[ok], HB[ok])
        if nv:
            pos = geom.pad_pos[pb]
            sel = pos >= 0
            IVpad[pos[sel]] = IV[sel]
    if np.any(zs):
        ii, jj, g = geom.gram
        np.add.at(Z, (ii, jj), zs * g)
    if nv:
        Z[nb:, :nb], Z[nb:, nb:] = _via_blocks(geom, greens, IVpad, zs if zs_via is None else zs_via)
        Z[:nb, nb:] = Z[nb:, :nb].T
    return 0.5 * (Z + Z.T)


def _fil_z(a1, a2, b1, b2, d):
    f = lambda s: s * np.arcsinh(s / d) - np.sqrt(s * s + d * d)  # noqa: E731
    return MU0 / (4 * np.pi) * (f(a2 - b1) - f(a2 - b2) - f(a1 - b1) + f(a1 - b2))


def via_open_impedance(vias, spans, omega, zb, zt):
    """jw L of via pieces in one open section with images (solver.py:125-199, k None)."""
    xy = np.array([v.xy for v in vias], float).reshape(-1, 2)
    rho = np.linalg.norm(xy[:, None] - xy[None], axis=-1)
    rad = np.array([0.5 * v.drill for v in vias])
    rho = np.where(rho < 1e-12, np.minimum(rad[:, None], rad[None, :]), rho)
    np.fill_diagonal(rho, rad)
    spans = np.asarray(spans, float).reshape(-1, 2)
    a1 = np.clip(spans[:, 0], zb, zt) - zb
    a2 = np.clip(spans[:, 1], zb, zt) - zb
    p1, p2, q1, q2 = a1[:, None], a2[:, None], a1[None, :], a2[None, :]
    return 1j * omega * (_fil_z(p1, p2, q1, q2, rho) + _fil_z(p1, p2, -q2, -q1, rho))


def _via_blocks(geom, greens, IVpad, zs=0.0):
    """Via rows of Z (solver.py:527-566); via currents per stack section."""
    omega = greens.omega
    ijwe = 1 / (1j * omega * EPS0)
    D = geom.Dvia
    Y = (geom.Cdiv.T @ IVpad.T).T
    Zvb = ijwe * (D.T @ Y)
    Zvv = ijwe * (D.T @ IVpad[:, geom.pad_tris] @ D)
    im0 = {}
    for k, ek in enumerate(geom.via_ends):
        for m, em in enumerate(geom.via_ends):
            acc = 0.0
            for la, sa in ek:
                for lb, sb in em:
                    if (la, lb) not in im0:
                        pr = greens.pair(la, lb)
                        rho = 0.0 if pr.dz else 1e-6 * greens.h
                        im0[la, lb] = pr.gV(np.array([rho]))[0].imag
                    acc += sa * sb * im0[la, lb]
            Zvv[k, m] -= ijwe * 1j * acc
    h = greens.h
    secs = greens.via_sections() if hasattr(greens, "via_sections") else [(0.0, h, None)]
    for zb, zt, k in secs:
        if k is None:
            Zvv += via_open_impedance(geom.vias, geom.via_spans, omega, zb, zt)
        else:
            # a section closed by planes: the package's parallel-plate mode sum, pinned on its own by the
            # closed forms of tests/test_planes.py (radial-line limit, static image series)
            from mom25d_amd.solver import via_impedance

            Zvv += via_impedance(geom.vias, geom.via_spans, omega, [(zb, zt, k)])
    lengths = np.array([z1 - z0 for z0, z1 in geom.via_spans])
    if zs:
        Zvv += np.diag([zs * l / (np.pi * v.drill) for l, v in zip(lengths, geom.vias)])
    return Zvb, Zvv
