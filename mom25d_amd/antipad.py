"""Antipads: a via passing a solid inner plane through a clearance hole
(the model of Engine/mom25d/antipad.py:1-384, restated).

The hole between the barrel (radius a) and the plate edge (radius b) is
closed by PEC carrying magnetic current M = E x n on its two faces, for an
axisymmetric aperture field E_r = -dU/dr.  U(r) = sum_j V_j w_j(r) with
nr radial functions on nodes r_i = a (b / a)^(i / nr): w_0 = 1 inside the
barrel falling linearly to 0 at r_1, w_j (j >= 1) a tent r_(j-1) .. r_(j+1);
V_0 is the barrel-to-plane voltage.  The via is split at the plane (no
charge there).  Continuity of tangential H across the aperture, tested with
each w_j (Ampere around the axis), gives per ring j

    sum_m C_mj I_m + j w sum_l Cff_jl V_l = 0,

where C holds -1 / +1 for the via segment above / below the plate (the
current through the loop) plus j w times the plate charge that unknown m
induces, weighted by w_j: for RWG and via charges on the levels of the
section, through the section's TM transmission line; for the vertical via
currents of a section closed by planes, through their parallel-plate modes.
Cff are the ring capacitances Int eps0 k^2 y_in(k) w~_j(k) w~_l(k)
J0(k d) k dk / 2 pi over the Sommerfeld path of each side, y_in the TM input
admittance of the section seen from the plate (with the transfer admittance
to the far plate between antipads on the two plates of one closed section).
The impedance matrix then reads [[Z, -C], [-C^T, -(F + F^T) / 2]] with
F = j w Cff (symmetric; the frill rows negated).
"""

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import hankel2, j0, j1, jv, kv, struve

from .constants import C0, EPS0, MU0
from .greens import gauss_panels, sommerfeld_path
from .layered import _interface_voltages, csqrt_re_pos

_GX, _GW = np.polynomial.legendre.leggauss(24)


def ring_nodes(a, b, nr):
    """Geometric radii a .. b of the nr ring functions."""
    return a * (b / a) ** (np.arange(nr + 1) / nr)


def ring_weight(j, r, nodes):
    """Ring function j at radii r (j = 0: 1 inside the barrel)."""
    r = np.asarray(r, float)
    lo = nodes[j - 1] if j > 0 else 0.0
    mid, hi = nodes[j], nodes[j + 1]
    rise = (r - lo) / (mid - lo) if j > 0 else np.ones_like(r)
    w = np.where(r <= mid, rise, (hi - r) / (hi - mid))
    return np.where((r >= lo) & (r <= hi), np.clip(w, 0.0, 1.0), 0.0)


def _struve_fixed(v, x):
    """Struve H_v(x); scipy 1.15 returns NaN exactly at some zeros of H_0
    (e.g. x = 25.7653728): such values are replaced by the mean at x -+ 1e-4
    (H_v is smooth there, error ~1e-9)."""
    h = np.asarray(struve(v, x), float)
    bad = ~np.isfinite(h)
    if bad.any():
        xb = np.broadcast_to(x, h.shape)[bad]
        h = h.copy()
        h[bad] = 0.5 * (struve(v, xb - 1e-4) + struve(v, xb + 1e-4))
    return h


def _j0_integral(x):
    """Int_0^x J0(t) dt for real x >= 0 (Abramowitz & Stegun 11.1.7)."""
    return x * j0(x) + 0.5 * np.pi * x * (j1(x) * _struve_fixed(0, x) - j0(x) * _struve_fixed(1, x))


def _pieces(j, nodes):
    """The linear pieces (r0, r1, alpha, beta) of ring j: w = alpha + beta r."""
    out = []
    if j > 0:
        lo, mid = nodes[j - 1], nodes[j]
        out.append((lo, mid, -lo / (mid - lo), 1.0 / (mid - lo)))
    mid, hi = nodes[j], nodes[j + 1]
    out.append((mid, hi, hi / (hi - mid), -1.0 / (hi - mid)))
    return out


def ring_spectrum(j, k, nodes, memo=None):
    """Hankel transform 2 pi Int w_j(r) J0(k r) r dr at (complex) k: a
    24-point Gauss rule per piece where k b is small or k complex (the
    lifted part of the path), and on the real axis beyond the closed form
    2 pi Int (alpha + beta r) J0 r dr = 2 pi [alpha r J1 / k + beta (r^2 J1 / k
    + r J0 / k^2 - Int_0^(kr) J0 / k^3)].  memo: a dict shared by the rings
    of one frill on one k array (the Struve functions dominate the cost)."""
    k = np.asarray(k, complex)
    out = np.zeros(k.shape, complex)
    real_far = (np.abs(k.imag) < 1e-12 * np.abs(k)) & (np.abs(k) * nodes[-1] > 40)
    pieces = _pieces(j, nodes)
    near = ~real_far
    if near.any():
        kn = k[near]
        acc = np.zeros(kn.shape, complex)
        if j == 0:  # the barrel disc, w = 1
            a = nodes[0]
            kz = np.where(np.abs(kn) * a < 1e-6, 1e-6 / a, kn)
            acc += 2 * np.pi * a * jv(1, kz * a) / kz
        cplx = kn.imag != 0
        for r0, r1, al, be in pieces:
            r = r0 + 0.5 * (r1 - r0) * (_GX + 1)
            w = 0.5 * (r1 - r0) * _GW * 2 * np.pi * r * (al + be * r)
            if cplx.any():
                acc[cplx] += jv(0, np.multiply.outer(kn[cplx], r)) @ w
            if (~cplx).any():
                acc[~cplx] += j0(np.multiply.outer(kn[~cplx].real, r)) @ w
        out[near] = acc
    if real_far.any():
        kf = k[real_far].real

        def prim(r, al, be):
            key = r  # the memo belongs to one k array
            if memo is not None and key in memo:
                b0, b1, bi = memo[key]
            else:
                x = kf * r
                b0, b1, bi = j0(x), j1(x), _j0_integral(x)
                if memo is not None:
                    memo[key] = (b0, b1, bi)
            return al * r * b1 / kf + be * (r * r * b1 / kf + r * b0 / kf ** 2 - bi / kf ** 3)

        acc = np.zeros(kf.shape)
        if j == 0:
            a = nodes[0]
            acc += 2 * np.pi * a * j1(kf * a) / kf
        for r0, r1, al, be in pieces:
            acc += 2 * np.pi * (prim(r1, al, be) - prim(r0, al, be))
        out[real_far] = acc
    return out


def ring_area(j, nodes):
    """Int w_j dA (the transform at k -> 0)."""
    return float(ring_spectrum(j, np.array([1e-9 / nodes[-1]]), nodes)[0].real)


def tm_side(eps, t, far, k0, kr):
    """TM line of one section side seen from the plate: layers eps, t listed
    from the plate outward, far end "pec" (a plane or cover) or "air".
    Returns the normalised input admittance y_in (y = eps / u per layer)
    and, for a PEC far end, the transfer admittance y_tr (current into the
    far plate per volt on the near one; None otherwise)."""
    kr = np.asarray(kr, complex)
    eps = np.asarray(eps, complex)
    t = np.asarray(t, float)
    n = len(t)
    u = np.stack([csqrt_re_pos(kr ** 2 - (k0 ** 2) * e) for e in eps])
    y = eps[:, None] / u
    q = np.exp(-2 * u * t[:, None])
    th = (1 - q) / (1 + q)  # tanh(u t)
    sh = 2 * np.exp(-u * t[:, None]) / (1 + q)  # sech(u t)
    Y = [None] * (n + 1)  # admittance looking outward at the near side of layer i
    if far == "pec":
        Y[n - 1] = y[n - 1] / th[n - 1]
    else:
        Y[n] = 1.0 / np.sqrt(kr ** 2 - k0 ** 2)
        Y[n - 1] = y[n - 1] * (Y[n] + y[n - 1] * th[n - 1]) / (y[n - 1] + Y[n] * th[n - 1])
    for i in range(n - 2, -1, -1):
        Y[i] = y[i] * (Y[i + 1] + y[i] * th[i]) / (y[i] + Y[i + 1] * th[i])
    y_tr = None
    if far == "pec":
        v = np.ones_like(kr)
        for i in range(n - 1):
            v = v * sh[i] / (1 + Y[i + 1] / y[i] * th[i])
        y_tr = y[n - 1] * v * sh[n - 1] / th[n - 1]
    return Y[0], y_tr


def side_of_plane(stack, plane, side):
    """(section, eps and t from the plate outward, far end, plate position
    'bottom' / 'top' of the section) on side +1 (above) / -1 (below)."""
    k = stack.planes[plane]
    s = int(np.searchsorted(stack.bounds, k)) - (1 if side < 0 else 0)
    lo, hi = stack.bounds[s], stack.bounds[s + 1]
    lay = stack.layers[lo:hi]
    eps = [d.eps_r * (1 - 1j * d.tand) for d in lay]
    t = [d.t for d in lay]
    if side > 0:
        return s, eps, t, "pec" if stack.section_top_pec(s) else "air", "bottom"
    return s, eps[::-1], t[::-1], "pec", "top"


def _mode_radial(n, kt, d, r):
    """Radial factor of parallel-plate mode n of G_zz at distances r:
    H0(2)(kt r) / 4j for the TEM mode, K0(q_n r) / 2 pi for TM_n."""
    r = np.maximum(np.asarray(r, float), 1e-12)
    if n == 0:
        return hankel2(0, kt * r) / 4j
    q = np.sqrt((n * np.pi / d) ** 2 - kt ** 2 + 0j)
    return kv(0, q * r) / (2 * np.pi)


class FrillModel:
    """The antipad (frill) blocks of one solver's impedance matrix."""

    def __init__(self, solver):
        self.solver = solver
        st, m = solver.stack, solver.mesh
        self.frills = list(m.frills)
        self.offset = np.cumsum([0] + [fr["nr"] for fr in self.frills])
        self.nf = int(self.offset[-1])
        self.nodes = [ring_nodes(fr["a"], fr["b"], fr["nr"]) for fr in self.frills]
        self.sides = [[(sd,) + side_of_plane(st, fr["plane"], sd) for sd in (1, -1)] for fr in self.frills]
        vz = {n: st.level_z(n) for n in m.layer_names + list(st.planes)}
        vz["gnd"] = 0.0
        self.spans = [(vz[g["bottom"]], vz[g["top"]]) for g in m.via_segments]
        self.vias = [m.vias[g["via"]] for g in m.via_segments]

    def _path(self, s, k0, extra_to, bmax):
        """The Sommerfeld path of section s (its tables' path when the
        section holds metal levels: lifted part and real tail; the lifted
        part only otherwise), extended along the real axis to extra_to in
        8-point panels about 0.5 / bmax wide (the ring spectra decay
        slowly)."""
        st = self.solver.stack
        lo, hi = st.bounds[s], st.bounds[s + 1]
        lay = st.layers[lo:hi]
        kk = k0 * np.sqrt(np.array([d.eps_r * (1 - 1j * d.tand) for d in lay]))
        t = np.array([d.t for d in lay])
        has_levels = any(st.section(n) == s for n in self.solver.mesh.layer_names)
        if has_levels:
            kc, wc, kr, wr = sommerfeld_path(k0, kk, t, self.solver.rho_max, 12)
            k1 = float(kr[-1])
        else:
            # the lifted part over [0, T] with T from the real parts of eps (no tables in this section)
            kmaxr = abs(k0 * np.sqrt(np.max([d.eps_r for d in lay])))
            kc, wc, kr, wr = sommerfeld_path(k0, np.array([kmaxr]), t, self.solver.rho_max, 12, tail=False)
            k1 = 1.3 * kmaxr + k0
        ks, ws = [kc, kr.astype(complex)], [wc, wr.astype(complex)]
        if extra_to > k1:
            n = int(np.ceil((extra_to - k1) * bmax / 0.5))
            t2, w2 = gauss_panels(k1, extra_to, max(n, 8), 8)
            ks.append(t2.astype(complex))
            ws.append(w2.astype(complex))
        return np.concatenate(ks), np.concatenate(ws), has_levels

    def blocks(self, omega):
        """C (N_J x nf) and F = j w Cff (nf x nf) at one frequency."""
        sol, st, m = self.solver, self.solver.stack, self.solver.mesh
        plan = sol.geometry()
        nb, nv, nf = m.n_basis, m.n_vias, self.nf
        k0 = omega / C0
        C = np.zeros((nb + nv, nf), complex)
        F = np.zeros((nf, nf), complex)
        bmax = max(fr["b"] for fr in self.frills)
        amin = min(fr["a"] for fr in self.frills)
        paths, spectra = {}, {}

        def spec(l, key, kr):
            if (l, key) not in spectra:
                memo = {}
                spectra[l, key] = np.stack([ring_spectrum(j, kr, self.nodes[l], memo)
                                            for j in range(self.frills[l]["nr"])])
            return spectra[l, key]

        for k, fr in enumerate(self.frills):
            o, nr = self.offset[k], fr["nr"]
            xy = np.asarray(fr["xy"])
            for sd, s, eps, t, far, plate in self.sides[k]:
                if (s, sd) not in paths:
                    kr, wk, has_levels = self._path(s, k0, 400.0 / amin, bmax)
                    yin, ytr = tm_side(eps, t, far, k0, kr)
                    paths[s, sd] = (kr, wk, has_levels, yin, ytr)
                kr, wk, has_levels, yin, ytr = paths[s, sd]
                wt = spec(k, (s, sd), kr)
                ys = EPS0 * kr ** 2 * yin
                for l2, fr2 in enumerate(self.frills):
                    o2, nr2 = self.offset[l2], fr2["nr"]
                    dist = float(np.hypot(*(np.asarray(fr2["xy"]) - xy)))
                    J = jv(0, kr * dist) if dist > 0 else 1.0
                    if fr2["plane"] == fr["plane"]:  # the same plate, the same side
                        wt2 = spec(l2, (s, sd), kr)
                        F[o:o + nr, o2:o2 + nr2] += 1j * omega * ((wt * (wk * ys * J * kr)) @ wt2.T) / (2 * np.pi)
                    elif far == "pec" and any(s2 == s and sd2 == -sd for sd2, s2, *_ in self.sides[l2]):
                        wt2 = spec(l2, (s, sd), kr)  # the far plate of this closed section
                        F[o:o + nr, o2:o2 + nr2] -= 1j * omega * ((wt * (wk * EPS0 * kr ** 2 * ytr * J * kr))
                                                                  @ wt2.T) / (2 * np.pi)
                if sd > 0:
                    C[nb + fr["seg_up"], o] -= 1.0  # downward via current against the normal +z
                else:
                    C[nb + fr["seg_low"], o] += 1.0  # downward via current along the normal -z
                if has_levels:
                    self._charges(C, o, nr, fr, s, plate, kr, wk, wt, plan, k0)
                if far == "pec":
                    self._vertical(C, o, nr, self.nodes[k], fr, s, sd, omega)
        return C, F

    def _charges(self, C, o, nr, fr, s, plate, kr, wk, wt, plan, k0):
        """C[:, ring j] -= Int (div f) Phi_j dA over the levels of section s,
        Phi_j(r) = Int sigma_1 w_j dA: the ring-weighted plate charge of a
        unit point charge on the level, through the section's TM line."""
        st, m = self.solver.stack, self.solver.mesh
        nb = m.n_basis
        sub = st.section_stack(s)
        names = [n for n in m.layer_names if st.section(n) == s]
        epsc = np.array([d.eps_r * (1 - 1j * d.tand) for d in sub.layers])
        t = np.array([d.t for d in sub.layers])
        u = np.stack([csqrt_re_pos(kr ** 2 - (k0 ** 2) * e) for e in epsc])
        u0 = np.sqrt(kr ** 2 - k0 ** 2)
        x = u * t[:, None]
        il = [sub.metals[n] for n in names]
        Ve = _interface_voltages(epsc[:, None] / u, x, 1.0 / u0, il, sub.top == "pec")
        y = epsc[:, None] / u
        q = np.exp(-2 * x)
        th = (1 - q) / (1 + q)
        sh = 2 * np.exp(-x) / (1 + q)
        n_lay = len(t)
        xy = np.asarray(fr["xy"])
        rcut = fr["b"] + 15 * float(t.sum())
        rr = np.linspace(0, rcut, 160)
        cen, qp7, qw7 = plan.cen, plan.qp7, plan.qw7
        for li, name in enumerate(names):
            if plate == "bottom":  # current into the plate from the layer next to it, per unit level current
                sig = -y[0] * Ve[li, 1] * sh[0] / th[0]
            else:
                sig = -y[n_lay - 1] * Ve[li, n_lay - 1] * sh[n_lay - 1] / th[n_lay - 1]
            # sig falls like exp(-k dz): the long real tail kept for the ring self terms adds nothing here
            keep = np.abs(sig) > 1e-16 * np.abs(sig).max()
            ks, cw = kr[keep], (wt * (wk * sig * kr))[:, keep]
            cplx = ks.imag != 0
            phi = (jv(0, np.multiply.outer(rr, ks[cplx])) @ cw[:, cplx].T
                   + j0(np.multiply.outer(rr, ks[~cplx].real)) @ cw[:, ~cplx].T) / (2 * np.pi)  # (160, nr)
            spl = CubicSpline(rr, phi, axis=0)
            lidx = m.layer_names.index(name)
            on = np.nonzero((m.tri_layer == lidx) & (np.hypot(*(cen - xy).T) < rcut + 1e-3))[0]
            if not on.size:
                continue
            d = np.hypot(qp7[on, :, 0] - xy[0], qp7[on, :, 1] - xy[1])
            vals = np.where((d < rcut)[..., None], spl(np.minimum(d, rcut)), 0.0)  # (n_on, 7, nr)
            iphi = np.einsum("tqj,tq->tj", vals, qw7[on])
            bt, cv = plan.tb[on], plan.cV[on]
            ok = bt >= 0
            for j in range(nr):
                np.add.at(C[:, o + j], bt[ok], (-cv * iphi[:, j][:, None])[ok])
            if plan.nv:
                pos = plan.pad_pos[on]
                sel = pos >= 0
                if sel.any():
                    C[nb:nb + plan.nv, o:o + nr] -= plan.Dvia[pos[sel]].T @ iphi[sel]

    def _vertical(self, C, o, nr, nodes, fr, s, sd, omega):
        """Vertical via currents of the closed section s: j w Int sigma^A w_j
        dA with sigma^A = eps E_z^A . n at the plate and E_z^A = -j w A_z
        from the parallel-plate modes of G_zz (TEM and TM_n)."""
        st, m = self.solver.stack, self.solver.mesh
        nb = m.n_basis
        lo, hi = st.bounds[s], st.bounds[s + 1]
        zb, zt = float(st.z[lo]), float(st.z[hi])
        d = zt - zb
        lay = st.layers[lo:hi]
        eps_s = d / sum(L.t / (L.eps_r * (1 - 1j * L.tand)) for L in lay)
        kt = omega / C0 * np.sqrt(eps_s)
        zp, nz = (0.0, 1.0) if sd > 0 else (d, -1.0)
        xy = np.asarray(fr["xy"])
        nmax = int(np.ceil(40 * d / (np.pi * fr["a"]))) + 1
        n = np.arange(0, nmax + 1)
        kn = n * np.pi / d
        epsn = np.where(n == 0, 1.0, 2.0)
        edges = np.r_[0.0, nodes]
        rr = np.concatenate([r0 + 0.5 * (r1 - r0) * (_GX + 1) for r0, r1 in zip(edges[:-1], edges[1:])])
        ww = np.concatenate([0.5 * (r1 - r0) * _GW for r0, r1 in zip(edges[:-1], edges[1:])])
        wj = np.stack([ring_weight(j, rr, nodes) for j in range(nr)])
        area = np.array([ring_area(j, nodes) for j in range(nr)])
        g_own = np.stack([_mode_radial(nn, kt, d, rr) for nn in n])  # (modes, R)
        G_own = (wj * (ww * 2 * np.pi * rr)) @ g_own.T  # (nr, modes)
        for seg, (z0, z1) in enumerate(self.spans):
            a1, a2 = np.clip(z0, zb, zt) - zb, np.clip(z1, zb, zt) - zb
            if a2 - a1 <= 1e-12 * d:
                continue
            dist = float(np.hypot(*(np.asarray(self.vias[seg].xy) - xy)))
            cn = np.where(n == 0, a2 - a1, (np.sin(kn * a2) - np.sin(kn * a1)) / np.where(n == 0, 1, kn))
            G = G_own if dist < 1e-9 else np.outer(area, [_mode_radial(nn, kt, d, np.array([dist]))[0] for nn in n])
            Az = -MU0 * (G @ (epsn / d * np.cos(kn * zp) * cn))  # Int w_j A_z dA per unit segment current
            C[nb + seg, o:o + nr] += 1j * omega * EPS0 * eps_s * (-1j * omega) * Az * nz

    def matrix_blocks(self, omega):
        """The blocks that extend Z: (-C, -(F + F^T) / 2)."""
        C, F = self.blocks(omega)
        return -C, -0.5 * (F + F.T)
