"""Oracle: the per-frequency sweep on the CPU (Engine/mom25d/solver.py:11-31,
984-1335): tables, assemble, LAPACK LU (scipy), charges, line voltages, wave
fit and S.  The port samples, the wave fit (with the reference's optimisers)
and the reference-impedance fit are the oracle's own (deembed_ref), pinned
by the reference's golden S / gamma / z0_line vectors.  Only the input data
model (Stackup: levels, heights, surface impedance) and the Kirschning-Jansen
start value (analytic.beta) come from the package.  Test infrastructure only."""

import time

import numpy as np
import scipy.linalg as sla

from .assemble_ref import EPS0, Geometry, assemble
from .greens_ref import LayeredGreensRef, SectionedGreensRef, SlabGreens
from .integrals_ref import TRI7_BARY, TRI7_W, points, wilton


def make_greens(stack, mesh, f, rho_max):
    single = len(stack.layers) == 1 and stack.top == "air" and mesh.layer_names == ["top"]
    zg = stack.surface_impedance(f) if stack.ground_loss else 0.0
    if single:
        d = stack.layers[0]
        return SlabGreens(f, d.eps_r, d.t, d.tand, rho_max, zs_ground=zg)
    cls = SectionedGreensRef if stack.planes else LayeredGreensRef
    return cls(f, stack, rho_max, zs_ground=zg, levels=mesh.layer_names)


class OracleSolver:
    """CPU restatement of Solver(mesh, stack) with method='dense', interp=None."""

    def __init__(self, mesh, substrate, use_c=True, lapack_threads=None):
        from mom25d_amd.layered import Stackup

        from .deembed_ref import PortSamplesRef

        self.mesh = mesh
        self.stack = substrate if isinstance(substrate, Stackup) else substrate.to_stackup()
        st = self.stack
        planes = list(getattr(mesh, "planes", []))
        order = sorted(mesh.layer_names + planes, key=st.level_z)
        if getattr(mesh, "level_order", None) != order:
            mesh.assign_levels(order)
        self.ports = mesh.layout.ports
        self.pm = [PortSamplesRef(mesh, p, st.level_h(p.layer)) for p in self.ports]
        v = mesh.nodes
        self.rho_max = 1.02 * np.hypot(*(v.max(0) - v.min(0)))
        self.zl = np.array([st.level_z(n) for n in mesh.layer_names])
        via_z = {n: st.level_z(n) for n in mesh.layer_names + planes}
        via_z["gnd"] = 0.0
        self.section = [st.section(n) for n in mesh.layer_names] if st.planes else None
        self.via_z = via_z
        self.use_c = use_c
        self._geom = None
        self.groups = {}
        for p, port in enumerate(self.ports):
            key = (round(port.width / st.level_h(port.layer), 9), port.layer, None, port.z0_from)
            self.groups.setdefault(key, []).append(p)
        self.v_weight = 1.0  # VOLTAGE_WEIGHT, Engine/mom25d/solver.py:615
        self.timing = {}

    def geometry(self):
        if self._geom is None:
            self._geom = Geometry(self.mesh, zl=self.zl, via_z=self.via_z, section=self.section)
        return self._geom

    def greens(self, f):
        return make_greens(self.stack, self.mesh, f, self.rho_max)

    def impedance(self, f, greens=None):
        g = greens or self.greens(f)
        zs = self.stack.surface_impedance(f)
        return assemble(self.mesh, g, zs, geom=self.geometry(), use_c=self.use_c)

    def port_excitations(self):
        return np.stack([pm.excitation for pm in self.pm], 1)

    def solve_currents(self, f, greens=None, Z=None):
        Z = self.impedance(f, greens) if Z is None else Z
        return sla.lu_solve(sla.lu_factor(Z, overwrite_a=True, check_finite=False), self.port_excitations())

    def charges(self, x, omega):
        m = self.mesh
        b = m.tri_basis
        lab = np.where(b >= 0, m.edge_len[np.maximum(b, 0)], 0.0)
        cV = m.tri_sign * lab / m.area[:, None]
        div = np.einsum("ti,tie->te", cV, np.where((b >= 0)[..., None], x[np.maximum(b, 0)], 0))
        for k, g in enumerate(m.via_segments):
            for key, sgn in (("top_tris", 1.0), ("bottom_tris", -1.0)):
                t = g[key]
                if t is not None:
                    div[t] += sgn * x[m.n_basis + k] / m.area[t].sum()
        return -div / (1j * omega)

    def line_voltage(self, greens, q, pts, layer):
        m = self.mesh
        li = m.layer_names.index(layer)
        out = 0.0
        for ls in range(len(m.layer_names)):
            on = np.nonzero(m.tri_layer == ls)[0]
            if not on.size:
                continue
            pr = greens.pair(li, ls)
            Vv = m.verts[on]
            qp = points(Vv, TRI7_BARY)
            qw = TRI7_W[None, :] * m.area[on][:, None]
            d = pts[None, :, None, :] - qp[:, None, :, :]
            K0W = np.einsum("qml,ql->qm", pr.smooth_w(np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2)), qw)
            if pr.cW:
                I0, _ = wilton(Vv, pts, pr.dz)
                K0W = K0W + pr.cW * I0 / (4 * np.pi)
            out = out + (K0W.T @ q[on]) / EPS0
        return out

    def sweep(self, freqs):
        from mom25d_amd import analytic

        from .deembed_ref import fit_waves_ref, z0_reciprocal_ref

        freqs = np.atleast_1d(np.asarray(freqs, float))
        nf, n = len(freqs), len(self.ports)
        S = np.zeros((nf, n, n), complex)
        gam = np.zeros((nf, n), complex)
        zl = np.zeros((nf, n), complex)
        zr = np.zeros((nf, n))
        res = np.zeros((nf, n))
        tt = dict(tables=0.0, fill=0.0, lu=0.0, sparams=0.0)
        for fi, f in enumerate(freqs):
            t0 = time.time()
            g = self.greens(f)
            t1 = time.time()
            Z = self.impedance(f, g)
            t2 = time.time()
            x = self.solve_currents(f, Z=Z)
            t3 = time.time()
            q = self.charges(x, g.omega)
            phi = [None] * n
            for lay in dict.fromkeys(p.layer for p in self.ports):
                ps = [p for p in range(n) if self.ports[p].layer == lay]
                vp = self.line_voltage(g, q, np.concatenate([self.pm[p].vpts for p in ps]), lay)
                for p, v in zip(ps, np.split(vp, np.cumsum([len(self.pm[p].vpts) for p in ps])[:-1])):
                    phi[p] = v
            A = np.zeros((n, n), complex)
            B = np.zeros((n, n), complex)
            for members in self.groups.values():
                port = self.ports[members[0]]
                hp = self.stack.level_h(port.layer)
                b0 = analytic.beta(port.width, hp, self.stack.eps_below(port.layer), f)
                bnd = (0.7, 1.3) if not self.stack.covered(port.layer) else (0.7, 1.8)
                s_l, I_l, V_l, keys = [], [], [], []
                for p in members:
                    Ip, Vp = self.pm[p].currents(x), self.pm[p].voltages(phi[p])
                    for e in range(n):
                        s_l.append(self.pm[p].s)
                        I_l.append(Ip[:, e])
                        V_l.append(Vp[:, e])
                        keys.append((p, e))
                gm, z0, coefs, r = fit_waves_ref(s_l, I_l, V_l, 1e-3 * b0 + 1j * b0, lossless=self.stack.lossless,
                                                 bounds=bnd, v_weight=self.v_weight)
                for (p, e), (a, bb) in zip(keys, coefs):
                    A[p, e], B[p, e] = a, bb
                for p in members:
                    gam[fi, p], zl[fi, p], res[fi, p] = gm, z0, r
            SI = -B @ np.linalg.inv(A)
            zr[fi] = z0_reciprocal_ref(SI, zl[fi].real, list(self.groups.values()), [p.z0_from for p in self.ports])
            d = np.sqrt(zr[fi])
            S[fi] = d[:, None] * SI / d[None, :]
            t4 = time.time()
            for k, v in zip(("tables", "fill", "lu", "sparams"), (t1 - t0, t2 - t1, t3 - t2, t4 - t3)):
                tt[k] += v
        self.timing = tt
        return dict(freqs=freqs, S=S, z0=zr, gamma=gam, z0_line=zl, fit_residual=res)
