"""Oracle: port de-embedding of one frequency as the reference computes it
(TEST INFRASTRUCTURE ONLY) -- the wave fit with the reference's own
optimisers and options, and the reciprocity-based reference impedances.

  fit_waves_ref       Engine/mom25d/solver.py:618-670  (scipy minimize_scalar 'bounded' with xatol
                      1e-12 on lossless stacks, Nelder-Mead with xatol 1e-10, fatol 1e-16,
                      maxiter 4000 otherwise; per data set numpy lstsq of the current samples)
  z0_voltage_ref      Engine/mom25d/solver.py:1399-1427 (weighted numpy lstsq of log-ratios + prior)
  z0_reciprocal_ref   Engine/mom25d/solver.py:1357-1397
  z0_anchor_ref       Engine/mom25d/solver.py:1337-1355
  port_samples_ref    Engine/mom25d/solver.py:570-602 (sample cuts, voltage points, gap excitation)

Pinned by the reference's golden S / gamma / z0_line (tests/golden/sweeps.npz
and config_*.npz through oracle.solve_ref.OracleSolver, test_oracle_sweep).
"""

import numpy as np
from scipy.optimize import minimize, minimize_scalar


def fit_waves_ref(s_list, I_list, V_list, gamma0, lossless, v_weight=1.0, bounds=(0.7, 1.3)):
    """(gamma, Z0, [(A, B)], relative residual) of the shared-wave fit."""
    beta_scale = abs(gamma0.imag)
    i_norm2 = sum(float(np.sum(np.abs(np.asarray(I)) ** 2)) for I in I_list)
    v_norm2 = sum(float(np.sum(np.abs(np.asarray(V)) ** 2)) for V in V_list)

    def objective(g):
        amps, shape, volts = [], [], []
        total = 0.0
        for s, I, V in zip(s_list, I_list, V_list):
            fwd, bwd = np.exp(g * s), np.exp(-g * s)
            basis = np.column_stack([fwd, bwd])
            ab = np.linalg.lstsq(basis, I, rcond=None)[0]
            amps.append(ab)
            total += float(np.sum(np.abs(basis @ ab - I) ** 2)) / i_norm2
            shape.append(ab[0] * fwd - ab[1] * bwd)
            volts.append(V)
        u, v = np.concatenate(shape), np.concatenate(volts)
        z0 = np.vdot(u, v) / np.vdot(u, u)
        total += v_weight * float(np.sum(np.abs(v - z0 * u) ** 2)) / v_norm2
        return total, z0, amps

    if lossless:
        opt = minimize_scalar(lambda b: objective(1j * beta_scale * b)[0], bounds=bounds, method="bounded",
                              options={"xatol": 1e-12})
        g = 1j * beta_scale * opt.x
    else:
        start = np.array([gamma0.real, gamma0.imag]) / beta_scale
        opt = minimize(lambda p: objective(beta_scale * complex(p[0], p[1]))[0], start, method="Nelder-Mead",
                       options={"xatol": 1e-10, "fatol": 1e-16, "maxiter": 4000})
        g = beta_scale * complex(opt.x[0], opt.x[1])
    total, z0, amps = objective(g)
    return g, z0, amps, np.sqrt(total)


def _w(SI, i, j, cmin):
    w = min(abs(SI[i, j]), abs(SI[j, i])) ** 2
    return w if (w > 0 and (cmin is None or w >= cmin ** 2)) else 0.0


def z0_voltage_ref(SI, zphi, groups, prior=1e-4, cmin=None):
    zphi = np.asarray(zphi, float)
    if len(groups) == 1:
        return zphi
    gid = np.zeros(len(zphi), int)
    for k, mem in enumerate(groups):
        gid[list(mem)] = k
    A, b, w = [], [], []
    for i in range(len(zphi)):
        for j in range(i + 1, len(zphi)):
            wij = _w(SI, i, j, cmin) if gid[i] != gid[j] else 0.0
            if wij:
                row = np.zeros(len(groups))
                row[gid[i]] += 1.0
                row[gid[j]] -= 1.0
                A.append(row)
                b.append(np.log(abs(SI[j, i] / SI[i, j])))
                w.append(wij)
    for k, mem in enumerate(groups):
        row = np.zeros(len(groups))
        row[k] = 1.0
        A.append(row)
        b.append(np.log(zphi[list(mem)].mean()))
        w.append(prior)
    sq = np.sqrt(np.array(w))
    x = np.linalg.lstsq(np.array(A) * sq[:, None], np.array(b) * sq, rcond=None)[0]
    x = x + np.log(zphi).mean() - x[gid].mean()
    return np.exp(x[gid])


def z0_reciprocal_ref(SI, zphi, groups, z0_from, prior=1e-4, cmin=None):
    recip = [z0_from[m[0]] == "reciprocity" for m in groups]
    if all(recip):
        raise ValueError("no voltage-referenced port")
    if not any(recip):
        return z0_voltage_ref(SI, zphi, groups, prior, cmin)
    z = np.array(zphi, float)
    vports = sorted(p for m, r in zip(groups, recip) if not r for p in m)
    vgroups = [[vports.index(p) for p in m] for m, r in zip(groups, recip) if not r]
    z[vports] = z0_voltage_ref(SI[np.ix_(vports, vports)], z[vports], vgroups, prior, cmin)
    for m, r in zip(groups, recip):
        if not r:
            continue
        num = den = 0.0
        for i in m:
            for j in vports:
                wij = _w(SI, i, j, cmin)
                if wij:
                    num += wij * np.log(z[j] * abs(SI[j, i] / SI[i, j]))
                    den += wij
        if den == 0:
            raise ValueError("a reciprocity-referenced group couples to no voltage port")
        z[m] = np.exp(num / den)
    return z


def z0_anchor_ref(SI, zfeed, zi, groups, prior=1e-4, cmin=None):
    npt = len(zfeed)
    z = np.array(zfeed, float)
    for m in groups:
        num, den = prior * np.log(z[m].mean()), prior
        for j in m:
            for k, zk in enumerate(zi):
                wij = _w(SI, npt + k, j, cmin)
                if wij:
                    num += wij * np.log(zk * abs(SI[npt + k, j] / SI[j, npt + k]))
                    den += wij
        z[m] = np.exp(num / den)
    return z


class PortSamplesRef:
    """Delta-gap column and sample cuts of one port: cuts in
    [margin, s_gap - margin], margin = max(1.5 W, 4 h) unless the port sets
    it; voltage points at +-d around each cut (d = half the distance to the
    nearest other cut) at (-0.3, -0.1, 0.1, 0.3) W across the strip."""

    def __init__(self, mesh, port, h):
        margin = max(1.5 * port.width, 4 * h) if port.margin is None else port.margin
        cuts = np.asarray(mesh.cut_positions(port), float)
        keep = (cuts >= margin - 1e-12) & (cuts <= port.s_gap - margin + 1e-12)
        self.s = cuts[keep]
        if self.s.size < 4:
            raise ValueError(f"port {port.name}: fewer than 4 sample cuts")
        self.cut_edges = [mesh.cut_edges(port, c) for c in self.s]
        gi, gs = mesh.cut_edges(port, port.s_gap)
        self.excitation = np.zeros(mesh.n_unknowns)
        self.excitation[gi] = gs * mesh.edge_len[gi]
        self.edge_len = mesh.edge_len
        pos = np.searchsorted(cuts, self.s)
        below = self.s - cuts[np.maximum(pos - 1, 0)]
        above = cuts[np.minimum(pos + 1, len(cuts) - 1)] - self.s
        half = 0.5 * np.minimum(np.where(below > 0, below, np.inf), np.where(above > 0, above, np.inf))
        across = np.array([-0.3, -0.1, 0.1, 0.3]) * port.width
        along = np.stack([self.s - half, self.s + half], 1)
        self.vpts = port.point(along[:, :, None], across).reshape(-1, 2)
        self.per_cut = 2 * len(across)

    def currents(self, x):
        return np.array([(sg * self.edge_len[i]) @ x[i] for i, sg in self.cut_edges])

    def voltages(self, phi):
        return phi.reshape(self.s.size, self.per_cut, -1).mean(1)
