"""Host side of the port de-embedding: the shared-gamma / Z0 wave fit, the
S-matrix of one frequency and the reference impedances
(Engine/mom25d/solver.py:618-670, 1284-1325, 1337-1427).

This module imports numpy, scipy and the analytic line formulas only, so
that worker processes can run the fits of a sweep (one frequency point per
task) next to the GPU work without loading torch.
"""

import warnings
from dataclasses import dataclass, field

import numpy as np
from scipy.optimize import minimize, minimize_scalar

from . import analytic

VOLTAGE_WEIGHT = 1.0


def fit_waves(s_list, I_list, V_list, gamma0, lossless=False, v_weight=VOLTAGE_WEIGHT, bounds=(0.7, 1.3)):
    """Shared-gamma / Z0 fit of I(s) = A e^{gs} + B e^{-gs}, V = Z0 (A e^{gs} - B e^{-gs})
    over all data sets (Engine/mom25d/solver.py:618-670): bounded scalar search
    of beta for lossless stacks, Nelder-Mead on (alpha, beta) otherwise.
    Returns gamma, Z0, [(A, B), ...], relative residual."""
    scale = abs(gamma0.imag)
    Inrm = sum(np.sum(np.abs(I) ** 2) for I in I_list)
    Vnrm = sum(np.sum(np.abs(V) ** 2) for V in V_list)

    def solve(g):
        coefs, res, U_all, V_all = [], 0.0, [], []
        for s, I, V in zip(s_list, I_list, V_list):
            ep, em = np.exp(g * s), np.exp(-g * s)
            M = np.stack([ep, em], 1)
            c, *_ = np.linalg.lstsq(M, I, rcond=None)
            coefs.append(c)
            res += np.sum(np.abs(M @ c - I) ** 2) / Inrm
            U_all.append(c[0] * ep - c[1] * em)
            V_all.append(V)
        U, V = np.concatenate(U_all), np.concatenate(V_all)
        z0 = np.vdot(U, V) / np.vdot(U, U)
        res += v_weight * np.sum(np.abs(V - z0 * U) ** 2) / Vnrm
        return coefs, z0, res

    if lossless:
        lo, hi = bounds
        opt = minimize_scalar(lambda b: solve(1j * scale * b)[2], bounds=(lo, hi), method="bounded",
                              options=dict(xatol=1e-12))
        if min(opt.x - lo, hi - opt.x) < 1e-3:
            warnings.warn(f"beta fit hit its search bound ({opt.x:.4f} x the Kirschning-Jansen estimate)",
                          RuntimeWarning, stacklevel=2)
        g = 1j * scale * opt.x
    else:
        p0 = np.array([gamma0.real, gamma0.imag]) / scale
        opt = minimize(lambda p: solve(scale * (p[0] + 1j * p[1]))[2], p0, method="Nelder-Mead",
                       options=dict(xatol=1e-10, fatol=1e-16, maxiter=4000))
        g = scale * (opt.x[0] + 1j * opt.x[1])
    coefs, z0, res = solve(g)
    return g, z0, coefs, np.sqrt(res)


@dataclass
class FitContext:
    """What the fit of one frequency needs about the ports of a solver."""

    groups: list            # port indices of each group (shared gamma / Z0)
    width: list             # strip width of each group
    h: list                 # height of the group's metal level over ground
    eps: list               # relative permittivity below the level
    covered: list           # a dielectric above the level
    z0_from: list           # per port: "voltage" or "reciprocity"
    s: list                 # per port: the sample cut positions
    lossless: bool
    v_weight: float
    z0_fixed: np.ndarray = None
    z0_recip_min: float = None
    extra: dict = field(default_factory=dict)
    zi: list = field(default_factory=list)  # reference impedance of each internal port


def sparams_point(ctx, f, I, V, Ii=None):
    """S and the fit results of one frequency from the feed currents and
    voltages at the sample cuts (I[p], V[p]: (Ns_p, Nexc)) and the currents
    through the internal ports' sources (Ii: (n_internal, Nexc)), as in
    Engine/mom25d/solver.py:1284-1325.  Feeds come first, then internal
    ports.  Returns (S, gamma, z0_line, residual, z0_ref, A) per port."""
    npt, ni = len(ctx.z0_from), len(ctx.zi)
    n = npt + ni
    gam = np.zeros(n, complex)
    zl = np.zeros(n, complex)
    res = np.zeros(n)
    Acoef = np.zeros((n, n), complex)
    Bcoef = np.zeros((n, n), complex)
    for k, members in enumerate(ctx.groups):
        b0 = analytic.beta(ctx.width[k], ctx.h[k], ctx.eps[k], f)
        bnd = (0.7, 1.3) if not ctx.covered[k] else (0.7, 1.8)
        s_list, I_list, V_list, keys = [], [], [], []
        for p in members:
            for e in range(n):
                s_list.append(ctx.s[p])
                I_list.append(I[p][:, e])
                V_list.append(V[p][:, e])
                keys.append((p, e))
        gm, z0, coefs, r = fit_waves(s_list, I_list, V_list, 1e-3 * b0 + 1j * b0, lossless=ctx.lossless,
                                     bounds=bnd, v_weight=ctx.v_weight)
        for (p, e), (a, bb) in zip(keys, coefs):
            Acoef[p, e], Bcoef[p, e] = a, bb
        for p in members:
            gam[p], zl[p], res[p] = gm, z0, r
    zr = np.zeros(n)
    if ni:  # internal ports: V = Vs - z_ref I; a = Vs / 2 sqrt(z), b = (V - z I) / 2 sqrt(z), current-normalised
        zi = np.asarray(ctx.zi, float)
        Vs = np.zeros((ni, n))
        Vs[:, npt:] = np.eye(ni)
        Acoef[npt:] = Vs / (2 * zi[:, None])
        Bcoef[npt:] = -(Vs - 2 * zi[:, None] * np.asarray(Ii)) / (2 * zi[:, None])
        zl[npt:] = zi
        zr[npt:] = zi
    SI = -Bcoef @ np.linalg.inv(Acoef)
    if npt and ctx.z0_fixed is None:
        zr[:npt] = reciprocal_z0(ctx, SI[:npt, :npt], zl[:npt].real)
        if ni:
            zr[:npt] = anchor_z0(ctx, SI, zr[:npt], ctx.zi)
    elif npt:
        zr[:npt] = np.asarray(ctx.z0_fixed, float)
    dd = np.sqrt(zr)
    return dd[:, None] * SI / dd[None, :], gam, zl, res, zr, Acoef


def anchor_z0(ctx, SI, zfeed, zi, prior=1e-4):
    """Feed reference impedances refined by reciprocity with the internal
    ports, whose z_ref is exact: per feed group the coupling-weighted mean of
    log(z_i |SI_ij / SI_ji|) over internal ports i and members j, with the
    feed-only estimate as a weak prior (Engine/mom25d/solver.py:1337-1355)."""
    npt = len(zfeed)
    z = np.array(zfeed, float)
    cmin = ctx.z0_recip_min
    for members in ctx.groups:
        num, den = prior * np.log(np.mean(z[members])), prior
        for j in members:
            for k, zk in enumerate(zi):
                i = npt + k
                w = min(abs(SI[i, j]), abs(SI[j, i])) ** 2
                if w > 0 and (cmin is None or w >= cmin ** 2):
                    num += w * np.log(zk * abs(SI[i, j] / SI[j, i]))
                    den += w
        z[members] = np.exp(num / den)
    return z


def reciprocal_z0(ctx, SI, zphi, prior=1e-4):
    """Port reference impedances from reciprocity ratios plus the
    scalar-potential level (Engine/mom25d/solver.py:1357-1397)."""
    groups = ctx.groups
    recip = np.array([ctx.z0_from[m[0]] == "reciprocity" for m in groups])
    if recip.all():
        raise ValueError("every port has z0_from='reciprocity': at least one needs 'voltage'")
    cmin = ctx.z0_recip_min
    if not recip.any():
        return voltage_z0(SI, zphi, groups, prior, cmin)
    gid = np.empty(len(ctx.z0_from), int)
    for k, members in enumerate(groups):
        gid[members] = k
    vport = ~recip[gid]
    z = np.array(zphi, float)
    vidx = np.nonzero(vport)[0]
    vgroups = [np.nonzero(np.isin(vidx, g))[0] for g, r in zip(groups, recip) if not r]
    z[vidx] = voltage_z0(SI[np.ix_(vidx, vidx)], z[vidx], vgroups, prior, cmin)
    for k in np.nonzero(recip)[0]:
        num = den = 0.0
        for i in groups[k]:
            for j in np.nonzero(vport)[0]:
                w = min(abs(SI[i, j]), abs(SI[j, i])) ** 2
                if w > 0 and (cmin is None or w >= cmin ** 2):
                    num += w * np.log(z[j] * abs(SI[j, i] / SI[i, j]))
                    den += w
        if den == 0:
            raise ValueError(f"port group {k} (z0_from='reciprocity') does not couple to any voltage port")
        z[groups[k]] = np.exp(num / den)
    return z


def voltage_z0(SI, zphi, groups, prior, cmin=None):
    """Reference impedances of voltage ports: weighted least squares of the
    reciprocity log-ratios with a prior on each group's scalar-potential
    level (Engine/mom25d/solver.py:1399-1427)."""
    if len(groups) == 1:
        return zphi
    gid = np.empty(len(zphi), int)
    for k, members in enumerate(groups):
        gid[members] = k
    rows, rhs, wts = [], [], []
    for i in range(len(gid)):
        for j in range(i + 1, len(gid)):
            if gid[i] != gid[j]:
                w = min(abs(SI[i, j]), abs(SI[j, i])) ** 2
                if w > 0 and (cmin is None or w >= cmin ** 2):
                    r = np.zeros(len(groups))
                    r[gid[i]], r[gid[j]] = 1, -1
                    rows.append(r)
                    rhs.append(np.log(abs(SI[j, i] / SI[i, j])))
                    wts.append(w)
    for k, members in enumerate(groups):
        r = np.zeros(len(groups))
        r[k] = 1
        rows.append(r)
        rhs.append(np.log(np.mean(zphi[members])))
        wts.append(prior)
    sw = np.sqrt(wts)
    logz, *_ = np.linalg.lstsq(np.array(rows) * sw[:, None], np.array(rhs) * sw, rcond=None)
    logz += np.mean(np.log(zphi)) - np.mean(logz[gid])
    return np.exp(logz[gid])
