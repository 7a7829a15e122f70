"""Host side of the port de-embedding: the shared-gamma / Z0 wave fit of the
sampled port currents and voltages, the S-matrix of one frequency and the
reference impedances (the quantities of Engine/mom25d/solver.py:618-670,
1284-1325, 1337-1427).

The wave fit is a separable nonlinear least-squares problem.  For a fixed
propagation constant g the wave amplitudes are linear: per data set the
current samples are projected onto span{e^{gs}, e^{-gs}} (a two-column QR
per port, all excitations of the port at once), and the line impedance is
the least-squares ratio of the voltage samples to the fitted voltage shape
A e^{gs} - B e^{-gs} over all data sets of the group.  What remains is a
function of g alone, minimised here by a Levenberg-Marquardt iteration on
the stacked, normalised current and voltage residuals (two real parameters,
or one on a lossless stack where g = j beta is searched inside its bounds
after a grid scan).  The objective is the reference's; the optimiser
reaches its minimum to round-off, where the reference's Nelder-Mead stops
within its xatol / fatol of it.

The reference impedances follow from reciprocity of the power-wave S:
with current-normalised SI, z_i SI_ij = z_j SI_ji, so every coupled pair
gives an estimate of log z_i - log z_j, weighted by the square of the
weaker coupling.  Voltage-referenced groups are found by weighted least
squares over those differences with a weak prior on their scalar-potential
level (a graph-Laplacian system), reciprocity-referenced groups and feeds
anchored to internal ports as weighted means of the estimates.

This module imports numpy and the analytic line formulas only, so that
worker processes can run the fits of a sweep next to the GPU work without
loading torch.
"""

import warnings
from dataclasses import dataclass, field

import numpy as np

from . import analytic

VOLTAGE_WEIGHT = 1.0


class _WaveProblem:
    """The data sets of one fit, grouped by their cut positions (one group
    per port: every excitation of a port is sampled at the same cuts)."""

    def __init__(self, s_list, I_list, V_list, v_weight):
        self.n_sets = len(s_list)
        blocks = []
        for k, s in enumerate(s_list):
            s = np.asarray(s, float)
            for b in blocks:
                if b[0].shape == s.shape and np.array_equal(b[0], s):
                    b[1].append(k)
                    break
            else:
                blocks.append((s, [k]))
        self.blocks = [(s, idx, np.stack([np.asarray(I_list[k], complex) for k in idx], 1),
                        np.stack([np.asarray(V_list[k], complex) for k in idx], 1)) for s, idx in blocks]
        self.wI = 1.0 / np.sqrt(sum(float(np.vdot(I, I).real) for _s, _i, I, _V in self.blocks))
        self.wV = np.sqrt(v_weight) / np.sqrt(sum(float(np.vdot(V, V).real) for _s, _i, _I, V in self.blocks))

    def evaluate(self, g):
        """Residual vector (real, stacked), amplitudes (n_sets x 2) and Z0 at
        propagation constant g."""
        amps = np.empty((self.n_sets, 2), complex)
        rI, shapes = [], []
        for s, idx, I, _V in self.blocks:
            ep = np.exp(g * s)
            em = 1.0 / ep
            # two-column QR of [ep, em] by Gram-Schmidt with one reorthogonalisation
            r11 = np.linalg.norm(ep)
            q1 = ep / r11
            r12 = np.vdot(q1, em)
            w = em - r12 * q1
            c = np.vdot(q1, w)
            w = w - c * q1
            r12 += c
            r22 = np.linalg.norm(w)
            q2 = w / r22
            h1 = q1.conj() @ I  # (K,)
            h2 = q2.conj() @ I
            b = h2 / r22
            a = (h1 - r12 * b) / r11
            amps[idx, 0], amps[idx, 1] = a, b
            rI.append((np.outer(q1, h1) + np.outer(q2, h2) - I).ravel())
            shapes.append(np.outer(ep, a) - np.outer(em, b))
        U = np.concatenate([u.ravel() for u in shapes])
        V = np.concatenate([blk[3].ravel() for blk in self.blocks])
        z0 = np.vdot(U, V) / np.vdot(U, U)
        r = np.concatenate([self.wI * np.concatenate(rI), self.wV * (V - z0 * U)])
        return np.concatenate([r.real, r.imag]), amps, z0


def _levenberg_marquardt(fun, x0, lo=None, hi=None, tol=1e-14, max_iter=200):
    """Minimise |fun(x)|^2 from x0 (x inside [lo, hi] when given) with
    central-difference Jacobians; returns x."""
    x = np.array(x0, float)
    r = fun(x)
    f = r @ r
    lam = 1e-4
    for _ in range(max_iter):
        h = 1e-6 * np.maximum(1.0, np.abs(x))
        J = np.empty((r.size, x.size))
        for j in range(x.size):
            e = np.zeros_like(x)
            e[j] = h[j]
            J[:, j] = (fun(x + e) - fun(x - e)) / (2 * h[j])
        A = J.T @ J
        g = J.T @ r
        while True:
            step = np.linalg.solve(A + lam * np.diag(np.diag(A) + 1e-300), -g)
            xn = x + step
            if lo is not None:
                xn = np.clip(xn, lo, hi)
            rn = fun(xn)
            fn = rn @ rn
            if fn <= f:
                lam = max(lam / 5, 1e-15)
                break
            lam *= 8
            if lam > 1e10:  # no descent direction left: at the minimum to round-off
                return x
        moved = np.max(np.abs(xn - x) / np.maximum(1.0, np.abs(x)))
        x, r, f = xn, rn, fn
        if moved <= tol:
            break
    return x


def fit_waves(s_list, I_list, V_list, gamma0, lossless=False, v_weight=VOLTAGE_WEIGHT, bounds=(0.7, 1.3)):
    """Shared-gamma / Z0 fit of I(s) = A e^{gs} + B e^{-gs}, V = Z0 (A e^{gs} - B e^{-gs})
    over all data sets (the objective of Engine/mom25d/solver.py:618-670: the
    normalised current residual of the per-set amplitude fit plus v_weight
    times the normalised voltage residual).  Lossless stacks: g = j beta
    with beta in bounds x Im(gamma0).  Returns gamma, Z0, [(A, B), ...],
    relative residual."""
    scale = abs(gamma0.imag)
    prob = _WaveProblem(s_list, I_list, V_list, v_weight)
    if lossless:
        lo, hi = bounds

        def res(x):
            return prob.evaluate(1j * scale * x[0])[0]

        grid = np.linspace(lo, hi, 25)
        f = [float(res([b]) @ res([b])) for b in grid]
        x = _levenberg_marquardt(res, [grid[int(np.argmin(f))]], lo, hi)
        if min(x[0] - lo, hi - x[0]) < 1e-3:
            warnings.warn(f"beta fit hit its search bound ({x[0]:.4f} x the Kirschning-Jansen estimate)",
                          RuntimeWarning, stacklevel=2)
        g = 1j * scale * x[0]
    else:
        x = _levenberg_marquardt(lambda p: prob.evaluate(scale * (p[0] + 1j * p[1]))[0],
                                 [gamma0.real / scale, gamma0.imag / scale])
        g = scale * (x[0] + 1j * x[1])
    r, amps, z0 = prob.evaluate(g)
    return g, z0, [(a, b) for a, b in amps], float(np.sqrt(r @ r))


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
    through the internal ports' sources (Ii: (n_internal, Nexc)); the
    quantities of Engine/mom25d/solver.py:1284-1325.  Feeds come first, then
    internal ports.  Returns (S, gamma, z0_line, residual, z0_ref, A)."""
    npt, ni = len(ctx.z0_from), len(ctx.zi)
    n = npt + ni
    gam = np.zeros(n, complex)
    zl = np.zeros(n, complex)
    res = np.zeros(n)
    # incident (A) and reflected (B) current-wave amplitudes: row = port, column = excitation
    A = np.zeros((n, n), complex)
    B = np.zeros((n, n), complex)
    for k, members in enumerate(ctx.groups):
        b0 = analytic.beta(ctx.width[k], ctx.h[k], ctx.eps[k], f)
        sets = [(p, e) for p in members for e in range(n)]
        gm, z0, amps, r = fit_waves([ctx.s[p] for p, _ in sets], [I[p][:, e] for p, e in sets],
                                    [V[p][:, e] for p, e in sets], b0 * (1e-3 + 1j), lossless=ctx.lossless,
                                    bounds=(0.7, 1.8) if ctx.covered[k] else (0.7, 1.3), v_weight=ctx.v_weight)
        for (p, e), (a, b) in zip(sets, amps):
            A[p, e], B[p, e] = a, b
        gam[members], zl[members], res[members] = gm, z0, r
    zr = np.zeros(n)
    if ni:
        # an internal port is a source Vs in series with z_ref: port voltage V = Vs - z_ref I, power waves
        # a = Vs / (2 sqrt(z)), b = (V - z_ref I) / (2 sqrt(z)) = (Vs - 2 z_ref I) / (2 sqrt(z)), kept
        # current-normalised (divided by sqrt(z)) like the feeds' amplitudes; B carries the sign of a
        # wave travelling away from the circuit
        zi = np.asarray(ctx.zi, float)
        src = np.zeros((ni, n))
        src[np.arange(ni), npt + np.arange(ni)] = 1.0
        A[npt:] = src / (2 * zi[:, None])
        B[npt:] = (2 * zi[:, None] * np.asarray(Ii) - src) / (2 * zi[:, None])
        zl[npt:] = zi
        zr[npt:] = zi
    # current-normalised scattering matrix: B = -SI A for every excitation
    SI = -np.linalg.solve(A.T, B.T).T
    if npt and ctx.z0_fixed is None:
        zr[:npt] = reciprocal_z0(ctx, SI[:npt, :npt], zl[:npt].real)
        if ni:
            zr[:npt] = anchor_z0(ctx, SI, zr[:npt], ctx.zi)
    elif npt:
        zr[:npt] = np.asarray(ctx.z0_fixed, float)
    root = np.sqrt(zr)
    return SI * np.outer(root, 1.0 / root), gam, zl, res, zr, A


def _pair_weight(SI, i, j, cmin):
    """Weight of the reciprocity estimate from ports i and j: the squared
    weaker coupling, 0 below the coupling floor cmin (or where one is 0)."""
    w = min(abs(SI[i, j]), abs(SI[j, i])) ** 2
    return w if w > 0 and (cmin is None or w >= cmin ** 2) else 0.0


def _log_estimate(SI, known, unknown, z_known):
    """log z of port `unknown` from port `known` by reciprocity: z_u = z_k SI_ku / SI_uk."""
    return np.log(z_known * abs(SI[known, unknown] / SI[unknown, known]))


def _weighted_mean_log(pairs, prior_log, prior):
    """exp of the prior-regularised weighted mean of log estimates [(w, log z)]."""
    num = prior * prior_log + sum(w * v for w, v in pairs)
    return np.exp(num / (prior + sum(w for w, _ in pairs)))


def anchor_z0(ctx, SI, zfeed, zi, prior=1e-4):
    """Feed reference impedances refined by reciprocity with the internal
    ports, whose z_ref is exact: each feed group takes the weighted mean of
    the estimates from every (internal port, member) pair, with its
    feed-only value as a weak prior (Engine/mom25d/solver.py:1337-1355)."""
    npt = len(zfeed)
    z = np.array(zfeed, float)
    for members in ctx.groups:
        pairs = []
        for j in members:
            for k, zk in enumerate(zi):
                w = _pair_weight(SI, npt + k, j, ctx.z0_recip_min)
                if w:
                    pairs.append((w, _log_estimate(SI, npt + k, j, zk)))
        z[members] = _weighted_mean_log(pairs, np.log(np.mean(z[members])), prior)
    return z


def reciprocal_z0(ctx, SI, zphi, prior=1e-4):
    """Port reference impedances (Engine/mom25d/solver.py:1357-1397): groups
    with z0_from='voltage' from their scalar-potential levels tied together
    by reciprocity (voltage_z0), groups with z0_from='reciprocity' as the
    weighted mean of the estimates from every voltage port they couple to."""
    groups = ctx.groups
    by_recip = [ctx.z0_from[m[0]] == "reciprocity" for m in groups]
    if all(by_recip):
        raise ValueError("every port has z0_from='reciprocity': at least one needs 'voltage'")
    z = np.array(zphi, float)
    vgroups = [g for g, r in zip(groups, by_recip) if not r]
    vports = np.array(sorted(p for g in vgroups for p in g))
    local = {p: i for i, p in enumerate(vports)}
    zv = voltage_z0(SI[np.ix_(vports, vports)], z[vports], [[local[p] for p in g] for g in vgroups], prior,
                    ctx.z0_recip_min)
    z[vports] = zv
    for k, members in enumerate(groups):
        if not by_recip[k]:
            continue
        pairs = [(w, _log_estimate(SI, j, i, z[j])) for i in members for j in vports
                 if (w := _pair_weight(SI, i, j, ctx.z0_recip_min))]
        if not pairs:
            raise ValueError(f"port group {k} (z0_from='reciprocity') does not couple to any voltage port")
        z[members] = _weighted_mean_log(pairs, 0.0, 0.0)
    return z


def voltage_z0(SI, zphi, groups, prior, cmin=None):
    """Reference impedances of voltage-referenced port groups
    (Engine/mom25d/solver.py:1399-1427): per group one unknown x = log z,
    minimising sum w_ij (x_gi - x_gj - log|SI_ji / SI_ij|)^2 over coupled
    ports of different groups plus prior (x_g - log mean zphi_g)^2 -- the
    weighted graph-Laplacian system (L + prior I) x = rhs -- then shifted so
    that the mean of log z over the ports equals that of zphi."""
    zphi = np.asarray(zphi, float)
    if len(groups) == 1:
        return zphi
    ng = len(groups)
    gid = np.empty(len(zphi), int)
    for k, members in enumerate(groups):
        gid[members] = k
    L = prior * np.eye(ng)
    rhs = prior * np.array([np.log(np.mean(zphi[m])) for m in groups])
    for i in range(len(gid)):
        for j in range(i + 1, len(gid)):
            a, b = gid[i], gid[j]
            if a == b:
                continue
            w = _pair_weight(SI, i, j, cmin)
            if not w:
                continue
            t = np.log(abs(SI[j, i] / SI[i, j]))  # x_a - x_b
            L[a, a] += w
            L[b, b] += w
            L[a, b] -= w
            L[b, a] -= w
            rhs[a] += w * t
            rhs[b] -= w * t
    x = np.linalg.solve(L, rhs)
    x += np.mean(np.log(zphi)) - np.mean(x[gid])
    return np.exp(x[gid])
