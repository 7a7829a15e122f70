"""Tabulated spatial Green's functions of a sweep, built with one GPU launch.

For each frequency and each pair of metal levels the MPIE kernels are split
into a closed-form leading part and a tabulated smooth remainder
(Engine/mom25d/greens.py:68-228, Engine/mom25d/layered.py:212-415):

    g(R) = c exp(-j k R) / (4 pi R) + S(rho),   R = sqrt(rho^2 + dz^2)

  * path   : a lifted path kr = t + j a sin(pi t / T) on [0, T] (12-point
             Gauss-Legendre panels), then the real axis up to kmax (8-point
             panels), as in Engine/mom25d/greens.py:138-155;
  * spectra: g~(kr) minus its asymptote c exp(-u dz) / (2u), on the host;
  * Sommerfeld: R(rho) = (1/2 pi) sum J0(kr rho) kr w g~ -- the one
             expensive step, run by the `mom_sommerfeld` GPU kernel for all
             rows of all frequencies of the sweep in one launch;
  * table  : S = c expm1(-j k R)/(4 pi R) + R(rho) (+ the ground-plane loss
             term in S_A), on the grid r_{i+1} = r_i + min(lam/80, max(h/25, r_i/25));
  * spline : not-a-knot cubic spline of [S_A, S_V] and of S_W (scipy
             CubicSpline semantics), solved for all frequencies at once.

The splines are uploaded in the layout the fill kernel reads (16 doubles per
interval: c0..c3 of S_A, S_V, re / im) together with a bucket table that
turns the interval search into one lookup and one comparison (bucket_table).
"""

import functools

import numpy as np
from scipy.linalg import solve_banded
import torch

from . import _lib
from .constants import C0, MU0
from .layered import csqrt_re_pos, layered_spectra

PAIR_DTYPE = np.dtype([("x_off", "<i8"), ("c_off", "<i8"), ("b_off", "<i8"), ("n", "<i4"), ("nbucket", "<i4"),
                       ("inv_dx", "<f8"), ("dz", "<f8"), ("cA_re", "<f8"), ("cA_im", "<f8"), ("cV_re", "<f8"),
                       ("cV_im", "<f8"), ("cW_re", "<f8"), ("cW_im", "<f8"), ("zero", "<i4"), ("pad_", "<i4")])


@functools.lru_cache(maxsize=None)
def _leggauss(order):
    return np.polynomial.legendre.leggauss(order)


def gauss_panels(a, b, n_panels, order):
    """Composite Gauss-Legendre nodes / weights on [a, b]."""
    x, w = _leggauss(order)
    edges = np.linspace(a, b, n_panels + 1)
    half = 0.5 * np.diff(edges)
    mid = 0.5 * (edges[1:] + edges[:-1])
    return (mid[:, None] + half[:, None] * x[None, :]).ravel(), (half[:, None] * w[None, :]).ravel()


def rho_grid(lam, h, rmax):
    """Table nodes: fine near the source, never coarser than lam / 80."""
    r, x = [0.0], 0.0
    fine, floor = lam / 80, h / 25
    while x < rmax:
        x = x + min(fine, max(floor, x / 25))
        r.append(x)
    return np.array(r)


BUCKET_DTYPE = np.dtype([("split", "<f8"), ("i0", "<i4"), ("pad", "<i4")])


def bucket_table(rho, dxb):
    """Interval lookup of the fill kernels for the nodes rho: entry b serves
    every distance R >= 0 with int(R * (1 / dxb)) == b in double arithmetic
    (the last entry also every larger R), so its range [lo_b, lo_b+1) is
    exact for the kernels' own rounding.  i0 is the spline interval of lo_b
    and split the one node inside the range (+inf if none, or if crossing it
    does not change the clamped interval); the interval of R is then
    i0 + (R >= split), equal to scipy's clamped
    searchsorted(rho, R, 'right') - 1.  dxb must be at most half the
    smallest node spacing, which leaves at most one node per range."""
    n = rho.size
    inv = 1.0 / dxb
    nb = int(rho[-1] * inv) + 2
    b = np.arange(nb, dtype=np.float64)
    lo = b * dxb  # then the smallest double with floor(lo * inv) >= b
    for _ in range(64):
        dn = np.nextafter(lo, -np.inf)
        m = (np.floor(dn * inv) >= b) & (dn >= 0)
        if not m.any():
            break
        lo = np.where(m, dn, lo)
    for _ in range(64):
        m = np.floor(lo * inv) < b
        if not m.any():
            break
        lo = np.where(m, np.nextafter(lo, np.inf), lo)
    hi = np.append(lo[1:], np.inf)
    raw = np.searchsorted(rho, lo, side="right") - 1
    j = raw + 1  # first node above lo
    inside = np.searchsorted(rho, hi, side="left") - j
    if np.any(inside > 1):
        raise RuntimeError("bucket_table: a bucket holds more than one node (width above half the node spacing)")
    t = np.zeros(nb, BUCKET_DTYPE)
    t["i0"] = np.clip(raw, 0, n - 2)
    moves = (inside >= 1) & (raw >= 0) & (j <= n - 2)
    t["split"] = np.where(moves, rho[np.minimum(j, n - 1)], np.inf)
    return t


def spline_coeffs(x, Y):
    """Not-a-knot cubic spline coefficients c (4, n-1, m) of Y (n, m) on
    nodes x (n >= 4), in scipy PPoly order (c[0] multiplies dx^3)."""
    n = x.size
    dx = np.diff(x)
    slope = np.diff(Y, axis=0) / dx[:, None]
    sub = np.zeros(n)
    dia = np.zeros(n)
    sup = np.zeros(n)
    rhs = np.empty(Y.shape, complex)
    dia[1:-1] = 2 * (dx[:-1] + dx[1:])
    sup[1:-1] = dx[:-1]
    sub[1:-1] = dx[1:]
    rhs[1:-1] = 3 * (dx[1:, None] * slope[:-1] + dx[:-1, None] * slope[1:])
    d0 = x[2] - x[0]
    dia[0], sup[0] = dx[1], d0
    rhs[0] = ((dx[0] + 2 * d0) * dx[1] * slope[0] + dx[0] ** 2 * slope[1]) / d0
    d1 = x[-1] - x[-3]
    dia[-1], sub[-1] = dx[-2], d1
    rhs[-1] = (dx[-1] ** 2 * slope[-2] + (2 * d1 + dx[-1]) * dx[-2] * slope[-1]) / d1
    # the tridiagonal system (first and last rows carry the not-a-knot conditions), banded LAPACK solve
    ab = np.zeros((3, n))
    ab[0, 1:] = sup[:-1]
    ab[1] = dia
    ab[2, :-1] = sub[1:]
    s = solve_banded((1, 1), ab, rhs, check_finite=False)
    t = (s[:-1] + s[1:] - 2 * slope) / dx[:, None]
    c = np.empty((4, n - 1, Y.shape[1]), complex)
    c[0] = t / dx[:, None]
    c[1] = (slope - s[:-1]) / dx[:, None] - t
    c[2] = s[:-1]
    c[3] = Y[:-1]
    return c


def ppoly_eval(x, c, R):
    """Evaluate PPoly coefficients c (4, n-1, ...) at R (interval by search,
    clamped to the end intervals)."""
    R = np.asarray(R, float)
    i = np.clip(np.searchsorted(x, R, side="right") - 1, 0, x.size - 2)
    d = (R - x[i])[(...,) + (None,) * (c.ndim - 2)]
    return ((c[0, i] * d + c[1, i]) * d + c[2, i]) * d + c[3, i]


class KernelPair:
    """Kernels between an observer and a source level at one frequency:
    g = c / (4 pi R) + S(rho) with R = sqrt(rho^2 + dz^2)."""

    zero = False

    def __init__(self, dz, cA, cV, cW, x, c, cw):
        self.dz, self.cA, self.cV, self.cW = float(dz), complex(cA), complex(cV), complex(cW)
        self.cv = self.cV
        self.x, self.c, self.cw = x, c, cw

    def smooth(self, rho):
        S = ppoly_eval(self.x, self.c, rho)
        return S[..., 0], S[..., 1]

    def smooth_w(self, rho):
        return ppoly_eval(self.x, self.cw, rho)

    def _R(self, rho):
        return np.sqrt(np.asarray(rho) ** 2 + self.dz ** 2)

    def gA(self, rho):
        return self.cA / (4 * np.pi * self._R(rho)) + self.smooth(rho)[0]

    def gV(self, rho):
        return self.cV / (4 * np.pi * self._R(rho)) + self.smooth(rho)[1]

    def gW(self, rho):
        return self.cW / (4 * np.pi * self._R(rho)) + self.smooth_w(rho)


def sommerfeld_path(k0, kk, t, rho_max, order=12, tail=True):
    """Nodes and weights of the Sommerfeld integrals of a stack
    (Engine/mom25d/layered.py:212-236): a path lifted into the complex plane
    by a sin(pi k / T) bump of height a = min(k0 / 2, 1 / rho_max) over
    [0, T], T = 1.3 max |k_i| + k0, in order-point Gauss panels (100..3000),
    then (tail) the real axis from T to kmax = max(40 max |k_i|, 20 / t_min,
    1.5 T) in 8-point panels of width min(pi / rho_max, 0.5 / t_min,
    max |k_i|).  Returns kc, wc, kr, wr (kr, wr empty without the tail)."""
    kmaxr = np.abs(kk).max()
    tmin = np.min(t)
    T = 1.3 * kmaxr + k0
    a = min(0.5 * k0, 1.0 / rho_max)
    n1 = int(np.clip(np.ceil(2 * T / a), 100, 3000))
    tt, ww = gauss_panels(0.0, T, n1, order)
    kc = tt + 1j * a * np.sin(np.pi * tt / T)
    wc = ww * (1 + 1j * a * np.pi / T * np.cos(np.pi * tt / T))
    if not tail:
        return kc, wc, np.zeros(0), np.zeros(0)
    kmax = max(40 * kmaxr, 20.0 / tmin, T * 1.5)
    dk = min(np.pi / rho_max, 0.5 / tmin, kmaxr)
    n2 = int(np.ceil((kmax - T) / dk))
    kr, wr = gauss_panels(T, kmax, n2, 8)
    return kc, wc, kr, wr


class FreqGreens:
    """Green's functions of all level pairs at one frequency (host view plus
    the location of its tables in the sweep's device arrays)."""

    def __init__(self, freq, h, levels, pairs, stack, tables, index):
        self.freq = freq
        self.omega = 2 * np.pi * freq
        self.k0 = self.omega / C0
        self.h = h
        self.levels = list(levels)
        self._pairs = pairs
        self.stack = stack
        self.tables = tables
        self.index = index

    def index_of(self, level):
        return level if isinstance(level, (int, np.integer)) else self.levels.index(level)

    def pair(self, i=0, j=0):
        return self._pairs[self.index_of(i), self.index_of(j)]

    @property
    def pairs_dev(self):
        """Device PairTab array (L*L entries, observer-major) of this frequency."""
        return self.tables.pairs_dev[self.index]

    def via_sections(self):
        return self.stack.via_sections(self.freq)


class SweepTables:
    """Tabulated kernels of a stack at every frequency of a sweep.

    single : one dielectric with copper on top and only the "top" level:
             the grounded-slab closed-form spectra (Engine/mom25d/greens.py:111-135);
             otherwise the transmission-line spectra of the stack.
    """

    def __init__(self, stack, freqs, rho_max, levels, single, device="cuda", gl_order=12, surrogate=None,
                 surrogate_tol=5e-5):
        self._setup(stack, rho_max, levels, single, device)
        self.freqs = np.atleast_1d(np.asarray(freqs, float))
        jobs = [self._prepare(f, gl_order) for f in self.freqs]
        self.integral_nodes_total = self.integral_nodes(jobs)
        if surrogate is None:
            res = self._sommerfeld(jobs)
            self.surrogate_report = None
        else:
            res = self._surrogate_remainders(jobs, surrogate, surrogate_tol)
        self._finish(jobs, res)

    def _setup(self, stack, rho_max, levels, single, device):
        self.stack = stack
        self.rho_max = float(rho_max)
        self.levels = list(levels)
        self.single = bool(single)
        self.device = torch.device(device)
        L = len(self.levels)
        self.L = L
        st = stack
        self.il = np.array([st.metals[n] for n in self.levels])
        self.zl = np.array([st.level_z(n) for n in self.levels])
        self.epsc = np.array([d.eps_r * (1 - 1j * d.tand) for d in st.layers])
        self.t = np.array([d.t for d in st.layers])
        self.h = float(st.z[-1])

    @classmethod
    def spectra_only(cls, stack, rho_max, levels, single=False, device="cuda"):
        """An instance with the stack set up and no tables (for sampling the
        remainders at arbitrary points)."""
        obj = cls.__new__(cls)
        obj._setup(stack, rho_max, levels, single, device)
        return obj

    # ---------------------------------------------------------- per frequency
    def _prepare(self, f, order):
        st, L = self.stack, self.L
        omega = 2 * np.pi * f
        k0 = omega / C0
        kk = k0 * np.sqrt(self.epsc)
        tmin = self.t.min()
        kmaxr = np.abs(kk).max()
        kc, wc, kr, wr = sommerfeld_path(k0, kk, self.t, self.rho_max, order)

        dz = np.abs(self.zl[:, None] - self.zl[None, :])
        c = np.zeros((3, L, L), complex)
        kx = np.zeros((2, L, L), complex)
        if self.single:
            epsc, k1 = self.epsc[0], kk[0]
            cv = 2.0 / (epsc + 1.0)
            c[:, 0, 0] = (1.0, cv, cv)
            kx[0, 0, 0] = np.sqrt(0.5 * (k0 ** 2 + k1 ** 2))
            kx[1, 0, 0] = np.sqrt((epsc * k0 ** 2 + k1 ** 2) / (epsc + 1.0))
        else:
            self._leading(k0, kk, dz, c, kx)

        def remainders(k):
            if self.single:
                g = _slab_spectra(k, k0, kk[0], self.epsc[0], self.t[0])
                g = [gi[None, None] for gi in g]
            else:
                g = layered_spectra(k, k0, kk, self.epsc, self.t, self.il, st.top == "pec")
            out = []
            for m in range(3):
                ux = np.sqrt(k[None, None] ** 2 - kx[min(m, 1)][..., None] ** 2)
                asym = c[m][..., None] / (2 * ux) if not dz.any() else c[m][..., None] * np.exp(-ux * dz[..., None]) / (2 * ux)
                out.append(g[m] - asym)
            return np.stack(out).reshape(3 * L * L, k.size)

        Cc = (wc * kc / (2 * np.pi))[None] * remainders(kc.astype(complex))
        Cr = (wr.astype(complex) * kr / (2 * np.pi))[None] * remainders(kr.astype(complex))
        lam = 2 * np.pi / kmaxr
        rho = rho_grid(lam, tmin, self.rho_max)
        zs = st.surface_impedance(f) if st.ground_loss else 0.0
        return dict(f=f, omega=omega, k0=k0, kc=kc, Cc=Cc.T.copy(), kr=kr, Cr=Cr.T.copy(), rho=rho, c=c, kx=kx,
                    dz=dz, zs=zs)

    def remainders_at(self, freqs, rhos, gl_order=12):
        """Sommerfeld remainders (n_i, 3, L, L) at rhos[i] for freqs[i], and
        the leading-term wavenumbers kx (2, L, L) of each frequency."""
        jobs = [self._prepare(f, gl_order) for f in np.atleast_1d(freqs)]
        parts = self._sommerfeld(jobs, rhos=rhos)
        L = self.L
        return [p.reshape(-1, 3, L, L) for p in parts], [j["kx"] for j in jobs]

    def _leading(self, k0, kk, dz, c, kx):
        """Leading terms per pair (Engine/mom25d/layered.py:292-331)."""
        L, n = self.L, len(self.t)
        Kbig = 15.0 / self.t.min()
        big = None
        for i in range(L):
            for j in range(L):
                if i == j:
                    m = self.il[i]
                    eb, kb = self.epsc[m - 1], kk[m - 1]
                    ea, ka = (1.0, k0) if m == n else (self.epsc[m], kk[m])
                    cv = 2.0 / (ea + eb)
                    c[:, i, j] = (1.0, cv, cv)
                    kx[0, i, j] = np.sqrt(0.5 * (ka ** 2 + kb ** 2))
                    kx[1, i, j] = np.sqrt((eb * ka ** 2 + ea * kb ** 2) / (ea + eb))
                    continue
                if big is None:
                    g = layered_spectra(np.array([Kbig], complex), k0, kk, self.epsc, self.t, self.il,
                                        self.stack.top == "pec")
                    big = [gi[..., 0] for gi in g]
                ux = np.sqrt(Kbig ** 2 - k0 ** 2 + 0j)
                f = 2 * ux * np.exp(ux * dz[i, j])
                cA, cV, cW = (f * g[i, j] for g in big)
                if self.zl[j] < self.zl[i]:
                    cW = 0.0
                c[:, i, j] = (cA, cV, cW)
                kx[:, i, j] = k0
        small = np.abs(c.imag) < 1e-9 * np.abs(c)
        c[small] = c[small].real

    # ------------------------------------------------------------ Sommerfeld
    def _sommerfeld(self, jobs, cols=None, rhos=None):
        """Remainders of every job at its table nodes (or at rhos[i] for job
        i), columns cols of the 3 L^2 kernels (default: all), one launch."""
        dev = self.device
        ncol_all = 3 * self.L * self.L
        cols = np.arange(ncol_all) if cols is None else np.asarray(cols)
        ncol = cols.size
        rl = [j["rho"] for j in jobs] if rhos is None else [np.asarray(r, float) for r in rhos]
        rho = np.concatenate(rl)
        if rho.size == 0:
            return [np.zeros((0, ncol), complex) for _ in jobs]
        prob = np.concatenate([np.full(r.size, p, np.int32) for p, r in enumerate(rl)])
        kc_len = np.array([j["kc"].size for j in jobs], np.int32)
        kr_len = np.array([j["kr"].size for j in jobs], np.int32)
        kc_off = np.concatenate([[0], np.cumsum(kc_len)[:-1]]).astype(np.int64)
        kr_off = np.concatenate([[0], np.cumsum(kr_len)[:-1]]).astype(np.int64)
        t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)  # noqa: E731
        d_rho = t(rho, torch.float64)
        d_prob = t(prob, torch.int32)
        d_kc_off, d_kc_len = t(kc_off, torch.int64), t(kc_len, torch.int32)
        d_kr_off, d_kr_len = t(kr_off, torch.int64), t(kr_len, torch.int32)
        d_kc = t(np.concatenate([j["kc"] for j in jobs]).astype(complex), torch.complex128)
        d_Cc = t(np.concatenate([j["Cc"][:, cols] for j in jobs]), torch.complex128)
        d_kr = t(np.concatenate([j["kr"] for j in jobs]), torch.float64)
        d_Cr = t(np.concatenate([j["Cr"][:, cols] for j in jobs]), torch.complex128)
        out = torch.empty((rho.size, ncol), dtype=torch.complex128, device=dev)
        P = _lib.ptr
        _lib.call("mom_sommerfeld", int(rho.size), P(d_rho), P(d_prob), P(d_kc_off), P(d_kc_len), P(d_kc), P(d_Cc),
                  P(d_kr_off), P(d_kr_len), P(d_kr), P(d_Cr), ncol, P(out), _lib.stream_ptr())
        res = out.cpu().numpy()
        parts, o = [], 0
        for r in rl:
            parts.append(res[o:o + r.size])
            o += r.size
        return parts

    @staticmethod
    def surrogate_nodes(rho, n_fit, n_check):
        """Table-node indices where the surrogate's tables take the integrals:
        fit nodes log-spaced in rho over the whole table (both ends, about
        n_fit of them once coincident ones merge), and up to n_check check
        nodes, each midway between two fit nodes, spread over the table."""
        n = rho.size
        x = np.log(np.maximum(rho, 0.5 * rho[1]))
        fit = np.unique(np.round(np.interp(np.linspace(x[0], x[-1], n_fit), x, np.arange(n))).astype(int))
        gaps = np.nonzero(np.diff(fit) >= 2)[0]
        mids = (fit[gaps] + fit[gaps + 1]) // 2
        chk = mids[np.unique(np.round(np.linspace(0, mids.size - 1, min(n_check, mids.size))).astype(int))] \
            if mids.size else np.zeros(0, int)
        return fit, chk, x

    def _surrogate_remainders(self, jobs, sur, tol, n_fit=64, n_check=16):
        """Remainder columns of every job from the surrogate networks (R_A,
        R_V, R_W at every table node), corrected by the integrals at the fit
        nodes and judged by them at the check nodes (surrogate_nodes):

          - the networks' error in the demodulated remainder T at the fit
            nodes, interpolated over log rho by a cubic spline, is taken off
            every node (the network is smooth in log rho, so is its error),
            and the fit nodes keep the integrals' values;
          - a frequency whose largest error at the check nodes, relative to
            the largest |R| of the same kernel and level pair, exceeds tol
            takes all its columns from the integrals."""
        import time

        from scipy.interpolate import CubicSpline

        L = self.L
        LL = L * L
        f_lo, f_hi = sur.f_range
        if self.rho_max > sur.rho_range[1] or any(not (f_lo * (1 - 1e-9) <= j["f"] <= f_hi * (1 + 1e-9)) for j in jobs):
            self.surrogate_report = dict(check_err=np.full(len(jobs), np.inf), fallback=list(range(len(jobs))),
                                         tol=tol, n_fit=n_fit, n_check=n_check, seconds={}, integral_nodes=None,
                                         reason="outside the trained rho / frequency range")
            return self._sommerfeld(jobs)
        t0 = time.perf_counter()
        nodes = [self.surrogate_nodes(j["rho"], n_fit, n_check) for j in jobs]
        exact = self._sommerfeld(jobs, rhos=[j["rho"][np.concatenate([fi, ci])] for j, (fi, ci, _) in zip(jobs, nodes)])
        t1 = time.perf_counter()
        cols = []
        for j in jobs:
            n = j["rho"].size
            for i in range(L):
                for k in range(L):
                    cols.append((j["rho"], np.full(n, self.zl[i]), np.full(n, self.zl[k]), np.full(n, j["k0"]),
                                 np.full(n, j["kx"][0, i, k]), np.full(n, j["kx"][1, i, k])))
        rho, zi, zj, k0, kxA, kxV = (np.concatenate(c) for c in zip(*cols))
        RA, RV, RW = sur.remainders(rho, zi, zj, k0, kxA, kxV)
        t2 = time.perf_counter()
        res, errs, fallback = [], [], []
        o = 0
        for fi, j in enumerate(jobs):
            n = j["rho"].size
            r = j["rho"]
            R = np.empty((n, 3 * LL), complex)
            for p in range(LL):
                sl = slice(o + p * n, o + (p + 1) * n)
                R[:, p], R[:, LL + p], R[:, 2 * LL + p] = RA[sl], RV[sl], RW[sl]
            o += LL * n
            # demodulation of each column: A with kx_A, V and W with kx_V of its pair
            kx = np.concatenate([j["kx"][0].ravel(), j["kx"][1].ravel(), j["kx"][1].ravel()])
            dem = (4 * np.pi * np.sqrt(r ** 2 + sur.h ** 2))[:, None] * np.exp(1j * np.outer(r, kx))
            fit, chk, x = nodes[fi]
            ex_fit, ex_chk = exact[fi][:fit.size], exact[fi][fit.size:]
            E = (ex_fit - R[fit]) * dem[fit]
            sp = CubicSpline(x[fit], np.concatenate([E.real, E.imag], 1))(x)
            R = R + (sp[:, :3 * LL] + 1j * sp[:, 3 * LL:]) / dem
            R[fit] = ex_fit
            scale = np.maximum(np.abs(exact[fi]).max(0), 1e-300)
            err = float(np.max(np.abs(R[chk] - ex_chk).max(0) / scale)) if chk.size else np.inf
            errs.append(err)
            if not err <= tol:
                fallback.append(fi)
            res.append(R)
        if fallback:
            full = self._sommerfeld([jobs[fi] for fi in fallback])
            for fi, F in zip(fallback, full):
                res[fi] = F
        t3 = time.perf_counter()
        used = sum((f.size + c.size) * (j["kc"].size + j["kr"].size) for j, (f, c, _) in zip(jobs, nodes))
        used += sum(jobs[fi]["rho"].size * (jobs[fi]["kc"].size + jobs[fi]["kr"].size) for fi in fallback)
        self.surrogate_report = dict(check_err=np.array(errs), fallback=fallback, tol=tol, n_fit=n_fit,
                                     n_check=n_check, seconds=dict(integrals=t1 - t0 + (t3 - t2), network=t2 - t1),
                                     integral_nodes=used, reason=None)
        return res

    def integral_nodes(self, jobs):
        """Bessel evaluations of the plain tables (table nodes x spectral nodes)."""
        return sum(j["rho"].size * (j["kc"].size + j["kr"].size) for j in jobs)

    # ----------------------------------------------------------------- tables
    def _finish(self, jobs, res):
        L = self.L
        xs, cs, cws, bks = [], [], [], []
        x_off = c_off = b_off = 0
        pair_rows = []
        self.greens = []
        for fi, (j, RES) in enumerate(zip(jobs, res)):
            rho = j["rho"]
            RES = RES.T.reshape(3, L, L, rho.size)
            pairs = {}
            recs = np.zeros(L * L, PAIR_DTYPE)
            dxb = 0.5 * np.diff(rho).min()
            bucket = bucket_table(rho, dxb)
            nbucket = bucket.size
            xs.append(rho)
            bks.append(bucket)
            for i in range(L):
                for k in range(L):
                    dz = j["dz"][i, k]
                    R = np.sqrt(rho ** 2 + dz ** 2)
                    S = []
                    for m in range(3):
                        kw = j["kx"][min(m, 1), i, k]
                        with np.errstate(divide="ignore", invalid="ignore"):
                            f = np.where(R > 0, np.expm1(-1j * kw * R) / (4 * np.pi * R), -1j * kw / (4 * np.pi))
                        S.append(j["c"][m, i, k] * f + RES[m, i, k])
                    if j["zs"]:
                        d = self.zl[i] + self.zl[k]
                        S[0] = S[0] + j["zs"] / (1j * j["omega"] * MU0) * d / (2 * np.pi * (rho ** 2 + d ** 2) ** 1.5)
                    coef = spline_coeffs(rho, np.stack(S, -1))  # (4, n-1, 3)
                    c2, cw = coef[..., :2], coef[..., 2]
                    pairs[i, k] = KernelPair(dz, *j["c"][:, i, k], rho, c2, cw)
                    cs.append(np.ascontiguousarray(c2.transpose(1, 0, 2)).view(np.float64).reshape(-1))
                    cws.append(np.ascontiguousarray(cw.T).view(np.float64).reshape(-1))
                    r = recs[i * L + k]
                    r["x_off"], r["c_off"], r["b_off"] = x_off, c_off, b_off
                    r["n"], r["nbucket"], r["inv_dx"], r["dz"] = rho.size, nbucket, 1.0 / dxb, dz
                    cA, cV = complex(j["c"][0, i, k]), complex(j["c"][1, i, k])
                    cW = complex(j["c"][2, i, k])
                    r["cA_re"], r["cA_im"], r["cV_re"], r["cV_im"] = cA.real, cA.imag, cV.real, cV.imag
                    r["cW_re"], r["cW_im"] = cW.real, cW.imag
                    c_off += 16 * (rho.size - 1)
            x_off += rho.size
            b_off += nbucket
            pair_rows.append(recs)
            self.greens.append(FreqGreens(j["f"], self.h, self.levels, pairs, self.stack, self, fi))
        dev = self.device
        self.tab_x = torch.from_numpy(np.concatenate(xs)).to(dev)
        self.tab_c = torch.from_numpy(np.concatenate(cs)).to(dev)
        self.tab_cw = torch.from_numpy(np.concatenate(cws)).to(dev)
        self.tab_b = torch.from_numpy(np.concatenate(bks).view(np.uint8).copy()).to(dev)  # BUCKET_DTYPE entries
        allrec = np.concatenate(pair_rows)
        self._pairs_blob = torch.from_numpy(allrec.view(np.uint8).copy()).to(dev)
        rec_bytes = PAIR_DTYPE.itemsize * L * L
        self.pairs_dev = [self._pairs_blob[fi * rec_bytes:(fi + 1) * rec_bytes] for fi in range(len(jobs))]
        self.pair_records = pair_rows

    def __getitem__(self, fi):
        return self.greens[fi]

    def __len__(self):
        return len(self.greens)


class ZeroPair:
    """Kernels between levels of different stack sections: a solid plane
    between them, so none (Engine/mom25d/layered.py:418-434)."""

    zero = True

    def __init__(self, dz):
        self.dz = float(dz)
        self.cA = self.cV = self.cW = self.cv = 0j

    def smooth(self, rho):
        z = np.zeros(np.shape(rho), complex)
        return z, z

    def smooth_w(self, rho):
        return np.zeros(np.shape(rho), complex)

    gA = gV = gW = smooth_w


class SectionedTables:
    """Tables of a stack with inner planes (Engine/mom25d/layered.py:437-474):
    one SweepTables per section -- its own stack over the plane below it,
    the levels of that section -- merged into one set of device arrays in
    which the records of level pairs in different sections are marked as
    uncoupled.  The fill forms no work for those pairs (its plan keeps
    sources in the test level's section) and the line-voltage kernel skips
    them; on the host they are ZeroPair."""

    def __init__(self, stack, freqs, rho_max, levels, device="cuda", gl_order=12):
        self.stack = stack
        self.levels = list(levels)
        self.freqs = np.atleast_1d(np.asarray(freqs, float))
        self.L = L = len(self.levels)
        self.zl = np.array([stack.level_z(n) for n in self.levels])
        sec = [stack.section(n) for n in self.levels]
        subs = {}
        for q in sorted(set(sec)):
            names = [n for n, t in zip(self.levels, sec) if t == q]
            subs[q] = (names, SweepTables(stack.section_stack(q), self.freqs, rho_max, names, False, device, gl_order))
        base, xs, cs, cws, bs = {}, [], [], [], []
        x0 = c0 = b0 = 0
        for q, (names, t) in subs.items():
            base[q] = (x0, c0, b0)
            xs.append(t.tab_x)
            cs.append(t.tab_c)
            cws.append(t.tab_cw)
            bs.append(t.tab_b)
            x0 += t.tab_x.numel()
            c0 += t.tab_c.numel()
            b0 += t.tab_b.numel() // BUCKET_DTYPE.itemsize
        self.tab_x, self.tab_c, self.tab_cw, self.tab_b = (torch.cat(a) for a in (xs, cs, cws, bs))
        self.greens, rows = [], []
        for fi, f in enumerate(self.freqs):
            recs = np.zeros(L * L, PAIR_DTYPE)
            pairs = {}
            for i in range(L):
                for k in range(L):
                    if sec[i] != sec[k]:
                        recs[i * L + k]["dz"] = abs(self.zl[i] - self.zl[k])
                        recs[i * L + k]["zero"] = 1
                        pairs[i, k] = ZeroPair(abs(self.zl[i] - self.zl[k]))
                        continue
                    names, t = subs[sec[i]]
                    a, b = names.index(self.levels[i]), names.index(self.levels[k])
                    r = t.pair_records[fi][a * len(names) + b].copy()
                    xb, cb, bb = base[sec[i]]
                    r["x_off"] += xb
                    r["c_off"] += cb
                    r["b_off"] += bb
                    recs[i * L + k] = r
                    pairs[i, k] = t.greens[fi].pair(a, b)
            rows.append(recs)
            self.greens.append(FreqGreens(f, float(stack.z[-1]), self.levels, pairs, stack, self, fi))
        allrec = np.concatenate(rows)
        self._pairs_blob = torch.from_numpy(allrec.view(np.uint8).copy()).to(device)
        rec_bytes = PAIR_DTYPE.itemsize * L * L
        self.pairs_dev = [self._pairs_blob[fi * rec_bytes:(fi + 1) * rec_bytes] for fi in range(len(self.freqs))]
        self.pair_records = rows
        self.surrogate_report = None
        self.integral_nodes_total = sum(t.integral_nodes_total for _, t in subs.values())
        self.sections = subs

    def __getitem__(self, fi):
        return self.greens[fi]

    def __len__(self):
        return len(self.greens)


def sommerfeld_remainders(stack, levels, freqs, rhos, rho_max, device="cuda"):
    """Remainders of a (multi-level) stack at arbitrary points: list of
    (n_i, 3, L, L) arrays for rhos[i] at freqs[i], and the kx (2, L, L)."""
    tabs = SweepTables.spectra_only(stack, rho_max, levels, False, device)
    return tabs.remainders_at(freqs, rhos)


def _slab_spectra(kr, k0, k1, epsc, h):
    """(gA~, gV~, gW~) of a grounded slab with conductors on top
    (Engine/mom25d/greens.py:111-129)."""
    u0 = np.sqrt(kr ** 2 - k0 ** 2)
    u1 = csqrt_re_pos(kr ** 2 - k1 ** 2)
    x = u1 * h
    e = np.exp(-2 * x)
    small = np.abs(x) < 1e-4
    with np.errstate(divide="ignore", invalid="ignore"):
        u1coth = np.where(small, 1 / h + u1 ** 2 * h / 3, u1 * (1 + e) / (1 - e))
        u1tanh = np.where(small, u1 ** 2 * h, u1 * (1 - e) / (1 + e))
        tanh_over_u1 = np.where(small, h, u1tanh / u1 ** 2)
    dte = u0 + u1coth
    dtm = epsc * u0 + u1tanh
    return 1.0 / dte, (u0 + u1tanh) / (dte * dtm), u0 * tanh_over_u1 / dtm
