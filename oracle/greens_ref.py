"""Oracle: tabulated MPIE Green's functions of a grounded slab and of a
layered stack, with scipy's Bessel functions and cubic spline, following
Engine/mom25d/greens.py:53-228 and Engine/mom25d/layered.py:183-415.
Test infrastructure only."""

import numpy as np
from scipy.interpolate import CubicSpline
from scipy.special import j0, jv

C0 = 299792458.0
MU0 = 4e-7 * np.pi


def panels(a, b, n, order):
    """Gauss-Legendre composite rule on [a, b] with n panels (greens.py:53-60)."""
    x, w = np.polynomial.legendre.leggauss(order)
    e = np.linspace(a, b, n + 1)
    h = 0.5 * (e[1:] - e[:-1])
    m = 0.5 * (e[1:] + e[:-1])
    return (m[:, None] + h[:, None] * x).ravel(), (h[:, None] * w).ravel()


def sqrt_right(z):
    s = np.sqrt(z)
    return np.where(s.real < 0, -s, s)


def bessel_sum(rho, C, kr, real_axis):
    """(n_rho, ncol) = J0(rho kr) @ C, chunked (greens.py:157-166)."""
    out = np.empty((rho.size, C.shape[1]), complex)
    step = max(1, int(4e6 // kr.size))
    for s in range(0, rho.size, step):
        arg = np.outer(rho[s:s + step], kr)
        out[s:s + step] = (j0(arg.real) if real_axis else jv(0, arg)) @ C
    return out


def table_nodes(lam, h, rmax):
    r = [0.0]
    while r[-1] < rmax:
        r.append(r[-1] + min(lam / 80, max(h / 25, r[-1] / 25)))
    return np.array(r)


class _PairRef:
    def __init__(self, dz, cA, cV, cW, spl, splw):
        self.dz, self.cA, self.cV, self.cW = dz, cA, cV, cW
        self._spl, self._splw = spl, splw
        self.zero = False

    def smooth(self, rho):
        S = self._spl(rho)
        return S[..., 0], S[..., 1]

    def smooth_w(self, rho):
        return self._splw(rho)

    def _R(self, rho):
        return np.sqrt(np.asarray(rho) ** 2 + self.dz ** 2)

    def gA(self, rho):
        return self.cA / (4 * np.pi * self._R(rho)) + self.smooth(rho)[0]

    def gV(self, rho):
        return self.cV / (4 * np.pi * self._R(rho)) + self.smooth(rho)[1]

    def gW(self, rho):
        return self.cW / (4 * np.pi * self._R(rho)) + self.smooth_w(rho)


class SlabGreens:
    """Grounded slab, conductors on top (greens.py:68-228)."""

    def __init__(self, freq, eps_r, h, tand=0.0, rho_max=0.1, gl_order=12, zs_ground=0.0):
        self.freq, self.h, self.zs_ground = freq, h, zs_ground
        self.omega = 2 * np.pi * freq
        self.k0 = k0 = self.omega / C0
        self.epsc = ec = eps_r * (1 - 1j * tand)
        self.k1 = k1 = k0 * np.sqrt(ec)
        self.ka = np.sqrt(0.5 * (k0 ** 2 + k1 ** 2))
        self.cv = 2.0 / (ec + 1.0)
        self.kv = np.sqrt((ec * k0 ** 2 + k1 ** 2) / (ec + 1.0))
        self.rho_max = rho_max
        # path
        k1r = abs(k1)
        T = 1.3 * k1r + k0
        a = min(0.5 * k0, 1.0 / rho_max)
        n1 = int(np.clip(np.ceil(2 * T / a), 100, 3000))
        t, w = panels(0.0, T, n1, gl_order)
        self.kc = t + 1j * a * np.sin(np.pi * t / T)
        self.wc = w * (1 + 1j * a * np.pi / T * np.cos(np.pi * t / T))
        kmax = max(40 * k1r, 20.0 / h, T * 1.5)
        dk = min(np.pi / rho_max, 0.5 / h, k1r)
        self.kr, wr = panels(T, kmax, int(np.ceil((kmax - T) / dk)), 8)
        self.wr = wr.astype(complex)
        # table
        rho = table_nodes(2 * np.pi / k1r, h, rho_max)
        RA, RV, RW = self.residual(rho)
        with np.errstate(divide="ignore", invalid="ignore"):
            fa = np.where(rho > 0, np.expm1(-1j * self.ka * rho) / (4 * np.pi * rho), -1j * self.ka / (4 * np.pi))
            fv = np.where(rho > 0, np.expm1(-1j * self.kv * rho) / (4 * np.pi * rho), -1j * self.kv / (4 * np.pi))
        gl = 0.0
        if zs_ground:
            gl = zs_ground / (1j * self.omega * MU0) * 2 * h / (2 * np.pi * (rho ** 2 + (2 * h) ** 2) ** 1.5)
        self.rho_tab = rho
        self.SA_tab = fa + RA + gl
        self.SV_tab = self.cv * fv + RV
        self.SW_tab = self.cv * fv + RW
        self._pair = _PairRef(0.0, 1.0, self.cv, self.cv, CubicSpline(rho, np.stack([self.SA_tab, self.SV_tab], -1)),
                              CubicSpline(rho, self.SW_tab))
        self.cA, self.cV, self.cW, self.dz = 1.0, self.cv, self.cv, 0.0
        self.levels = ["top"]

    def spectra(self, kr):
        k0, k1, h, ec = self.k0, self.k1, self.h, self.epsc
        u0 = np.sqrt(kr ** 2 - k0 ** 2)
        u1 = sqrt_right(kr ** 2 - k1 ** 2)
        x = u1 * h
        e = np.exp(-2 * x)
        small = np.abs(x) < 1e-4
        with np.errstate(divide="ignore", invalid="ignore"):
            coth = np.where(small, 1 / h + u1 ** 2 * h / 3, u1 * (1 + e) / (1 - e))
            tanh = np.where(small, u1 ** 2 * h, u1 * (1 - e) / (1 + e))
            tou = np.where(small, h, tanh / u1 ** 2)
        dte, dtm = u0 + coth, ec * u0 + tanh
        return 1.0 / dte, (u0 + tanh) / (dte * dtm), u0 * tou / dtm

    def remainders(self, kr):
        gA, gV, gW = self.spectra(kr)
        ua, uv = np.sqrt(kr ** 2 - self.ka ** 2), np.sqrt(kr ** 2 - self.kv ** 2)
        return gA - 0.5 / ua, gV - 0.5 * self.cv / uv, gW - 0.5 * self.cv / uv

    def residual(self, rho):
        rho = np.atleast_1d(np.asarray(rho, float))
        Cc = np.stack([self.wc * self.kc * f / (2 * np.pi) for f in self.remainders(self.kc)], 1)
        Cr = np.stack([self.wr * self.kr * f / (2 * np.pi) for f in self.remainders(self.kr)], 1)
        out = bessel_sum(rho, Cc, self.kc, False) + bessel_sum(rho, Cr, self.kr, True)
        return tuple(out.T)

    def pair(self, i=0, j=0):
        return self._pair

    def smooth(self, R):
        return self._pair.smooth(R)

    def gA(self, R):
        return self._pair.gA(R)

    def gV(self, R):
        return self._pair.gV(R)

    def gW(self, R):
        return self._pair.gW(R)


class LayeredGreensRef:
    """Kernels between all metal levels of a stack without inner planes
    (layered.py:212-415); `stack` provides layers, metals, top, z."""

    def __init__(self, freq, stack, rho_max=0.1, gl_order=12, zs_ground=0.0, levels=None):
        self.freq, self.stack, self.zs_ground = freq, stack, zs_ground
        self.omega = 2 * np.pi * freq
        self.k0 = self.omega / C0
        self.levels = list(levels) if levels is not None else list(stack.metals)
        self.zl = np.array([stack.level_z(n) for n in self.levels])
        self.il = np.array([stack.metals[n] for n in self.levels])
        self.epsc = np.array([d.eps_r * (1 - 1j * d.tand) for d in stack.layers])
        self.kk = self.k0 * np.sqrt(self.epsc)
        self.t = np.array([d.t for d in stack.layers])
        self.h = float(stack.z[-1])
        self.rho_max = rho_max
        kmaxr = np.abs(self.kk).max()
        T = 1.3 * kmaxr + self.k0
        a = min(0.5 * self.k0, 1.0 / rho_max)
        n1 = int(np.clip(np.ceil(2 * T / a), 100, 3000))
        t, w = panels(0.0, T, n1, gl_order)
        self.kc = t + 1j * a * np.sin(np.pi * t / T)
        self.wc = w * (1 + 1j * a * np.pi / T * np.cos(np.pi * t / T))
        tmin = self.t.min()
        kmax = max(40 * kmaxr, 20.0 / tmin, T * 1.5)
        dk = min(np.pi / rho_max, 0.5 / tmin, kmaxr)
        self.kr, wr = panels(T, kmax, int(np.ceil((kmax - T) / dk)), 8)
        self.wr = wr.astype(complex)
        self._leading()
        self._tabulate()

    def _tl(self, kr):
        N, K, L = len(self.t), kr.size, len(self.levels)
        u = np.stack([sqrt_right(kr ** 2 - k ** 2) for k in self.kk])
        u0 = np.sqrt(kr ** 2 - self.k0 ** 2)
        x = u * self.t[:, None]
        e = np.exp(-2 * x)
        th = (1 - e) / (1 + e)
        sc = 2 * np.exp(-x) / (1 + e)
        pec = self.stack.top == "pec"
        res = []
        for y, ya in ((u, u0), (self.epsc[:, None] / u, 1.0 / u0)):
            dn = np.empty((N + 1, K), complex)
            dn[1] = y[0] / th[0]
            for k in range(2, N + 1):
                dn[k] = y[k - 1] * (dn[k - 1] + y[k - 1] * th[k - 1]) / (y[k - 1] + dn[k - 1] * th[k - 1])
            up = np.empty((N + 1, K), complex)
            if not pec:
                up[N] = ya
            for k in range(N - 1, 0, -1):
                if k == N - 1 and pec:
                    up[k] = y[k] / th[k]
                else:
                    up[k] = y[k] * (up[k + 1] + y[k] * th[k]) / (y[k] + up[k + 1] * th[k])
            V = np.zeros((L, N + 1, K), complex)
            for l, s in enumerate(self.il):
                V[l, s] = 1.0 / (dn[s] + up[s])
                for k in range(s, 1, -1):
                    V[l, k - 1] = V[l, k] * sc[k - 1] / (1 + dn[k - 1] / y[k - 1] * th[k - 1])
                for k in range(s, N):
                    if k + 1 == N and pec:
                        break
                    V[l, k + 1] = V[l, k] * sc[k] / (1 + up[k + 1] / y[k] * th[k])
            res.append(V)
        return res[0], res[1], u

    def spectra(self, kr):
        kr = np.asarray(kr, complex)
        Vh, Ve, u = self._tl(kr)
        il = self.il
        gA = Vh[:, il].transpose(1, 0, 2)
        gV = ((Ve[:, il] + self.k0 ** 2 * Vh[:, il]) / kr ** 2).transpose(1, 0, 2)
        gW = np.cumsum(np.diff(Ve, axis=1) / u[None] ** 2, axis=1)[:, il - 1].transpose(1, 0, 2)
        return gA, gV, gW

    def _leading(self):
        L, N = len(self.levels), len(self.t)
        self.dz = np.abs(self.zl[:, None] - self.zl[None, :])
        self.c = np.zeros((3, L, L), complex)
        self.kx = np.zeros((2, L, L), complex)
        Kbig = 15.0 / self.t.min()
        big = None
        for i in range(L):
            for j in range(L):
                if i == j:
                    n = self.il[i]
                    eb, kb = self.epsc[n - 1], self.kk[n - 1]
                    ea, ka = (1.0, self.k0) if n == N else (self.epsc[n], self.kk[n])
                    cv = 2.0 / (ea + eb)
                    self.c[:, i, j] = (1.0, cv, cv)
                    self.kx[0, i, j] = np.sqrt(0.5 * (ka ** 2 + kb ** 2))
                    self.kx[1, i, j] = np.sqrt((eb * ka ** 2 + ea * kb ** 2) / (ea + eb))
                    continue
                if big is None:
                    big = [g[..., 0] for g in self.spectra(np.array([Kbig]))]
                ux = np.sqrt(Kbig ** 2 - self.k0 ** 2 + 0j)
                f = 2 * ux * np.exp(ux * self.dz[i, j])
                cA, cV, cW = (f * g[i, j] for g in big)
                if self.zl[j] < self.zl[i]:
                    cW = 0.0
                self.c[:, i, j] = (cA, cV, cW)
                self.kx[:, i, j] = self.k0
        self.c = np.where(np.abs(self.c.imag) < 1e-9 * np.abs(self.c), self.c.real, self.c)

    def _remainder(self, kr):
        out = []
        for m, g in enumerate(self.spectra(kr)):
            ux = np.sqrt(kr[None, None] ** 2 - self.kx[min(m, 1)][..., None] ** 2)
            out.append(g - self.c[m][..., None] * np.exp(-ux * self.dz[..., None]) / (2 * ux))
        return out

    def _tabulate(self):
        L = len(self.levels)
        rho = table_nodes(2 * np.pi / np.abs(self.kk).max(), self.t.min(), self.rho_max)
        res = 0.0
        for kr, w, real in ((self.kc, self.wc, False), (self.kr, self.wr, True)):
            rem = np.stack(self._remainder(kr.astype(complex)))
            C = (w * kr / (2 * np.pi))[None] * rem.reshape(-1, kr.size)
            res = res + bessel_sum(rho, C.T, kr.real if real else kr, real)
        RES = res.T.reshape(3, L, L, rho.size)
        self.rho_tab = rho
        self._pairs = {}
        for i in range(L):
            for j in range(L):
                R = np.sqrt(rho ** 2 + self.dz[i, j] ** 2)
                S = []
                for m in range(3):
                    k = self.kx[min(m, 1), i, j]
                    with np.errstate(divide="ignore", invalid="ignore"):
                        f = np.where(R > 0, np.expm1(-1j * k * R) / (4 * np.pi * R), -1j * k / (4 * np.pi))
                    S.append(self.c[m, i, j] * f + RES[m, i, j])
                if self.zs_ground:
                    d = self.zl[i] + self.zl[j]
                    S[0] = S[0] + self.zs_ground / (1j * self.omega * MU0) * d / (2 * np.pi * (rho ** 2 + d ** 2) ** 1.5)
                self._pairs[i, j] = _PairRef(self.dz[i, j], *self.c[:, i, j],
                                             CubicSpline(rho, np.stack([S[0], S[1]], -1)), CubicSpline(rho, S[2]))

    def pair(self, i, j):
        i = i if isinstance(i, (int, np.integer)) else self.levels.index(i)
        j = j if isinstance(j, (int, np.integer)) else self.levels.index(j)
        return self._pairs[i, j]

    def via_sections(self):
        return self.stack.via_sections(self.freq)


class _ZeroPairRef:
    """No kernels across a solid plane (layered.py:418-434)."""

    zero = True

    def __init__(self, dz):
        self.dz, self.cA, self.cV, self.cW = dz, 0.0, 0.0, 0.0

    def smooth(self, rho):
        z = np.zeros(np.shape(rho), complex)
        return z, z

    def smooth_w(self, rho):
        return np.zeros(np.shape(rho), complex)

    gA = gV = gW = smooth_w


class SectionedGreensRef:
    """Kernels of a stack with inner planes (layered.py:437-474): one
    LayeredGreensRef per section over its own section stack (levels measured
    from the plane below), none between sections."""

    def __init__(self, freq, stack, rho_max=0.1, gl_order=12, zs_ground=0.0, levels=None):
        self.freq, self.stack = freq, stack
        self.omega = 2 * np.pi * freq
        self.k0 = self.omega / C0
        self.levels = list(levels) if levels is not None else list(stack.metals)
        self.zl = np.array([stack.level_z(n) for n in self.levels])
        self.h = float(stack.z[-1])
        self.sec = [stack.section(n) for n in self.levels]
        self._sub = {}
        for q in sorted(set(self.sec)):
            names = [n for n, t in zip(self.levels, self.sec) if t == q]
            self._sub[q] = LayeredGreensRef(freq, stack.section_stack(q), rho_max, gl_order, zs_ground, names)

    def pair(self, i, j):
        i = i if isinstance(i, (int, np.integer)) else self.levels.index(i)
        j = j if isinstance(j, (int, np.integer)) else self.levels.index(j)
        if self.sec[i] != self.sec[j]:
            return _ZeroPairRef(abs(self.zl[i] - self.zl[j]))
        return self._sub[self.sec[i]].pair(self.levels[i], self.levels[j])

    def via_sections(self):
        return self.stack.via_sections(self.freq)
