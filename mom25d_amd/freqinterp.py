"""Frequency interpolation of the impedance matrix on the GPU
(Engine/mom25d/freqinterp.py:27-170).

Z is filled exactly at evenly spaced anchor frequencies and kept on the
device as W = j w Z exp(+j k_e R) (R the distance of the basis-function
centres, 0 on the via-via block); at any other frequency W is
Lagrange-interpolated from the order + 1 nearest anchors and
Z = W exp(-j k_e R) / (j w).  Both steps are one pass over the matrices
(`mom_interp_anchor`, `mom_interp_combine`).  The window and weight rules and
the defaults of step, order and k_e are the reference's, so a sweep uses the
same anchors and the same interpolation formula.
"""

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .constants import C0


class FreqInterp:
    """Interpolated impedance matrices for Solver.sweep(..., interp=FreqInterp(...)).

    step     : anchor spacing [Hz] (anchors span the requested band evenly);
               None: 0.35 c / (k_ratio D), D the diagonal of the box of the
               basis-function centres (freqinterp.py:60-64)
    order    : interpolation order (3 = cubic, 4 anchors per point)
    k_ratio  : phase-extraction wavenumber k_e = k_ratio * k0; None: sqrt of
               the quasi-static eps_eff under the first port
    check    : fill the first frequency that lies between anchors (at least a
               quarter step from both) exactly as well, and print the relative
               interpolation error of Z
    block    : kept for the reference's signature (the GPU kernels need no
               row blocking)
    dtype    : anchor storage, numpy.complex128 or complex64; None:
               complex128 unless order + 4 matrices would take more than half
               of the free device memory
    """

    def __init__(self, step=None, order=3, k_ratio=None, check=False, block=1024, dtype=None):
        if not 1 <= order <= 7:
            raise ValueError("order must be between 1 and 7")
        self.step, self.order, self.k_ratio, self.check, self.block = step, order, k_ratio, check, block
        self.dtype = dtype
        self.solver = None

    # ----------------------------------------------------------- setup
    def setup(self, solver, freqs):
        self.solver = solver
        f = np.asarray(freqs, float)
        lo, hi = f.min(), f.max()
        self._W = {}
        self._ready = {}
        self._windows = {}
        self.n_fills = 0
        self.fill_time = 0.0
        self.pos = solver._basis_positions()
        self.nb = solver.mesh.n_basis
        self._pos = torch.from_numpy(np.ascontiguousarray(self.pos)).to(solver.device)
        if self.k_ratio is None:
            self.k_ratio = float(np.sqrt(eps_eff(solver)))
        if self.step is None:
            ext = np.ptp(self.pos, axis=0)
            self.step = 0.35 * C0 / (max(self.k_ratio, 1.0) * float(np.linalg.norm(ext)))
        n = max(self.order + 1, int(np.ceil((hi - lo) / self.step - 1e-9)) + 1)
        self.anchors = np.linspace(lo, hi, n) if hi > lo else lo + self.step * (np.arange(n) - n // 2)
        self._checked = not self.check
        N = len(self.pos)
        if self.dtype is None:
            need = (self.order + 1) * N * N * 16 + 3 * N * N * 16
            free, _ = torch.cuda.mem_get_info(solver.device)
            self.dtype = np.complex128 if need < 0.5 * free else np.complex64
        self._prec = 64 if np.dtype(self.dtype) == np.complex128 else 32
        return self

    def _window(self, f):
        """Indices of the order+1 anchors used at f: centred on the anchor
        interval holding f (freqinterp.py:102-109)."""
        n, k = len(self.anchors), self.order + 1
        j = int(np.clip(np.searchsorted(self.anchors, f, side="right") - 1, 0, n - 2))
        i0 = min(max(j - (k - 1) // 2, 0), n - k)
        return list(range(i0, i0 + k))

    def weights(self, f, idx):
        """Lagrange weights of the anchors idx at f (freqinterp.py:111-118)."""
        fa = self.anchors[idx]
        L = np.ones(len(idx))
        for a in range(len(idx)):
            for b in range(len(idx)):
                if a != b:
                    L[a] *= (f - fa[b]) / (fa[a] - fa[b])
        return L

    # ----------------------------------------------------------- device
    def _phase(self, f):
        w = 2 * np.pi * f
        return w, self.k_ratio * w / C0

    def _anchor(self, i, lane):
        """W of anchor i on the device, filled on the current stream if it is
        not kept yet; an event marks when it is complete."""
        if i in self._W:
            return self._W[i]
        s = self.solver
        fa = float(self.anchors[i])
        t = time.time()
        n = s.mesh.n_unknowns
        Z = s._buf("Z", (n, n), torch.complex128, lane)
        s._fill(s.greens(fa), Z, None, lane)
        tdt = torch.complex128 if self._prec == 64 else torch.complex64
        W = torch.empty((n, n), dtype=tdt, device=s.device)
        w, ke = self._phase(fa)
        _lib.call("mom_interp_anchor", self._prec, n, self.nb, _lib.ptr(Z), Z.stride(0), _lib.ptr(self._pos), w, ke,
                  _lib.ptr(W), W.stride(0), _lib.stream_ptr())
        ev = torch.cuda.Event()
        ev.record()
        self._W[i], self._ready[i] = W, ev
        self.n_fills += 1
        self.fill_time += time.time() - t
        return W

    def _evict(self, keep):
        """Drop the anchors no lane's current window needs (each lane walks
        its frequencies in one direction, so they are not needed again)."""
        for i in list(self._W):
            if i not in keep:
                del self._W[i], self._ready[i]

    def fill(self, f, Z, lane=0):
        """Z (device complex128, N x N) <- the interpolated impedance matrix
        at f, enqueued on the current stream."""
        idx = self._window(f)
        self._windows[lane] = idx
        self._evict(set().union(*self._windows.values()))
        stream = torch.cuda.current_stream()
        Ws = []
        for i in idx:
            W = self._anchor(i, lane)
            stream.wait_event(self._ready[i])
            W.record_stream(stream)  # another lane may drop it while this stream still reads it
            Ws.append(W)
        L = self.weights(f, idx)
        w, ke = self._phase(f)
        n = Z.shape[0]
        ptrs = (ctypes.c_void_p * len(Ws))(*[W.data_ptr() for W in Ws])
        Lc = (ctypes.c_double * len(L))(*L)
        _lib.call("mom_interp_combine", self._prec, n, self.nb, len(Ws), ptrs, Lc, Ws[0].stride(0),
                  _lib.ptr(self._pos), w, ke, _lib.ptr(Z), Z.stride(0), _lib.stream_ptr())
        if self.check and not self._checked and np.min(np.abs(self.anchors - f)) > 0.25 * self.step:
            self._checked = True
            self.report_error(f, Z)
        return Z

    # ---------------------------------------------------------- public
    def impedance(self, f):
        """Interpolated Z at f (numpy complex128)."""
        n = self.solver.mesh.n_unknowns
        Z = torch.empty((n, n), dtype=torch.complex128, device=self.solver.device)
        return self.fill(f, Z).cpu().numpy()

    def report_error(self, f, Zi=None):
        """Relative Frobenius error of the interpolated Z at f against an
        exact fill, per block (RWG, RWG-via, via-via), the largest returned
        (the check of Engine/mom25d/freqinterp.py:141-162).  Both matrices
        stay on the device; Zi: an interpolated matrix already at hand."""
        s = self.solver
        n = s.mesh.n_unknowns
        dev = s.device
        Zint = torch.empty((n, n), dtype=torch.complex128, device=dev)
        if Zi is None:
            self.fill(f, Zint)
        else:
            Zint.copy_(torch.as_tensor(Zi))
        Zex = torch.empty((n, n), dtype=torch.complex128, device=dev)
        s._fill(s.greens(f), Zex, None, lane="interp-check")  # buffers of its own: lanes may be mid-solve
        nb = self.nb
        errs = {}
        for name, rows, cols in (("RWG", slice(0, nb), slice(0, nb)), ("RWG-via", slice(0, nb), slice(nb, n)),
                                 ("via-via", slice(nb, n), slice(nb, n))):
            ref = float(torch.linalg.vector_norm(Zex[rows, cols]))
            if ref > 0:
                errs[name] = float(torch.linalg.vector_norm(Zint[rows, cols] - Zex[rows, cols])) / ref
        del Zint, Zex
        self.last_error = max(errs.values())
        print(f"  FreqInterp check at {f / 1e9:.4f} GHz, Z error per block: "
              + ", ".join(f"{k} {v:.1e}" for k, v in errs.items())
              + f" (anchor spacing {self.step / 1e9:.3g} GHz, order {self.order})", flush=True)
        return self.last_error


def eps_eff(solver):
    """Quasi-static effective permittivity that sets the phase reference
    k_e = sqrt(eps_eff) k0 of the interpolation
    (Engine/mom25d/freqinterp.py:199-209): at the metal level of the first
    feed (or of the first gap port, or the first mesh level), the mean of
    the dielectric below and the air above for an uncovered level, the
    dielectric's own value under a cover."""
    levels = [p.layer for p in solver.ports]
    if not levels:
        gap_level = {q.name: q.layer for q in solver.lumped}
        levels = [gap_level[ip["name"]] for ip in solver.iports if ip["kind"] == "gap"]
    level = levels[0] if levels else solver.mesh.layer_names[0]
    below = solver.stack.eps_below(level)
    return below if solver.stack.covered(level) else 0.5 * (below + 1.0)
