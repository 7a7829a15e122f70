"""MPIE / RWG method-of-moments solver on one MI355X: the reference's Python
API (Engine/mom25d/solver.py:769-1335) over the gfx950 kernels of libmom25d.

Per frequency of a sweep:
  tables (all frequencies at once, greens.SweepTables)
  -> FP64 fill of Z on the GPU (tiers near / mid / far, atomics), Gram loss
     term, via rows / columns, symmetrisation Z <- (Z + Z^T) / 2
  -> LU with partial pivoting on the GPU: complex128 ("fp64"), or complex64
     on the matrix cores with FP64 iterative refinement against the FP64 Z
     ("mixed")
  -> triangle charges, line voltages at the port sample points (GPU)
  -> shared-gamma wave fit per port group and S (host, FP64).
"""

import ctypes
import math
import multiprocessing
import os
import sys
import time
import types
import warnings
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, analytic
from .constants import C0, EPS0, MU0
from .deembed import VOLTAGE_WEIGHT, FitContext, fit_waves, reciprocal_z0, sparams_point
from .greens import SectionedTables, SweepTables
from .plan import FillPlan, MomFillCall, check_struct_sizes


@dataclass
class Substrate:
    """Grounded dielectric slab with conductors on top (Engine/mom25d/solver.py:42-76)."""

    eps_r: float
    h: float
    tand: float = 0.0
    sigma: float = None
    t: float = 35e-6
    ground_loss: bool = True

    @property
    def lossless(self):
        return not self.tand and not self.sigma

    def surface_impedance(self, freq):
        if not self.sigma:
            return 0.0
        w = 2 * np.pi * freq
        kc = np.sqrt(1j * w * MU0 * self.sigma)
        return kc / self.sigma / np.tanh(kc * self.t)

    def to_stackup(self):
        from .layered import Stackup

        return Stackup.microstrip(self.eps_r, self.h, self.tand, self.sigma, self.t, self.ground_loss)


# --------------------------------------------------------------------- vias
def _grover(a1, a2, b1, b2, d):
    """Partial mutual inductance [H] of two parallel z-directed filaments
    occupying [a1, a2] and [b1, b2] at lateral distance d: the double
    integral of mu0 / (4 pi R) over both lengths, whose antiderivative in the
    offset u = z - z' is F(u) = u asinh(u / d) - sqrt(u^2 + d^2) (Grover);
    broadcasts over array arguments."""
    def F(u):
        return u * np.arcsinh(u / d) - np.hypot(u, d)

    return MU0 / (4 * np.pi) * (F(a2 - b1) + F(a1 - b2) - F(a2 - b2) - F(a1 - b1))


def filament_mutual(l, d):
    """Partial mutual inductance of two parallel, aligned filaments of length
    l at distance d (d = the radius: the partial self inductance of a tube)."""
    return _grover(0.0, l, 0.0, l, d)


def filament_mutual_z(a1, a2, b1, b2, d):
    """Partial mutual inductance of parallel z-filaments spanning [a1, a2] and
    [b1, b2] at lateral distance d."""
    return _grover(a1, a2, b1, b2, d)


def _via_geometry(vias):
    """Lateral distances between via axes, with each via's radius on the
    diagonal and the smaller radius for coaxial pairs."""
    xy = np.array([v.xy for v in vias], float).reshape(-1, 2)
    rad = np.array([0.5 * v.drill for v in vias], float)
    dist = np.hypot(*(xy[:, None, :] - xy[None, :, :]).transpose(2, 0, 1))
    dist = np.where(dist < 1e-12, np.minimum.outer(rad, rad), dist)
    dist[np.diag_indices_from(dist)] = rad
    return dist


def _with_image(a1, a2, b1, b2, d):
    """Filament [a1, a2] against filament [b1, b2] plus the image of the
    latter in the ground plane z = 0 (the span [-b2, -b1], same current
    direction for a vertical current)."""
    return _grover(a1, a2, b1, b2, d) + _grover(a1, a2, -b2, -b1, d)


def via_inductance(vias, h, spans=None):
    """Quasi-static partial inductance matrix [H] of vertical via currents
    with their ground-plane images (the k = 0 case of
    Engine/mom25d/solver.py:93-122).  Without spans every via runs from the
    plane to height h: the via and its image form one filament of 2 h, of
    which each half holds half the flux."""
    d = _via_geometry(vias)
    if spans is None:
        return 0.5 * _grover(0.0, 2 * h, 0.0, 2 * h, d)
    z = np.asarray(spans, float).reshape(-1, 2)
    return _with_image(z[:, None, 0], z[:, None, 1], z[None, :, 0], z[None, :, 1], d)


def _parallel_plate_z(a1, a2, r, height, k, omega):
    """Impedance between z-directed via pieces [a1_i, a2_i] inside a
    parallel-plate section of the given height (PEC on both sides): the TEM
    mode, omega mu0 / (4 height) l_i l_j H0^(2)(k r), plus the evanescent
    TM_n modes, j omega mu0 / (pi height) K0(q_n r) c_ni c_nj with
    k_n = n pi / height, q_n = sqrt(k_n^2 - k^2) and c_n = int cos(k_n z) dz
    over the piece; the sum runs until k_n r_min passes 45 (K0 < 1e-20)."""
    from scipy.special import hankel2, kv

    length = a2 - a1
    Z = omega * MU0 / (4 * height) * np.outer(length, length) * hankel2(0, k * r)
    n_modes = int(np.ceil(45 * height / (np.pi * r.min()))) + 1
    kn = np.pi / height * np.arange(1, n_modes + 1)
    c = (np.sin(np.outer(a2, kn)) - np.sin(np.outer(a1, kn))) / kn  # (pieces, modes)
    if np.abs(c).max() > 1e-12 * height:
        qn = np.sqrt(kn ** 2 - k ** 2 + 0j)
        Z = Z + 1j * omega * MU0 / (np.pi * height) * np.einsum("ijn,in,jn->ij", kv(0, r[..., None] * qn), c, c)
    return Z


def via_impedance(vias, spans, omega, sections, cache=None):
    """Vector-potential impedance [ohm] between the via segments, summed over
    the sections of the stack (Engine/mom25d/solver.py:125-199): in a section
    open at the top, j omega times the partial inductance of the pieces with
    their ground images; in a section closed by PEC, the parallel-plate mode
    sum.  sections: [(z_bottom, z_top, k_tem or None)].  cache: a dict the
    caller keeps for one set of vias and spans, holding the
    frequency-independent inductances of the open sections."""
    dist = None if cache is None else cache.get("dist")
    if dist is None:
        dist = _via_geometry(vias)
        if cache is not None:
            cache["dist"] = dist
    spans = np.asarray(spans, float).reshape(-1, 2)
    Z = np.zeros((len(spans), len(spans)), complex)
    for zb, zt, k in sections:
        height = zt - zb
        lo = np.clip(spans[:, 0], zb, zt) - zb
        hi = np.clip(spans[:, 1], zb, zt) - zb
        inside = np.flatnonzero(hi - lo > 1e-12 * height)
        if inside.size == 0:
            continue
        sel = slice(None) if inside.size == len(spans) else np.ix_(inside, inside)
        a1, a2 = lo[inside], hi[inside]
        if k is None:
            key = (zb, zt)
            L = None if cache is None else cache.get(key)
            if L is None:
                L = _with_image(a1[:, None], a2[:, None], a1[None, :], a2[None, :], dist[sel])
                if cache is not None:
                    cache[key] = L
            Z[sel] += (1j * omega) * L
        else:
            Z[sel] += _parallel_plate_z(a1, a2, dist[sel], height, k, omega)
    return Z


# -------------------------------------------------------------------- ports
class PortModel:
    """Excitation and sample cuts of one feed (the delta-gap model of
    Engine/mom25d/solver.py:570-610).

    The gap column puts 1 V across the RWG edges of the cut at s_gap (each
    coefficient is the signed edge length).  The wave fit samples the line
    at the mesh's own cuts inside [margin, s_gap - margin] (margin, unless
    the port sets one, max(1.5 W, 4 h): clear of the gap's and the
    circuit's near fields): the current through each cut, and the scalar
    potential averaged over 8 points around it, at +-d along the line (d half
    the distance to the nearest other cut) and (-0.3, -0.1, 0.1, 0.3) W
    across it."""

    ACROSS = (-0.3, -0.1, 0.1, 0.3)

    def __init__(self, mesh, port, h):
        self.port = port
        margin = max(1.5 * port.width, 4 * h) if port.margin is None else port.margin
        all_cuts = np.asarray(mesh.cut_positions(port), float)
        window = (all_cuts >= margin - 1e-12) & (all_cuts <= port.s_gap - margin + 1e-12)
        s = all_cuts[window]
        if s.size < 4:
            raise ValueError(f"port {port.name}: feed line too short or mesh too coarse ({s.size} sample cuts in "
                             f"the fitting window [{margin * 1e3:.3g}, {(port.s_gap - margin) * 1e3:.3g}] mm)")
        self.s = s
        gap_idx, gap_sgn = mesh.cut_edges(port, port.s_gap)
        if gap_idx.size == 0:
            raise ValueError(f"port {port.name}: no mesh edges on the gap cut")
        self.excitation = np.zeros(mesh.n_unknowns)
        self.excitation[gap_idx] = gap_sgn * mesh.edge_len[gap_idx]
        # every sample cut's edges as one flat list: current(cut) = sum of w * x[edge] over its entries
        per_cut = [mesh.cut_edges(port, c) for c in s]
        self._edge = np.concatenate([i for i, _ in per_cut]).astype(np.int64)
        self._w = np.concatenate([sg * mesh.edge_len[i] for i, sg in per_cut])
        self._cut = np.repeat(np.arange(s.size), [len(i) for i, _ in per_cut])
        # voltage points: the two neighbours of each cut at half the gap to its nearest other cut
        pos = np.flatnonzero(window)
        gaps = np.diff(all_cuts)
        left = np.where(pos > 0, gaps[np.maximum(pos - 1, 0)], np.inf)
        right = np.where(pos < all_cuts.size - 1, gaps[np.minimum(pos, gaps.size - 1)], np.inf)
        left = np.where(left > 0, left, np.inf)
        right = np.where(right > 0, right, np.inf)
        half = 0.5 * np.minimum(left, right)
        along = np.stack([s - half, s + half], 1)
        self.vpts = port.point(along[:, :, None], np.asarray(self.ACROSS) * port.width).reshape(-1, 2)
        self.nper = 2 * len(self.ACROSS)

    def currents(self, x):
        """Total current toward the circuit through each sample cut, (Ns, Nexc)."""
        out = np.zeros((self.s.size,) + x.shape[1:], complex)
        np.add.at(out, self._cut, self._w.reshape((-1,) + (1,) * (x.ndim - 1)) * x[self._edge])
        return out

    def voltages(self, phi):
        """Mean potential of each cut's 8 points, (Ns, Nexc)."""
        return phi.reshape(self.s.size, self.nper, -1).mean(1)


class FitPool:
    """Worker processes for the wave fits of a sweep.  They are spawned (not
    forked: this process holds a GPU context) and import numpy, scipy and the
    de-embedding module only.  The caller's main module is hidden while a
    worker starts, so that a script without a `__main__` guard is not run
    again in the workers."""

    def __init__(self, workers):
        self.workers = workers
        self.ex = ProcessPoolExecutor(workers, mp_context=multiprocessing.get_context("spawn"))

    def submit(self, fn, *args):
        main = sys.modules.get("__main__")
        sys.modules["__main__"] = types.ModuleType("__main__")
        try:
            return self.ex.submit(fn, *args)
        finally:
            sys.modules["__main__"] = main


_POOL = None


def fit_pool(workers):
    """The process's FitPool (kept for its lifetime, replaced by a larger one
    when more workers are asked for); None for workers == 0."""
    global _POOL
    if workers <= 0:
        return None
    if _POOL is None or _POOL.workers < workers:
        if _POOL is not None:
            _POOL.ex.shutdown(wait=False)
        _POOL = FitPool(workers)
    return _POOL


@dataclass
class SweepResult:
    """Result of a sweep (Engine/mom25d/solver.py:673-765)."""

    freqs: np.ndarray
    S: np.ndarray
    z0: np.ndarray
    gamma: np.ndarray
    z0_line: np.ndarray
    fit_residual: np.ndarray
    port_names: list
    currents: list = None
    waves: np.ndarray = None
    timing: dict = None

    def matched_currents(self, fi, port):
        if self.currents is None:
            raise ValueError("run sweep(..., keep_currents=True)")
        e = np.zeros(self.waves.shape[1])
        e[port] = 1.0
        return self.currents[fi] @ np.linalg.solve(self.waves[fi], e)

    @property
    def eeff(self):
        k0 = 2 * np.pi * self.freqs / C0
        return (self.gamma.imag / k0[:, None]) ** 2

    def renormalize(self, zref=50.0):
        """S referenced to a common real impedance zref."""
        out = np.empty_like(self.S)
        eye = np.eye(self.S.shape[1])
        for f, S in enumerate(self.S):
            sq = np.diag(np.sqrt(self.z0[f]))
            Zm = sq @ (eye + S) @ np.linalg.inv(eye - S) @ sq
            out[f] = (Zm - zref * eye) @ np.linalg.inv(Zm + zref * eye)
        return out

    def db(self, i, j, S=None):
        S = self.S if S is None else S
        return 20 * np.log10(np.abs(S[:, i, j]))

    def write_touchstone(self, path, zref=50.0):
        S = self.S if zref is None else self.renormalize(zref)
        z = zref if zref is not None else float(np.mean(self.z0))
        n = S.shape[1]
        with open(path, "w") as fh:
            fh.write(f"! mom25d-mi355x 2.5D MoM. ports: {', '.join(self.port_names)}\n")
            fh.write(f"# HZ S RI R {z:g}\n")
            for f, Sf in zip(self.freqs, S):
                vals = Sf.T.ravel() if n == 2 else Sf.ravel()
                fh.write(f"{f:.9e} " + " ".join(f"{v.real:.9e} {v.imag:.9e}" for v in vals) + "\n")


def _check_abi():
    L = _lib.lib()
    check_struct_sizes()
    from .greens import PAIR_DTYPE

    if L.mom_pair_tab_size() != PAIR_DTYPE.itemsize:
        raise RuntimeError("libmom25d PairTab layout differs from the Python mirror")


def givens_column(h, cs, sn, g, j):
    """Bring column j of an upper-Hessenberg least-squares problem to upper
    triangular form in place (GMRES; one problem per right-hand side along the
    last axis): apply the earlier rotations G_i = [[c_i*, s_i*], [-s_i, c_i]]
    to h (length j + 2), make the rotation that zeroes h[j + 1], and apply it
    to the right-hand side g.  |g[j + 1]| is then the least-squares residual.
    Scalar Python complex arithmetic: the arrays are short (a few right-hand
    sides), where per-element numpy calls would cost more than the work."""
    for c in range(h.shape[1]):
        col = h[:, c].tolist()
        csc, snc = cs[:j, c].tolist(), sn[:j, c].tolist()
        for i in range(j):
            ci, si = csc[i], snc[i]
            a, b = col[i], col[i + 1]
            col[i] = ci.conjugate() * a + si.conjugate() * b
            col[i + 1] = -si * a + ci * b
        a, b = col[j], col[j + 1]
        den = math.sqrt(abs(a) ** 2 + abs(b) ** 2) or 1.0
        cj, sj = a / den, b / den
        col[j] = cj.conjugate() * a + sj.conjugate() * b
        col[j + 1] = 0j
        cs[j, c], sn[j, c] = cj, sj
        h[:, c] = col
        gj = complex(g[j, c])
        g[j + 1, c] = -sj * gj
        g[j, c] = cj.conjugate() * gj


@dataclass
class Precond:
    """A factorisation that preconditions the mixed path's GMRES: the LU in
    precision prec (32 or 64) of D Z D, its row permutation, and the solve
    workspace -- with sb > 0 that of the super-block solve (mom_sbsolve_*:
    inverses of the diagonal super-blocks and the scaled couplings), else
    that of the 64-row chain with inverted blocks (mom_getrs_inv_z)."""

    prec: int
    A: object
    perm: object
    sw: object
    sb: int = 0

    def solve(self, n, nrhs, V, out, stream):
        """out = M^-1 V (complex128 vectors), enqueued on stream."""
        P = _lib.ptr
        if self.sb:
            _lib.call("mom_sbsolve_z", self.prec, n, nrhs, self.sb, P(self.perm), P(V), P(out), P(self.sw), stream)
        else:
            _lib.call("mom_getrs_inv_z", self.prec, n, nrhs, P(self.A), self.A.stride(0), P(self.perm), P(V), P(out),
                      P(self.sw), stream)

    def arnoldi(self, n, nrhs, j, m, Z, dv, swk, wbs, Vb, Wb, Hd, Hh, stream):
        """One GMRES step (mom_gmres_arnoldi / _sb)."""
        P = _lib.ptr
        if self.sb:
            _lib.call("mom_gmres_arnoldi_sb", self.prec, n, nrhs, j, m, self.sb, P(self.perm), P(self.sw), P(Z),
                      Z.stride(0), P(dv), P(swk), wbs, P(Vb), P(Wb), P(Hd), P(Hh), stream)
        else:
            _lib.call("mom_gmres_arnoldi", self.prec, n, nrhs, j, m, P(self.A), self.A.stride(0), P(self.perm),
                      P(self.sw), P(Z), Z.stride(0), P(dv), P(swk), wbs, P(Vb), P(Wb), P(Hd), P(Hh), stream)


class Solver:
    """Solver(mesh, Substrate | Stackup).sweep(freqs) -> SweepResult.

    precision : "fp64" -- complex128 LU (the plain FP64 path);
                "mixed" -- complex64 LU on the matrix cores with FP64
                iterative refinement against the FP64 impedance matrix
                (GMRES preconditioned by the LU, restarted every
                gmres_restart steps, at most refine_maxit cycles, to a
                normwise backward error of refine_tol).
    reuse_lu  : mixed precision, in a sweep: a lane keeps its last complex64
                LU as the preconditioner of its next frequency points (the
                FP64 GMRES reaches the same tolerance with a few more steps)
                and factors afresh once the GMRES steps of the last reused
                point exceed the lane's average cost per point since that
                factorisation (costs in GMRES steps; one factorisation counts
                lu_equiv steps, by default a model of this implementation:
                LU time / step time at this N).  lu_stats counts
                factorisations, reuses, GMRES steps and complex128 LUs.
    lu64_steps: mixed precision, in a sweep: when the solve of a point with
                its own fresh complex64 LU has not converged after lu64_steps
                GMRES steps, the complex64 factors have lost too much (the
                low-frequency end of a sweep, where the charge term of the
                MPIE outweighs the current-loop term by many orders and
                complex64 keeps too few digits of the latter): that solve is
                abandoned, the point is factored in complex128 (which also
                preconditions its neighbours), and new anchors up to twice
                its frequency are complex128 LUs (GMRES then needs ~10-20
                steps per reused point instead of 50-120).  None: complex64
                anchors only.
    nopiv     : mixed precision: factor the preconditioner LUs (complex64
                and complex128 anchors) without row interchanges
                (mom_getrf_nopiv: no cross-workgroup pivot search per
                column, so the factorisation is bounded by its update GEMM).
                On the diagonally scaled MoM matrices the unpivoted factors
                precondition as well as pivoted ones (element growth ~2); a
                solve whose GMRES has not converged after NOPIV_CAP steps
                is factored again with pivoting.
    ldlt      : mixed precision, with nopiv: D Z D is complex symmetric
                (Galerkin testing with the basis itself), so factor it as
                L D L^T (mom_sytrf_nopiv: the rows of U come from the columns
                of L and the trailing updates skip the tiles above the
                diagonal, about half the update flops of the LU).  The
                factors are stored as L and U = D L^T, so every solve takes
                them unchanged.
    solve_sb  : mixed precision: rows of the super-blocks of the
                preconditioner's triangular solves (512: the chain of
                super-blocks with precomputed inverses and scaled couplings,
                mom_sbsolve_*; 0: the chain of 64-row blocks, mom_getrs_inv_z).
                Systems of fewer than 2 super-blocks use the 64-row chain.
    method    : "auto" / "dense" (the H-matrix solver is not part of this build).
    surrogate : a trained GreensSurrogate of this stack: the remainders of
                the Green's tables come from its networks, corrected and
                judged per frequency by the Sommerfeld integrals at a few
                nodes (relative tolerance surrogate_tol; the integrals are
                used throughout where it misses).
    autotune  : time the LU's panel widths and update tiles for this N and
                precision on the device (mom25d_amd.autotune, cached per size
                class) and use the fastest; panel then only sets the default.
    graphs    : in a sweep, replay each lane's factorisation (and the FP64
                solve) from a HIP graph captured once per lane (off by
                default: on this ROCm the replay of these multi-stream graphs
                is slower than issuing the calls).
    workers   : processes that run the wave fits of a sweep beside the GPU
                work (None: up to 8 for sweeps of 16 points or more; 0: none).
    lanes     : frequency points in flight during a sweep, each on its own
                stream with its own matrices (None: 4 up to N = 4000, else 3
                in mixed precision and 2 in FP64, as many as fit in a third of
                the free device memory).  The latency-bound
                steps of one point (LU panels, triangular solves, host
                de-embedding) then overlap the matrix-core work of another.
    """

    OPEN = 1e12
    HMATRIX_N = 7000

    def __init__(self, mesh, substrate, z0=None, verbose=True, method="auto", v_weight=None, z0_recip_min=None,
                 tand_floor=None, precision="fp64", device="cuda", panel=None, refine_tol=1e-15, refine_maxit=20,
                 lanes=None, surrogate=None, surrogate_tol=5e-5, workers=None, graphs=False, autotune=False,
                 gmres_restart=120, reuse_lu=True, lu_equiv=None, lu64_steps=20, nopiv=True, solve_sb=512,
                 ldlt=True):
        from .layered import Stackup

        if method not in ("auto", "dense"):
            raise ValueError("method must be 'auto' or 'dense' (no H-matrix solver in this build)")
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.mesh = mesh
        self.sub = substrate
        self.stack = substrate if isinstance(substrate, Stackup) else substrate.to_stackup()
        st = self.stack
        missing = [n for n in mesh.layer_names if n not in st.metals]
        if missing:
            raise ValueError(f"mesh layers {missing} are not metal levels of the stackup {list(st.metals)} "
                             "(declare inner planes with Layout.add_plane)")
        planes = list(getattr(mesh, "planes", []))
        missing = [n for n in planes if n not in st.planes]
        if missing:
            raise ValueError(f"layout planes {missing} are not planes of the stackup {list(st.planes)}")
        self.single = len(st.layers) == 1 and st.top == "air" and mesh.layer_names == ["top"]
        self.zl = np.array([st.level_z(n) for n in mesh.layer_names])
        # levels of different stack sections (a solid plane between them) do not couple
        self.section = [st.section(n) for n in mesh.layer_names] if st.planes else None
        order = sorted(mesh.layer_names + planes, key=st.level_z)
        if getattr(mesh, "level_order", None) != order:
            mesh.assign_levels(order)
        self.ports = mesh.layout.ports
        self.iports = list(getattr(mesh.layout, "internal_ports", []))
        if not self.ports and not self.iports:
            raise ValueError("layout has no ports")
        # lumped components: delta gaps loaded after the solve (Woodbury); internal ports: S-parameter
        # ports with a fixed reference impedance across a delta gap, via ports adding their own columns
        # after the lumped gaps (Engine/mom25d/solver.py:848-868)
        self.lumped = list(getattr(mesh.layout, "lumped", []))
        self.components = list(getattr(mesh.layout, "components", []))
        self.Lm = mesh.lumped_vectors() if self.lumped else None
        self.iport_index = []
        if self.iports:
            names = [q.name for q in self.lumped]
            V = mesh.internal_port_vectors()
            extra = [k for k, ip in enumerate(self.iports) if ip["kind"] == "via"]
            if extra:
                self.Lm = V[:, extra] if self.Lm is None else np.hstack([self.Lm, V[:, extra]])
            col = {k: len(names) + i for i, k in enumerate(extra)}
            self.iport_index = [names.index(ip["name"]) if ip["kind"] == "gap" else col[k]
                                for k, ip in enumerate(self.iports)]
        self.values = {}
        self.pm = [PortModel(mesh, p, st.level_h(p.layer)) for p in self.ports]
        self.method = "dense"
        self.precision = precision
        self.z0_fixed = None if z0 is None else np.asarray(z0, float)
        v = mesh.nodes
        self.rho_max = 1.02 * np.hypot(*(v.max(0) - v.min(0)))
        self.verbose = verbose
        self.v_weight = VOLTAGE_WEIGHT if v_weight is None else float(v_weight)
        self.z0_recip_min = z0_recip_min
        self.device = torch.device(device)
        self.panel = panel or 64
        self.tile = None
        self.block = None
        if autotune:
            from .autotune import tune_lu

            # the mixed path factors its preconditioners without row interchanges when nopiv is set
            nopiv_mixed = precision == "mixed" and nopiv
            best = tune_lu(mesh.n_unknowns, 64 if precision == "fp64" else 32, self.device,
                           pivot=not nopiv_mixed, symmetric=nopiv_mixed and ldlt)
            self.panel, self.block, self.tile = best["jb"], best["nb"], tuple(best["tile"])
        self.tuning = dict(jb=self.panel, nb=self.block, tile=self.tile)
        self.refine_tol = refine_tol
        self.refine_maxit = refine_maxit
        self.gmres_restart = int(gmres_restart)
        self.reuse_lu = bool(reuse_lu)
        self.lu_equiv = lu_equiv
        self._anchor = {}
        self._history = {}
        self._lane_pc = {}  # lane -> Precond of its last factorisation
        # mixed path, per solved point: frequency, lane and the (preconditioner, GMRES steps) of its solves
        self.point_log = []
        self.lu64_steps = lu64_steps
        self.nopiv = bool(nopiv)
        self.ldlt = bool(ldlt)
        self.solve_sb = int(solve_sb or 0)
        self._f_poor = 0.0
        self._c64_refs = []  # (frequency, GMRES steps) of converged fresh complex64 factorisations
        self._c64_pending = {}  # lane -> frequency of its complex64 factorisation in progress
        self._active_lanes = 1
        self.lu_stats = dict(factorisations=0, reused=0, fallbacks=0, gmres_steps=0, fp64_lu=0, pivoted_retries=0,
                             abandoned=0, predicted_poor=0)
        self.lanes = lanes
        self.workers = workers
        self.graphs = graphs
        self._graphs = {}
        self._graph_seen = set()
        self.surrogate = surrogate
        self.surrogate_tol = surrogate_tol
        self.last_tables = None
        self._frills = None  # antipad.FrillModel, built at the first fill with antipads
        self._plan = None
        self._bufs = {}
        self._keep = {}
        self._via_cache = {}
        _check_abi()
        self.groups = {}
        for p, port in enumerate(self.ports):
            hp = st.level_h(port.layer)
            key = (round(port.width / hp, 9), port.layer,
                   port.ground.key(hp) if port.ground is not None else None, port.z0_from)
            self.groups.setdefault(key, []).append(p)

    def __del__(self):
        for _sig, h in getattr(self, "_graphs", {}).values():
            try:
                _lib.call("mom_graph_destroy", h)
            except Exception:  # interpreter shutdown: the library may be gone
                pass

    @property
    def lossless(self):
        return self.stack.lossless

    def surface_impedance(self, freq):
        return self.stack.surface_impedance(freq)

    # ------------------------------------------------------------- set-up
    def geometry(self):
        if self._plan is None:
            self._plan = FillPlan(self.mesh, zl=self.zl, section=self.section)
        return self._plan

    def tables(self, freqs):
        """Green's-function tables of all frequencies (one GPU launch)."""
        if self.stack.planes:
            if self.surrogate is not None:
                raise ValueError("the remainder surrogate is trained for stacks without inner planes")
            t = SectionedTables(self.stack, freqs, self.rho_max, self.mesh.layer_names, self.device)
        else:
            t = SweepTables(self.stack, freqs, self.rho_max, self.mesh.layer_names, self.single, self.device,
                            surrogate=self.surrogate, surrogate_tol=self.surrogate_tol)
        self.last_tables = t
        return t

    def greens(self, freq):
        return self.tables([freq])[0]

    def _buf(self, name, shape, dtype, lane=0, zero=False):
        """Device buffer kept between frequencies (one set per lane); zero:
        a new buffer starts zeroed (the solve workspaces, whose status words
        the library reads before it writes them)."""
        key = (name, lane)
        b = self._bufs.get(key)
        if b is None or tuple(b.shape) != tuple(shape) or b.dtype != dtype:
            self._bufs.pop(key, None)
            b = (torch.zeros if zero else torch.empty)(shape, dtype=dtype, device=self.device)
            self._bufs[key] = b
        return b

    SOLVE_WORKSPACES = ("swork32", "swork64", "swork64i", "sbwork32", "sbwork64")

    def _clear_status(self, lane=0):
        """Zero the failure words of the lane's solve workspaces (on the
        current stream): the solves only ever set them, so a check after a
        frequency point sees a failure of any solve the point ran."""
        for k in self.SOLVE_WORKSPACES:
            b = self._bufs.get((k, lane))
            if b is not None:
                b[16:20].zero_()

    def _pinned(self, name, shape, dtype, lane=0):
        """Pinned host buffer kept between calls (one per lane)."""
        key = ("pinned", name, lane)
        b = self._bufs.get(key)
        if b is None or tuple(b.shape) != tuple(shape) or b.dtype != dtype:
            b = torch.empty(shape, dtype=dtype, pin_memory=True)
            self._bufs[key] = b
        return b

    def _upload(self, name, a, lane=0):
        """A host array in the lane's device buffer `name`, enqueued on the
        current stream without waiting for the stream's earlier work (as a
        copy from pageable memory would): staged through a pinned buffer of
        the lane, whose previous copy is waited for before it is overwritten
        (long done when the lane comes back to it)."""
        a = np.ascontiguousarray(a)
        dt = torch.from_numpy(a[:0]).dtype
        key = ("stage", name, lane)
        h, ev = self._bufs.get(key, (None, None))
        if h is None or h.numel() < a.nbytes:
            h, ev = torch.empty(max(a.nbytes, 1 << 12), dtype=torch.uint8, pin_memory=True), None
        if ev is not None:
            ev.synchronize()
        hv = h[:a.nbytes]
        hv.numpy()[:] = a.reshape(-1).view(np.uint8)
        d = self._buf(name, a.shape, dt, lane)
        d.view(torch.uint8).view(-1).copy_(hv, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._bufs[key] = (h, ev)
        return d

    PANEL_LDS = 152 * 1024  # LDS of one LU panel workgroup (the library's bound)
    # GMRES products from the tiles on and above the diagonal of Z (mom_scaled_symv):
    # valid because every Z the solver builds is symmetric (the fill ends with
    # Z <- (Z + Z^T) / 2 as Engine/mom25d/solver.py:522-524 does; the frequency
    # interpolation preserves it); False: the row-wise product over all of Z
    symmetric_products = True
    # a list: every fill tier's launch is bracketed by two events on its stream,
    # appended here as (tier, start, end, algorithmic flops) (the bench's roofline report)
    fill_events = None
    # spline intervals a fill work item may stage in LDS (those between its
    # smallest and largest source-test distance; an item that needs more
    # reads the table from global memory; 0: never stage).  The values are
    # the same either way.
    fill_stage = 64

    def _worker_count(self, nf):
        """Processes for the wave fits of a sweep (0: in this process)."""
        if self.workers is not None:
            return max(0, int(self.workers))
        return min(8, max(1, (os.cpu_count() or 2) // 2)) if nf >= 4 else 0

    def _lane_count(self):
        """Lanes of a sweep: as asked (default 4 in mixed precision and, up
        to N = 4000, in FP64, else 2; as many as fit in a third of the free
        device memory), and no more than allow every lane's LU panel to be
        resident at once."""
        n = self.mesh.n_unknowns
        if self.lanes is not None:
            lanes = max(1, int(self.lanes))
        else:
            per = n * n * (16 + (8 if self.precision == "mixed" else 0))
            free, _total = torch.cuda.mem_get_info(self.device)
            # mixed precision spends most of a sweep in latency-bound GMRES steps and factorisations
            # bound by their panels, which further lanes overlap (measured at N = 12 428, 20 points:
            # 0.98 / 0.94 / 0.86 / 0.88 / 0.92 s with 2 / 3 / 4 / 5 / 6 lanes)
            want = 4 if n <= 4000 or self.precision == "mixed" else 2
            lanes = max(1, min(want, int(free / 3 // per)))
        if self.precision == "mixed" and self.nopiv:
            # unpivoted factorisations wait for no other workgroup; a pivoted
            # one is only a fallback, whose panel narrows (down to 16 columns
            # of complex128) until its rows fit its share of the CUs
            elem, jb = 16, 16
        else:
            elem, jb = (16 if self.precision == "fp64" else 8), self.panel
        rows = max(1, self.PANEL_LDS // (jb * elem))
        groups = -(-n // rows)
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return max(1, min(lanes, min(cus, 256) // groups))

    @staticmethod
    def _run(gen):
        """Drive a solve generator to its end on the current stream."""
        try:
            while True:
                next(gen)
        except StopIteration as stop:
            return stop.value

    # --------------------------------------------------------------- fill
    def _via_extra(self, g, zs_via):
        """Host part of the via-via block: -Im gV(0) correction, the
        vector-potential impedance of the segments, the via-wall loss."""
        plan = self.geometry()
        omega = g.omega
        inv_jwe = 1 / (1j * omega * EPS0)
        nv = plan.nv
        # sum over the end levels of both vias of s_a s_b Im gV_ab(0):
        # S^T M S with S[level, via] the summed end signs
        levels = sorted({la for ek in plan.via_ends for la, _ in ek})
        row = {la: i for i, la in enumerate(levels)}
        S = np.zeros((len(levels), nv))
        for k, ek in enumerate(plan.via_ends):
            for la, sa in ek:
                S[row[la], k] += sa
        M = np.zeros((len(levels), len(levels)))
        for la in levels:
            for lb in levels:
                pr = g.pair(la, lb)
                rho = 0.0 if pr.dz else 1e-6 * g.h
                M[row[la], row[lb]] = pr.gV(np.array([rho]))[0].imag
        E = (-inv_jwe * 1j) * (S.T @ M @ S).astype(complex)
        st = self.stack
        segs = self.mesh.via_segments
        vz = {n: st.level_z(n) for n in self.mesh.layer_names + list(st.planes)}
        vz["gnd"] = 0.0
        spans = [(vz[s["bottom"]], vz[s["top"]]) for s in segs]
        E += via_impedance(plan.vias, spans, omega, g.via_sections(), self._via_cache)
        if zs_via:
            lengths = np.array([z1 - z0 for z0, z1 in spans])
            E += np.diag([zs_via * l / (np.pi * v.drill) for l, v in zip(lengths, plan.vias)])
        return E

    def _fill(self, g, Z, Z32=None, lane=0):
        """Z (device complex128, zeroed here) <- the symmetric impedance matrix
        at g.freq; Z32 (complex64) receives the diagonally scaled copy when
        given.  Enqueued on the current stream."""
        plan = self.geometry()
        d = plan.upload(self.device)
        n = self.mesh.n_unknowns
        omega = g.omega
        jwmu = 1j * omega * MU0
        ijwe = 1 / (1j * omega * EPS0)
        Z.zero_()
        nv = plan.nv
        ivpad = None
        if nv:
            ivpad = self._buf("ivpad", (plan.pad_tris.size, plan.nt), torch.complex128, lane)
            ivpad.zero_()
        P = _lib.ptr
        tabs = g.tables
        call = MomFillCall(jwmu_re=jwmu.real, jwmu_im=jwmu.imag, ijwe_re=ijwe.real, ijwe_im=ijwe.imag,
                           pairs=g.pairs_dev.data_ptr(), tab_x=tabs.tab_x.data_ptr(), tab_c=tabs.tab_c.data_ptr(),
                           tab_b=tabs.tab_b.data_ptr(), Z=Z.data_ptr(), ldz=Z.stride(0),
                           IVpad=ivpad.data_ptr() if ivpad is not None else None, stage_max=self.fill_stage)
        stream = _lib.stream_ptr()
        if self.fill_events is None:
            plan.fill(call, stream)
        else:  # one launch per tier, timed on the launch's stream for a roofline report
            for tier in (0, 1, 2):
                if not plan.items[tier].size:
                    continue
                ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                ev[0].record()
                plan.fill(call, stream, 1 << tier)
                ev[1].record()
                self.fill_events.append((tier, ev[0], ev[1], float(plan.FLOPS_PER_PAIR[tier] * plan.tier_pairs[tier])))
        zs = self.stack.surface_impedance(g.freq)
        if zs:
            _lib.call("mom_gram_add", plan.nt, P(d["gram"]), P(d["tb"]), float(np.real(zs)), float(np.imag(zs)),
                      P(Z), Z.stride(0), stream)
        if nv:
            E = self._upload("via_extra", self._via_extra(g, zs), lane)
            _lib.call("mom_via_blocks", plan.nb, nv, plan.nt, P(ivpad), P(d["d_off"]), P(d["d_pad"]), P(d["d_val"]),
                      P(d["pad_tri"]), P(d["tp"]), P(d["vp"]), P(d["tm"]), P(d["vm"]), P(d["cV"]), P(E),
                      ijwe.real, ijwe.imag, P(Z), Z.stride(0), stream)
            self._keep[lane] = E
        if self.mesh.n_frills:
            self._frill_blocks(g, Z, lane)
        if Z32 is not None:
            self._scale_copy(Z, Z32, lane)
        else:
            _lib.call("mom_symmetrize", n, P(Z), Z.stride(0), None, 0, None, stream)
        return Z

    def _frill_blocks(self, g, Z, lane=0):
        """The antipad rows and columns of Z (antipad.FrillModel, host):
        Z[:nj, nj:] = -C, Z[nj:, :nj] = -C^T, Z[nj:, nj:] = -(F + F^T) / 2
        with nj the RWG and via unknowns (Engine/mom25d/solver.py:946-961)."""
        from .antipad import FrillModel

        if self._frills is None:
            self._frills = FrillModel(self)
        C, F = self._frills.matrix_blocks(g.omega)
        nj = C.shape[0]
        blk = self._upload("frill", np.ascontiguousarray(np.vstack([C, F])), lane)
        Z[:nj, nj:].copy_(blk[:nj])
        Z[nj:, :nj].copy_(blk[:nj].T)
        Z[nj:, nj:].copy_(blk[nj:])

    def _scale_copy(self, Z, Z32, lane=0):
        """Z <- (Z + Z^T) / 2 and Z32 <- D Z D in complex64 with
        D = diag(|Z_ii|^-1/2) (kept in the lane's dscale buffer).  Idempotent
        on Z, so it also serves a Z that is already symmetric."""
        n = Z.shape[0]
        P = _lib.ptr
        stream = _lib.stream_ptr()
        dsc = self._buf("dscale", (n,), torch.float64, lane)
        _lib.call("mom_diag_scale", n, P(Z), Z.stride(0), P(dsc), stream)
        _lib.call("mom_symmetrize", n, P(Z), Z.stride(0), P(Z32), Z32.stride(0), P(dsc), stream)

    def impedance(self, freq, greens=None):
        """The impedance matrix at freq (numpy complex128)."""
        g = greens or self.greens(freq)
        n = self.mesh.n_unknowns
        Z = self._buf("Z", (n, n), torch.complex128)
        self._fill(g, Z)
        return Z.cpu().numpy()

    # -------------------------------------------------------------- solve
    def port_excitations(self):
        """One column per port: 1 V delta gap at each feed, then a 1 V source
        (in series with z_ref) at each internal port (Engine/mom25d/solver.py:1029-1034)."""
        cols = [pm.excitation for pm in self.pm]
        cols += [self.Lm[:, c] for c in self.iport_index]
        return np.stack(cols, 1)

    NOPIV_CAP = 200  # GMRES steps after which an unpivoted preconditioner is replaced by a pivoted one

    def _lu_solve(self, Z, B, Z32=None, lane=0, graph=False, pivot=False, cap=None):
        """Generator: X = Z^-1 B on the device (the value it returns).  Z
        (complex128) is overwritten by its LU in the fp64 path; in the mixed
        path Z32 holds the complex64 LU and Z stays intact for the refinement
        residuals.  It yields wherever the host would wait for the GPU.
        graph: replay the factorisation (and the FP64 solve) from a HIP graph
        of this lane's buffers, captured on the second call (the first runs
        the same calls directly); the current stream must not be the default
        stream.  The mixed path's complex64 LU has no row interchanges when
        nopiv is set (pivot=False); should its GMRES not converge within
        NOPIV_CAP steps, the point is factored again with pivoting.  cap: the
        mixed path returns None once its GMRES has taken that many steps
        without converging (the caller decides what follows)."""
        n, nrhs = B.shape
        stream = _lib.stream_ptr()
        P = _lib.ptr
        jb = self.panel
        prec = 64 if Z32 is None else 32
        A = Z if Z32 is None else Z32
        ipiv = self._buf("ipiv", (n,), torch.int32, lane)
        perm = self._buf("perm", (n,), torch.int32, lane)
        info = self._buf("info", (4,), torch.int32, lane)
        nopiv = Z32 is not None and self.nopiv and not pivot
        groups = self._panel_groups()
        elem = 16 if prec == 64 else 8
        # a pivoted panel's rows must fit the LDS of `groups` workgroups
        while not nopiv and groups and jb > 16 and -(-n // max(1, self.PANEL_LDS // (elem * jb))) > groups:
            jb //= 2
        wbytes = int(_lib.lib().mom_getrf_work_bytes(prec, n, jb))
        work = self._buf(f"work{prec}", (wbytes,), torch.uint8, lane)
        # the complex64 solves (GMRES preconditioner): the super-block solve's
        # inverses and scaled couplings, or the 64-row chain's inverted blocks
        # kept after the getrs workspace
        sb = self._sb_rows(n) if prec == 32 else 0
        if sb:
            swb = _lib.lib().mom_sbsolve_work_bytes(prec, n, sb)
            sw = self._buf(f"sbwork{prec}", (int(swb),), torch.uint8, lane, zero=True)
        else:
            swb = (_lib.lib().mom_getrs_work_bytes(prec, n) if prec == 64
                   else _lib.lib().mom_getrs_inv_work_bytes(prec, n))
            sw = self._buf(f"swork{prec}", (int(swb),), torch.uint8, lane, zero=True)
        X = self._buf("X", (n, nrhs), torch.complex128, lane)

        if self.tile is not None:
            _lib.call("mom_set_gemm_tile", prec, self.tile[0], self.tile[1])
        if self.block is not None:
            _lib.call("mom_set_lu_block", prec, self.block)

        def calls():
            if nopiv:
                _lib.call(self._nopiv_entry, prec, n, P(A), A.stride(0), P(ipiv), P(info), P(work), wbytes, jb,
                          stream)
            else:
                _lib.call("mom_getrf", prec, n, P(A), A.stride(0), P(ipiv), P(info), P(work), wbytes, jb, groups,
                          stream)
            _lib.call("mom_build_perm", n, P(ipiv), P(perm), stream)
            if Z32 is None:
                for c0, c1 in _column_chunks(nrhs):  # the solves take up to 4 right-hand sides
                    if nrhs <= 4:
                        _lib.call("mom_getrs", 64, n, nrhs, P(A), A.stride(0), P(perm), P(B), P(X), P(sw), stream)
                    else:
                        Bc = B[:, c0:c1].contiguous()
                        Xc = torch.empty_like(Bc)
                        _lib.call("mom_getrs", 64, n, c1 - c0, P(A), A.stride(0), P(perm), P(Bc), P(Xc), P(sw),
                                  stream)
                        X[:, c0:c1] = Xc
            elif sb:
                _lib.call("mom_sbsolve_setup", 32, n, P(A), A.stride(0), sb, P(sw), int(swb), stream)
            else:
                _lib.call("mom_trtri_blocks", 32, n, P(A), A.stride(0), P(sw), stream)

        if graph and nrhs <= 4:
            key = (lane, prec)
            sig = (stream, A.data_ptr(), B.data_ptr(), X.data_ptr(), work.data_ptr(), sw.data_ptr(), ipiv.data_ptr(),
                   groups, jb, n, nrhs, nopiv)
            g = self._graphs.get(key)
            if g is not None and g[0] != sig:
                _lib.call("mom_graph_destroy", g[1])
                g = None
            if g is None and key in self._graph_seen:
                h = ctypes.c_void_p()
                _lib.call("mom_graph_begin", stream)
                try:
                    calls()
                finally:
                    _lib.call("mom_graph_end", stream, ctypes.byref(h))
                g = self._graphs[key] = (sig, h.value)
            self._graph_seen.add(key)
            if g is not None:
                _lib.call("mom_graph_launch", g[1], stream)
            else:
                calls()
        else:
            yield from _on_launcher(calls, self.device)
        if Z32 is None:
            self.last_refine = 0
            return X
        pc = Precond(32, A, perm, sw, sb)
        self._lane_pc[lane] = pc  # per lane: other lanes factor while this one's GMRES yields
        if cap is not None:  # the caller handles a solve that has not converged within cap steps
            return (yield from self._gmres_chunks(Z, pc, B, lane, max_its=cap))
        X = yield from self._gmres_chunks(Z, pc, B, lane, max_its=self.NOPIV_CAP if nopiv else None)
        if X is None:  # the unpivoted factors precondition too poorly here: factor again with pivoting
            self.lu_stats["pivoted_retries"] += 1
            self._scale_copy(Z, Z32, lane)
            X = yield from self._lu_solve(Z, B, Z32, lane, graph, pivot=True)
        return X

    @property
    def _nopiv_entry(self):
        """The unpivoted factorisation of the preconditioners: L D L^T of the
        complex-symmetric D Z D, or the general LU."""
        return "mom_sytrf_nopiv" if self.ldlt else "mom_getrf_nopiv"

    def _sb_rows(self, n):
        """Super-block rows of the preconditioner solves at this size (0:
        the 64-row chain)."""
        return self.solve_sb if self.solve_sb and n >= 2 * self.solve_sb else 0

    def _gmres_chunks(self, Z, pc, B, lane=0, max_its=None, X0=None):
        """_gmres_ir over groups of at most 4 right-hand sides (the solve
        kernels' limit); None if a group does not converge within max_its.
        X0: an initial guess of the solution (or None)."""
        nrhs = B.shape[1]
        if nrhs <= 4:
            return (yield from self._gmres_ir(Z, pc, B, lane, max_its, X0))
        X = torch.empty_like(B)
        its = 0
        for c0, c1 in _column_chunks(nrhs):
            Xc = yield from self._gmres_ir(Z, pc, B[:, c0:c1].contiguous(), lane, max_its,
                                           None if X0 is None else X0[:, c0:c1].contiguous())
            its += self.last_refine
            if Xc is None:
                self.last_refine = its
                return None
            X[:, c0:c1] = Xc
        self.last_refine = its
        return X

    def _status_tensors(self, lane=0):
        """The device status words of the lane's factorisation and solves."""
        return [self._bufs[("info", lane)]] + [self._bufs[(k, lane)][:20] for k in self.SOLVE_WORKSPACES
                                               if (k, lane) in self._bufs]

    def _check_solve(self, lane=0, status=None):
        """Status words of the lane's last factorisation and solves (read
        after the solve's result has been copied to the host; status: their
        host copies, as _status_tensors lists them, when already read): raise
        if a bounded wait of the LU kernels ran out, warn on an exactly zero
        pivot (as LAPACK's zgetrf reports; the solution is then not finite)."""
        if status is None:
            status = [t.cpu().numpy() for t in self._status_tensors(lane)]
        info = status[0]
        waits = [w.view(np.uint32)[4] for w in status[1:]]
        if info[1] or any(waits):
            raise RuntimeError("LU kernels: a bounded wait ran out (workgroups did not all run); result invalid")
        if info[0]:
            warnings.warn(f"LU: diagonal number {int(info[0])} is exactly zero (singular matrix)", RuntimeWarning,
                          stacklevel=2)

    def _panel_groups(self):
        """Workgroup bound of the LU panel: every lane's panel resident at once."""
        lanes = getattr(self, "_active_lanes", 1)
        if lanes <= 1:
            return 0
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        return max(1, cus // lanes)

    def _gmres_ir(self, Z, pc, B, lane=0, max_its=None, X0=None):
        """Restarted GMRES in FP64 on the diagonally scaled system
        (D Z D) y = D b, x = D y, right-preconditioned by the complex64 LU of
        D Z D.  All right-hand sides advance together (one Krylov basis per
        column).  Target: normwise backward error of the scaled system
        |D b - D Z D y| <= refine_tol |D Z D|_inf |y| (FP64 round-off); a
        restart cycle that does not halve the true residual ends the
        iteration (the FP64 residual floor is reached).  With max_its it
        returns None once that many steps have not reached the target (the
        caller then factors this matrix itself).  X0: an initial guess of x;
        the iteration then starts from y0 + M (D b - D Z D y0), y0 = D^-1 X0,
        instead of M D b.  pc: the preconditioning factorisation (Precond:
        complex64, or complex128 for the anchors of the low-frequency end)."""
        n, nrhs = B.shape
        P = _lib.ptr
        stream = _lib.stream_ptr()
        dv = self._bufs[("dscale", lane)]
        d = dv[:, None]
        m = self.gmres_restart
        rows = self._buf("rowsum", (n,), torch.float64, lane)
        Bs = d * B

        def precond(V, out=None):  # complex64 LU solve of complex128 vectors
            out = torch.empty_like(V) if out is None else out
            pc.solve(n, nrhs, V, out, stream)
            return out

        if self.symmetric_products:  # Z = Z^T (the fill symmetrises it): half the HBM traffic
            wbs = int(_lib.lib().mom_scaled_symv_work_bytes(n, nrhs))
            swk = self._buf("symv_work", (wbs,), torch.uint8, lane)

        def matvec(Y):  # (D Z D) Y
            W = torch.empty_like(B)
            if self.symmetric_products:
                _lib.call("mom_scaled_symv", n, nrhs, P(Z), Z.stride(0), P(dv), P(Y), P(W), P(swk), wbs, stream)
            else:
                _lib.call("mom_scaled_matvec", n, nrhs, P(Z), Z.stride(0), P(dv), P(Y), P(W), stream)
            return W

        def residual(Y, rowsum=None):  # D (b - Z D Y)
            if rowsum is None and self.symmetric_products:  # D b - (D Z D) Y from the upper tiles of Z
                return Bs - matvec(Y)
            R = torch.empty_like(B)
            _lib.call("mom_residual", n, nrhs, P(Z), Z.stride(0), P(d * Y), P(B), P(R), P(rowsum), P(dv), stream)
            return d * R

        if X0 is None:
            Y = precond(Bs)
        else:
            Y0 = X0 / d
            Y = Y0 + precond(residual(Y0))
        R = residual(Y, rows)
        (zmax,) = yield from _host(rows.max())
        znorm = float(zmax)
        its = 0
        prev = None
        Vb = torch.empty((m + 1, n, nrhs), dtype=torch.complex128, device=self.device)
        Wb = torch.empty((m, n, nrhs), dtype=torch.complex128, device=self.device)
        Hd = self._buf("gmres_h", (int(_lib.lib().mom_gmres_work_elems(m, nrhs)),), torch.complex128, lane)
        Hh = self._pinned("gmres_hh", ((m + 1) * nrhs,), torch.complex128, lane)
        ev = torch.cuda.Event()
        for _restart in range(self.refine_maxit):
            beta = torch.linalg.vector_norm(R, dim=0)
            ynorm = torch.linalg.vector_norm(Y, dim=0)
            bh, yh = yield from _host(beta, ynorm)
            target = self.refine_tol * znorm * yh
            # converged, or stalled at the FP64 residual floor (within 1e3 of the target)
            if np.all(bh <= target) or (prev is not None and np.all((bh > 0.5 * prev) & (bh < 1e3 * target))):
                break
            prev = bh
            Vb[0] = R / beta
            H = np.zeros((m + 1, m, nrhs), complex)
            g = np.zeros((m + 1, nrhs), complex)
            g[0] = bh
            cs = np.zeros((m, nrhs), complex)
            sn = np.zeros((m, nrhs), complex)
            k = 0
            for j in range(m):
                # one call: preconditioner solve, product, Gram-Schmidt twice, normalisation, and
                # the new Hessenberg column copied to pinned host memory
                pc.arnoldi(n, nrhs, j, m, Z, dv, swk if self.symmetric_products else None,
                           wbs if self.symmetric_products else 0, Vb, Wb, Hd, Hh, stream)
                ev.record()
                while not ev.query():
                    yield
                hc = Hh.numpy()[:(j + 2) * nrhs].reshape(j + 2, nrhs).copy()
                its += 1
                givens_column(hc, cs, sn, g, j)
                H[:j + 2, j] = hc
                k = j + 1
                if np.all(np.abs(g[j + 1]) <= 0.25 * target):
                    break
                if max_its is not None and its >= max_its:
                    self.last_refine = its
                    self.lu_stats["gmres_steps"] += its
                    return None
            y = np.zeros((k, nrhs), complex)
            for c in range(nrhs):
                y[:, c] = np.linalg.solve(np.triu(H[:k, :k, c]), g[:k, c])
            yd = self._upload("gmres_y", -y, lane)
            _lib.call("mom_basis_sub", n, nrhs, k, P(Wb), P(yd), P(Y), stream)
            R = residual(Y)
        else:
            # restarts exhausted: report a result short of the FP64 floor
            beta = torch.linalg.vector_norm(R, dim=0)
            ynorm = torch.linalg.vector_norm(Y, dim=0)
            bh, yh = yield from _host(beta, ynorm)
            target = self.refine_tol * znorm * yh
            if not np.all(bh < 1e3 * target):
                warnings.warn(f"mixed-precision solve: GMRES stopped after {its} iterations at a normwise backward "
                              f"error of {np.max(bh / (znorm * yh)):.1e} (target {self.refine_tol:.0e})",
                              RuntimeWarning, stacklevel=2)
        self.last_refine = its
        self.lu_stats["gmres_steps"] += its
        return d * Y

    def solve_currents(self, freq, excitations=None, greens=None, Z=None, interp=None):
        """RWG (+ via segment) coefficients for the given excitations (default:
        one column per port), with the lumped loads applied; Z: a precomputed
        impedance matrix (numpy); interp: a FreqInterp (set up) supplying Z
        (Engine/mom25d/solver.py:984-1031)."""
        g = greens or self.greens(freq)
        if excitations is None:
            excitations = self.port_excitations()
        ne = excitations.shape[1]
        if Z is not None:
            interp = _GivenZ(Z)
        X = self._solve_dev(g, self._with_gaps(excitations), interp).cpu().numpy()
        return self._apply_lumped(X, ne, freq)

    # ------------------------------------------------------------- lumped
    def _with_gaps(self, excitations):
        """The excitations followed by one delta-gap column per lumped gap."""
        E = np.asarray(excitations)
        return E if self.Lm is None else np.hstack([E, self.Lm])

    def lumped_impedance(self, freq):
        """D (n_gaps x n_gaps): the impedances across the lumped gaps at freq
        (the loads of Engine/mom25d/solver.py:1039-1061).  An internal port's
        gap is terminated in its z_ref; an N-port component fills the block
        of its gaps; every other gap takes its value (a sweep's `values`
        override the layout's), and a gap without one stays open (OPEN)."""
        from .lumped import as_value, parse

        n_gaps = 0 if self.Lm is None else self.Lm.shape[1]
        D = np.zeros((n_gaps, n_gaps), complex)
        position = {q.name: k for k, q in enumerate(self.lumped)}
        taken = np.zeros(n_gaps, bool)  # filled by an internal port or a component
        internal = np.zeros(n_gaps, bool)
        for ip, c in zip(self.iports, self.iport_index):
            D[c, c] = ip["z_ref"]
            taken[c] = internal[c] = True
        for model, gaps in self.components:
            block = [position[g] for g in gaps]
            D[np.ix_(block, block)] = model.z(freq)
            taken[block] = True
        for k, q in enumerate(self.lumped):
            value = self.values.get(q.name, q.value)
            if taken[k] and (value is None or internal[k]):
                continue  # a component's own entry stays unless a value is given for the gap itself
            if value is None:
                D[k, k] = self.OPEN
            else:
                D[k, k] = as_value(parse(value) if isinstance(value, str) else value).z(freq)
        return D

    def _load_lumped(self, Xe, Xi, freq):
        """Currents with the loads D across the gaps, from the unloaded
        solutions Xe = Z^-1 e and Xi = Z^-1 L (L: the gaps' delta-gap
        columns).  The loaded system (Z + L D L^T) x = e gives Z x = e - L D y
        with y = L^T x the gap currents, so x = Xe - Xi D y, and applying L^T
        gives (I + L^T Xi D) y = L^T Xe: one small solve per frequency
        (Engine/mom25d/solver.py:1063-1069)."""
        D = self.lumped_impedance(freq)
        y = np.linalg.solve(np.eye(len(D)) + (self.Lm.T @ Xi) @ D, self.Lm.T @ Xe)
        return Xe - Xi @ (D @ y)

    def _apply_lumped(self, X, ne, freq):
        """Loaded solution of the first ne columns from the solve of
        [excitations, gap columns]."""
        return X[:, :ne] if self.Lm is None else self._load_lumped(X[:, :ne], X[:, ne:], freq)

    def lumped_currents(self, x):
        """Current a -> b through each lumped gap, (n_gaps, Nexc): the gap
        columns' signed edge lengths applied to x."""
        return np.einsum("ng,ne->ge", self.Lm, x)

    def _solve_dev(self, g, excitations, interp=None):
        B = torch.from_numpy(np.ascontiguousarray(excitations, complex)).to(self.device)
        X = self._run(self._solve_task(g, B, 0, interp=interp)).clone()
        self._check_solve(0)
        return X

    def _solve_task(self, g, B, lane, graph=False, reuse=False, interp=None):
        """Generator: fill, LU and port solve of one frequency on the current
        stream with the buffers of `lane`; returns the device currents (a
        buffer of the lane: read it before the lane's next solve).  reuse
        (mixed precision, within a sweep): precondition with the lane's last
        LU when the reuse policy says so (reuse_lu).  interp: a FreqInterp
        (set up for this solver) supplies Z instead of the fill."""
        n = self.mesh.n_unknowns
        self._clear_status(lane)
        Z = self._buf("Z", (n, n), torch.complex128, lane)
        # mixed precision relies on a symmetric Z (the fill's and the interpolation's are); a given
        # non-symmetric matrix is solved by the complex128 LU instead
        mixed = self.precision == "mixed" and getattr(interp, "symmetric", True)
        Z32 = self._buf("Z32", (n, n), torch.complex64, lane) if mixed else None

        def fill(Z32x):
            if interp is None:
                self._fill(g, Z, Z32x, lane)
                return
            interp.fill(g.freq, Z, lane)
            if Z32x is not None:
                self._scale_copy(Z, Z32x, lane)

        sweep = mixed and reuse and self.reuse_lu
        a = self._anchor.get(lane) if sweep else None
        if a is not None and self._jumped(lane, g.freq):
            a = None  # the lane moved to another part of the sweep: its factors are of a distant frequency
        filled = False
        steps = []  # (preconditioner, GMRES steps) of this point, for point_log

        def log(X):
            self.point_log.append(dict(freq=float(g.freq), lane=lane, solves=steps))
            return X

        if a is not None and self._reuse_next(a, g.freq):
            fill(None)
            filled = True
            cap = max(20, int(a["lu_equiv"] + a["its0"]))
            X = yield from self._gmres_chunks(Z, a["pc"], B, lane, max_its=cap, X0=self._extrapolate(lane, g.freq, B))
            steps.append((f"reuse{a['prec']}@{a['freq'] / 1e9:.3g}", self.last_refine))
            if X is not None:
                self._remember(lane, g.freq, X)
                a["cost"] += self.last_refine
                a["n"] += 1
                a["last"] = self.last_refine
                self.lu_stats["reused"] += 1
                return log(X)
            self.lu_stats["fallbacks"] += 1
        wide = self.lu64_steps is not None and sweep
        if wide and g.freq > 2 * self._f_poor:
            # a lane far below the complex64 points in progress waits for one
            # of them to converge, then predicts its own (a doomed complex64
            # factorisation costs an LU and lu64_steps steps).  Its own
            # intent is registered first, so that the lanes that start in the
            # same round see each other; waits only point to frequencies at
            # least twice as high, so no two lanes wait for each other.
            self._c64_pending[lane] = g.freq
            if self._active_lanes > 1 and not self._c64_refs_above(g.freq):
                yield  # the other lanes start their first points
                while not self._c64_refs_above(g.freq) and any(
                        f >= 2 * g.freq for ln, f in self._c64_pending.items() if ln != lane):
                    yield
            self._c64_pending.pop(lane, None)
            if self._c64_predicted_poor(g.freq):
                self.lu_stats["predicted_poor"] += 1
                self._f_poor = max(self._f_poor, g.freq)
        if not (wide and g.freq <= 2 * self._f_poor):
            if filled:
                self._scale_copy(Z, Z32, lane)
            else:
                fill(Z32)
                filled = True
            # in a sweep, a complex64 solve still short of the target after
            # lu64_steps steps is abandoned for a complex128 anchor
            if wide:
                self._c64_pending[lane] = g.freq
            try:
                X = yield from self._lu_solve(Z, B, Z32, lane, graph, cap=self.lu64_steps if wide else None)
            finally:
                self._c64_pending.pop(lane, None)
            if mixed:
                steps.append(("lu32", self.last_refine))
            if X is not None:
                if mixed:
                    self.lu_stats["factorisations"] += 1
                    if sweep:
                        self._c64_refs.append((float(g.freq), self.last_refine))
                        self._remember(lane, g.freq, X)
                        self._set_anchor(lane, g.freq, self._lane_pc[lane], self.last_refine)
                return log(X)
            self.lu_stats["abandoned"] += 1
            self._f_poor = max(self._f_poor, g.freq)
        # a band where complex64 factors precondition poorly: complex128 anchor
        if not filled:
            fill(None)
        pc = yield from self._lu64_factor(Z, lane)
        X = yield from self._gmres_chunks(Z, pc, B, lane, max_its=self.NOPIV_CAP if self.nopiv else None)
        steps.append(("lu64", self.last_refine))
        if X is None:  # unpivoted factors too poor here: factor again with pivoting
            self.lu_stats["pivoted_retries"] += 1
            pc = yield from self._lu64_factor(Z, lane, pivot=True)
            X = yield from self._gmres_chunks(Z, pc, B, lane)
            steps.append(("lu64-pivoted", self.last_refine))
        self._remember(lane, g.freq, X)
        self.lu_stats["factorisations"] += 1
        self._set_anchor(lane, g.freq, pc)
        return log(X)

    def _jumped(self, lane, freq):
        """True when freq is more than three of the lane's last spacings away
        from its last point (the lane was handed another part of the sweep)."""
        h = self._history.get(lane, [])
        if len(h) < 2:
            return False
        step = abs(h[-1][0] - h[-2][0])
        return step > 0 and abs(freq - h[-1][0]) > 3 * step

    def _c64_refs_above(self, freq):
        return any(f > freq for f, _ in self._c64_refs)

    # GMRES target of the step-count model below (relative residual reduction)
    _C64_TARGET = 1e-15

    def _c64_predicted_poor(self, freq):
        """True when the nearest higher frequency whose own complex64
        factorisation converged (in its_r GMRES steps) predicts more than
        lu64_steps steps at freq: its contraction per step
        rho_r = target^(1 / its_r) is scaled by (f_r / freq)^2 -- the charge
        term of the MPIE grows as 1 / f^2 against the current term, and the
        relative error of complex64 factors with the condition number."""
        refs = [(f, its) for f, its in self._c64_refs if f > freq]
        if not refs or self.lu64_steps is None:
            return False
        fr, its = min(refs)
        rho = self._C64_TARGET ** (1.0 / max(its, 1)) * (fr / freq) ** 2
        return rho >= 1.0 or np.log(self._C64_TARGET) / np.log(rho) > self.lu64_steps

    def _set_anchor(self, lane, freq, pc, its0=0):
        """The lane's factorisation (a Precond) that preconditions its next points."""
        eq = self._lu_equiv()
        self._anchor[lane] = dict(freq=freq, prec=pc.prec, pc=pc, lu_equiv=eq, its0=its0, cost=eq + its0, n=1,
                                  last=None)

    def _lu64_factor(self, Z, lane, pivot=False):
        """Generator: complex128 LU of D Z D (D = diag(|Z_ii|^-1/2), into the lane's
        dscale) with the solve data of its factors (super-block inverses
        and scaled couplings, or the 64-row inverted blocks), enqueued on the
        current stream; returns the Precond of the mixed path's GMRES.  Z
        must be symmetric (as the fill leaves it).  Without row interchanges
        when nopiv is set and pivot is False."""
        n = Z.shape[0]
        P = _lib.ptr
        stream = _lib.stream_ptr()
        L = _lib.lib()
        dsc = self._buf("dscale", (n,), torch.float64, lane)
        _lib.call("mom_diag_scale", n, P(Z), Z.stride(0), P(dsc), stream)
        A = self._buf("A64", (n, n), torch.complex128, lane)
        torch.mul(Z, dsc[:, None], out=A)
        A.mul_(dsc[None, :])
        ipiv = self._buf("ipiv64", (n,), torch.int32, lane)
        perm = self._buf("perm64", (n,), torch.int32, lane)
        info = self._buf("info", (4,), torch.int32, lane)
        nopiv = self.nopiv and not pivot
        groups = 0 if nopiv else self._panel_groups()
        jb = self.panel  # (tuned for complex64) halved until the panel's rows fit the LDS of `groups` workgroups
        while groups and jb > 16 and -(-n // max(1, self.PANEL_LDS // (16 * jb))) > groups:
            jb //= 2
        wbytes = int(L.mom_getrf_work_bytes(64, n, jb))
        work = self._buf("work64", (wbytes,), torch.uint8, lane)
        sb = self._sb_rows(n)
        if sb:
            swb = int(L.mom_sbsolve_work_bytes(64, n, sb))
            sw = self._buf("sbwork64", (swb,), torch.uint8, lane, zero=True)
        else:
            sw = self._buf("swork64i", (int(L.mom_getrs_inv_work_bytes(64, n)),), torch.uint8, lane, zero=True)
        def calls():
            if nopiv:
                _lib.call(self._nopiv_entry, 64, n, P(A), A.stride(0), P(ipiv), P(info), P(work), wbytes, jb, stream)
            else:
                _lib.call("mom_getrf", 64, n, P(A), A.stride(0), P(ipiv), P(info), P(work), wbytes, jb, groups,
                          stream)
            _lib.call("mom_build_perm", n, P(ipiv), P(perm), stream)
            if sb:
                _lib.call("mom_sbsolve_setup", 64, n, P(A), A.stride(0), sb, P(sw), swb, stream)
            else:
                _lib.call("mom_trtri_blocks", 64, n, P(A), A.stride(0), P(sw), stream)

        if self.block is not None:  # the tuned update depth serves the complex128 factors too
            _lib.call("mom_set_lu_block", 64, self.block)
        yield from _on_launcher(calls, self.device)
        self.lu_stats["fp64_lu"] += 1
        return Precond(64, A, perm, sw, sb)

    def _remember(self, lane, freq, X):
        """Keep the lane's last four solutions (device) for _extrapolate."""
        h = self._history.setdefault(lane, [])
        h.append((float(freq), X))
        del h[:-4]

    def _extrapolate(self, lane, freq, B):
        """Initial guess at freq from the lane's last solutions: the Lagrange
        polynomial through the (up to three) nearest of them within four
        spacings, evaluated at freq, when one lies within two spacings of
        freq (None otherwise; the spacing is the smallest between the kept
        solutions, which may lie on both sides of freq).  The solution varies
        smoothly with frequency away from resonances; GMRES corrects whatever
        the extrapolation misses."""
        h = list({f: X for f, X in self._history.get(lane, []) if X.shape == B.shape}.items())  # latest per f
        fs = sorted(f for f, _ in h)
        if len(fs) < 2:
            return None
        step = min(b - a for a, b in zip(fs, fs[1:]))
        if step <= 0 or min(abs(freq - f) for f in fs) > 2.0 * step:
            return None
        near = sorted((abs(freq - f), k) for k, (f, _) in enumerate(h) if abs(freq - f) <= 4.0 * step)
        pts = [h[k] for _, k in near[:3]]
        X0 = torch.zeros_like(B)
        for i, (fi, Xi) in enumerate(pts):
            w = 1.0
            for j, (fj, _) in enumerate(pts):
                if j != i:
                    w *= (freq - fj) / (fi - fj)
            X0 += w * Xi
        return X0

    def _lu_equiv(self):
        """Cost of one complex64 factorisation in GMRES steps at this N: a
        model of this implementation on MI355X -- the LU at ~50 TFLOP/s
        effective or ~2.5 us per column of panel latency, whichever is larger
        (unpivoted, nopiv; pivoted: ~32 TFLOP/s or ~8 us per column; L D L^T,
        ldlt: half the LU's flops at the same rate: 54 ms against 93 ms
        measured at n = 12 428), plus
        the super-block solve's set-up (8 n^2 sb flop of scaled couplings); a
        step = two triangular solves (64-row chain: ~2.9 us per 64 rows;
        super-block chain: the larger of ~4.5 us per super-block and one
        read of the factors at ~3 TB/s) + one FP64 read of Z (of its upper
        triangle with symmetric_products) + ~0.2 ms of host round trip --
        times 1.4, because the steps of
        concurrent lanes overlap better than factorisations do (measured on
        the 12.4 k-unknown pour, unpivoted: 17.3 points/s at the model's 78,
        16.9 at 123, 14.6 at 50)."""
        if self.lu_equiv is not None:
            return float(self.lu_equiv)
        n = self.mesh.n_unknowns
        rate, per_col = (50e12, 2.5e-6) if self.nopiv else (32e12, 8e-6)
        flops = (4 / 3 if self.nopiv and self.ldlt else 8 / 3) * n ** 3
        t_lu = max(flops / rate, per_col * n)
        sb = self._sb_rows(n)
        if sb:
            t_lu += 8 * n * n * sb / rate
            t_solve = 2 * max(n / sb * 4.5e-6, 4 * n * n / 3e12)
        else:
            t_solve = 2 * (n / 64) * 2.9e-6
        t_step = t_solve + (8 if self.symmetric_products else 16) * n * n / 5e12 + 2e-4
        return 1.4 * t_lu / t_step

    @staticmethod
    def _reuse_next(a, freq):
        """Reuse the anchor's LU for freq: always once, then while the last
        reuse cost fewer steps than the average per point since the
        factorisation; a complex64 LU never beyond half an octave from the
        anchor (a complex128 one preconditions several octaves well: its
        step count grows slowly, and the cost rule ends its reuse)."""
        if a["prec"] == 32 and abs(freq - a["freq"]) > 0.5 * a["freq"]:
            return False
        return a["last"] is None or a["last"] < a["cost"] / a["n"]

    # ----------------------------------------------------------- de-embed
    def _divergence(self):
        """Sparse map from the unknowns to the surface divergence of the
        current on each triangle, (Nt x N): an RWG basis contributes
        +-l / A on its two triangles, a via segment +-1 / (pad area) on the
        triangles of its top and bottom pads (its current leaves the top pad
        and enters the bottom one).  Built once per mesh."""
        if getattr(self, "_div_op", None) is None:
            import scipy.sparse as sp

            m = self.mesh
            b = m.tri_basis
            t_idx, slot = np.nonzero(b >= 0)
            cols = b[t_idx, slot]
            vals = m.tri_sign[t_idx, slot] * m.edge_len[cols] / m.area[t_idx]
            rows, cs, vs = [t_idx], [cols], [vals]
            for k, seg in enumerate(m.via_segments):
                for key, sign in (("top_tris", 1.0), ("bottom_tris", -1.0)):
                    pads = seg[key]
                    if pads is not None and len(pads):
                        pads = np.asarray(pads)
                        rows.append(pads)
                        cs.append(np.full(pads.size, m.n_basis + k))
                        vs.append(np.full(pads.size, sign / m.area[pads].sum()))
            self._div_op = sp.csr_matrix((np.concatenate(vs), (np.concatenate(rows), np.concatenate(cs))),
                                         shape=(len(m.tris), m.n_unknowns))
        return self._div_op

    def via_currents(self, x):
        """Downward current of each via segment (mesh.via_segments) in the
        solution x (Engine/mom25d/solver.py:1093-1095)."""
        return x[self.mesh.n_basis:self.mesh.n_basis + self.mesh.n_vias]

    def frill_voltages(self, x):
        """Ring voltages of each antipad (mesh.frills; the first of each the
        barrel-to-plane voltage) in the solution x
        (Engine/mom25d/solver.py:1097-1099)."""
        return x[self.mesh.n_basis + self.mesh.n_vias:]

    def charges(self, x, omega):
        """Surface charge per triangle, (Nt, Nexc), from continuity:
        q = -div(J) / (j omega) (Engine/mom25d/solver.py:1076-1091)."""
        return -(self._divergence() @ x) / (1j * omega)

    def line_voltage(self, greens, q, pts, layer="top"):
        """Voltage to the ground plane at points of a metal layer from the
        triangle charges q (Nt, Nexc) (GPU kernel)."""
        return torch.cat(self._line_voltage_dev(greens, q, pts, layer), 1).cpu().numpy()

    def _line_voltage_dev(self, greens, q, pts, layer, lane=0):
        """line_voltage enqueued on the current stream: the device results,
        in column groups of at most 4 (the kernel's limit); the inputs go up
        through pinned memory, so nothing here waits for the stream."""
        m = self.mesh
        plan = self.geometry()
        d = plan.upload(self.device)
        li = m.layer_names.index(layer)
        tabs = greens.tables
        pts = np.ascontiguousarray(pts, float)
        qd = self._upload("charges", np.ascontiguousarray(q, complex), lane)
        pd = self._upload("vpts", pts, lane)
        P = _lib.ptr
        outs = []
        for c0, c1 in _column_chunks(q.shape[1]):  # the kernel takes up to 4 columns
            qc = qd if q.shape[1] <= 4 else qd[:, c0:c1].contiguous()
            out = torch.empty((pts.shape[0], c1 - c0), dtype=torch.complex128, device=self.device)
            _lib.call("mom_line_voltage", pts.shape[0], P(pd), li, len(m.layer_names), plan.nt, c1 - c0,
                      P(d["tri_layer"]), P(d["V"]), P(d["qp7"]), P(d["qw7"]), P(greens.pairs_dev), P(tabs.tab_x),
                      P(tabs.tab_cw), P(tabs.tab_b), P(qc), 1.0 / EPS0, P(out), _lib.stream_ptr())
            outs.append(out)
        return outs

    def _basis_positions(self):
        """Centres of the basis functions (N x 3): RWG edge midpoints on their
        level, via segments at mid span, antipad rings on the via axis at the
        plane (Engine/mom25d/freqinterp.py:53-64)."""
        m = self.mesh
        pos = np.c_[m.edge_mid[:m.n_basis], self.zl[m.edge_layer[:m.n_basis]]]
        if m.n_vias:
            vz = {n: self.stack.level_z(n) for n in m.layer_names + list(self.stack.planes)}
            vz["gnd"] = 0.0
            zc = [0.5 * (vz[g["bottom"]] + vz[g["top"]]) for g in m.via_segments]
            xy = [m.vias[g["via"]].xy for g in m.via_segments]
            pos = np.r_[pos, np.c_[np.asarray(xy, float).reshape(-1, 2), zc]]
        for fr in getattr(m, "frills", []):  # antipad rings: on the via axis at the plane
            pos = np.r_[pos, [[*fr["xy"], self.stack.level_z(fr["plane"])]] * fr["nr"]]
        return np.ascontiguousarray(pos, float)

    def _fit_context(self):
        """The port description the wave fits need (picklable)."""
        groups = list(self.groups.values())
        first = [self.ports[m[0]] for m in groups]
        st = self.stack
        return FitContext(groups=groups, width=[p.width for p in first], h=[st.level_h(p.layer) for p in first],
                          eps=[st.eps_below(p.layer) for p in first], covered=[st.covered(p.layer) for p in first],
                          z0_from=[p.z0_from for p in self.ports], s=[pm.s for pm in self.pm],
                          lossless=bool(self.lossless), v_weight=self.v_weight, z0_fixed=self.z0_fixed,
                          z0_recip_min=self.z0_recip_min, zi=[ip["z_ref"] for ip in self.iports])

    def _internal_currents(self, x):
        """Current through each internal port's source, (n_internal, Nexc)."""
        return self.Lm[:, self.iport_index].T @ x if self.iports else None

    @property
    def port_names(self):
        return [p.name for p in self.ports] + [ip["name"] for ip in self.iports]

    def _port_iv(self, g, x):
        """Currents and voltages at every port's sample cuts (the voltages
        from the charges through the line-voltage kernel)."""
        return _drain(self._port_iv_task(g, x))

    def _port_iv_task(self, g, x, lane=0):
        """Generator form of _port_iv: yields while the line-voltage kernels
        run (the lane scheduler feeds the other lanes meanwhile)."""
        npt = len(self.ports)
        q = self.charges(x, g.omega)
        layers = list(dict.fromkeys(p.layer for p in self.ports))
        groups = [[p for p in range(npt) if self.ports[p].layer == lay] for lay in layers]
        outs = [self._line_voltage_dev(g, q, np.concatenate([self.pm[p].vpts for p in ps]), lay, lane)
                for lay, ps in zip(layers, groups)]
        host = yield from _host(*[o for grp in outs for o in grp])
        phi = [None] * npt
        k = 0
        for ps, grp in zip(groups, outs):
            vp = np.hstack(host[k:k + len(grp)])
            k += len(grp)
            for p, v in zip(ps, np.split(vp, np.cumsum([len(self.pm[p].vpts) for p in ps])[:-1])):
                phi[p] = v
        I = [self.pm[p].currents(x) for p in range(npt)]
        V = [self.pm[p].voltages(phi[p]) for p in range(npt)]
        return I, V

    @staticmethod
    def _store_point(out, fi, S, gam, zl, zr, res, waves):
        S[fi], gam[fi], zl[fi], res[fi], zr[fi], waves[fi] = out

    def _sparams(self, f, g, x, S, gam, zl, zr, res, waves, fi):
        I, V = self._port_iv(g, x)
        self._store_point(sparams_point(self._fit_context(), f, I, V, self._internal_currents(x)), fi, S, gam, zl, zr,
                          res, waves)

    def _schedule(self, task, count, lanes):
        run_lanes(task, count, lanes, self.device, [self], self.graphs)

    def sweep(self, freqs, keep_currents=False, interp="auto", values=None, checkpoint=None, wait_memory=None):
        """S-parameters at freqs (Engine/mom25d/solver.py:1189-1335).

        interp : "auto" (default) -> interpolate Z between anchor fills
                 (FreqInterp with its defaults) when the sweep has at least
                 twice as many frequencies as anchors, else fill every point;
                 a FreqInterp -> use it; None -> exact fill at every point.
        values : {lumped name: value} overriding the layout's components.
        checkpoint : path of an .npz rewritten (atomically) after every
                 frequency point; a sweep started again with the same path,
                 frequencies, mesh and values solves only the points the file
                 does not hold yet (with keep_currents every point is solved).
        wait_memory : seconds to wait for enough free device memory for one
                 lane's buffers (memory_estimate) before the first fill, then
                 MemoryError; None reads $MOM25D_WAIT_MEMORY (default 0: warn
                 only).  Engine/mom25d/solver.py:1137-1181, 1224-1256."""
        self.values = dict(values or {})
        freqs = np.atleast_1d(np.asarray(freqs, float))
        if isinstance(interp, str):
            if interp != "auto":
                raise ValueError("interp must be 'auto', None or a FreqInterp")
            from .freqinterp import FreqInterp

            interp = FreqInterp().setup(self, freqs)
            if len(freqs) < 2 * len(interp.anchors):
                interp = None
        elif interp is not None:
            interp.setup(self, freqs)
        if interp is not None and self.verbose:
            print(f"  FreqInterp: {len(interp.anchors)} anchor fills for {len(freqs)} frequencies "
                  f"(order {interp.order}, k_e = {interp.k_ratio:.3f} k0)")
        self.last_interp = interp
        if wait_memory is None:
            wait_memory = float(os.environ.get("MOM25D_WAIT_MEMORY", 0) or 0)
        ck = SweepCheckpoint(checkpoint, self, freqs) if checkpoint else None
        todo = np.arange(len(freqs)) if ck is None or keep_currents else np.flatnonzero(~ck.done)
        if len(todo) == 0:
            return ck.result(self.port_names)
        self._wait_memory(self.memory_estimate(), wait_memory)
        lanes = min(self._lane_count(), len(todo))
        # with a checkpoint the wave fits run in this process, so that each point's S is final when it is written
        workers = 0 if ck is not None else self._worker_count(len(todo))
        r = _sweep_many([(self, freqs[todo], interp)], lanes, workers, keep_currents, self.graphs,
                        on_point=None if ck is None else (lambda fi, row: ck.store(todo[fi], row)))[0]
        if ck is None or len(todo) == len(freqs):
            return r
        return ck.result(self.port_names, timing=r.timing)

    def memory_estimate(self):
        """Device memory [bytes] of one lane of a sweep: Z in complex128,
        in mixed precision its complex64 LU copy, and the LU and solve
        workspaces and GMRES basis (~5 % more); the lane count adapts to
        what is free beyond that (Engine/mom25d/solver.py:1137-1151 counts
        3 N x N complex128 on the host)."""
        n = int(self.mesh.n_unknowns)
        per = 16 * n * n + (8 * n * n if self.precision == "mixed" else 0)
        return int(1.05 * per)

    def _wait_memory(self, need, wait):
        """Free device memory against need before the first fill: wait <= 0
        warns, wait > 0 polls for up to `wait` seconds and then raises
        MemoryError (another process may hold the device for a while)."""
        free, _total = torch.cuda.mem_get_info(self.device)
        if need <= free:
            return
        msg = f"estimated {need / 1e9:.1f} GB > free device memory {free / 1e9:.1f} GB (N = {self.mesh.n_unknowns})"
        if not wait or wait <= 0:
            warnings.warn(f"{msg}; continuing (sweep(wait_memory=seconds) waits instead)", RuntimeWarning,
                          stacklevel=3)
            return
        t_end = time.time() + wait
        while time.time() < t_end:
            time.sleep(min(5.0, max(0.2, wait / 20)))
            free, _total = torch.cuda.mem_get_info(self.device)
            if need <= free:
                return
        raise MemoryError(f"{msg} after waiting {wait:.0f} s")

    # ---------------------------------------------------------- reference Z0
    def _reciprocal_z0(self, SI, zphi, prior=1e-4):
        """Port reference impedances (Engine/mom25d/solver.py:1357-1397)."""
        return reciprocal_z0(self._fit_context(), SI, zphi, prior)


def _host(*tensors):
    """Generator: numpy copies of device tensors, read without holding up the
    other lanes: asynchronous copies into pinned host memory on the current
    stream, then it yields until they have landed (a blocking read would stop
    the host from feeding the other lanes' streams meanwhile)."""
    bufs = []
    for t in tensors:
        h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
        h.copy_(t, non_blocking=True)
        bufs.append(h)
    ev = torch.cuda.Event()
    ev.record()
    while not ev.query():
        yield
    return [h.numpy().copy() for h in bufs]


_LAUNCHER = None


def _on_launcher(fn, device):
    """Generator: fn() -- library calls that only enqueue work on the streams
    they are given, such as a factorisation's several hundred launches -- on
    a launcher thread (ctypes releases the GIL), yielding until it has
    returned, so that the lane scheduler keeps feeding the other lanes
    meanwhile.  The caller enqueues nothing more on those streams until then."""
    global _LAUNCHER
    if _LAUNCHER is None:
        from concurrent.futures import ThreadPoolExecutor

        _LAUNCHER = ThreadPoolExecutor(max_workers=8, thread_name_prefix="mom-launch")

    def run():
        with torch.cuda.device(device):
            return fn()

    fut = _LAUNCHER.submit(run)
    while not fut.done():
        yield
    return fut.result()


def _drain(gen):
    """Run a generator task to its end (no other task to switch to)."""
    try:
        while True:
            next(gen)
    except StopIteration as e:
        return e.value


def _column_chunks(n, width=4):
    """Column ranges of at most `width` columns covering 0 .. n-1."""
    return [(c, min(c + width, n)) for c in range(0, n, width)]


class _GivenZ:
    """Stands in for a FreqInterp: supplies a given impedance matrix (numpy)."""

    def __init__(self, Z):
        self.Z = np.ascontiguousarray(Z, complex)
        # the mixed path's GMRES reads only one triangle of Z; another Z takes the complex128 LU
        self.symmetric = bool(np.array_equal(self.Z, self.Z.T))

    def fill(self, f, Zdev, lane=0):
        Zdev.copy_(torch.from_numpy(self.Z))
        return Zdev


class LaneCursor:
    """Hands the indices 0 .. count-1 to `lanes` lanes so that each lane
    works through neighbouring indices (neighbouring frequencies, whose
    factorisations can precondition each other): the range is cut into one
    segment per lane, and a lane starts at its segment's middle and takes
    the neighbours on alternate sides outward (above first), so that the
    points that reuse its first factorisation lie within half a segment of
    it.  The lowest segment is taken upward from its first index instead:
    at the low end of a sweep complex64 factors may fail, and the complex128
    one that then anchors it preconditions over octaves.  A lane whose
    segment is used up takes the untaken outer end (of any segment with at
    least two indices left) nearest to its own last index and works inward
    from there, toward that segment's owner, who keeps working outward; it
    gets None meanwhile when every segment has at most one index left (the
    owners take those).  Every index is handed out exactly once."""

    def __init__(self, count, lanes):
        nseg = min(lanes, count)  # lanes beyond the count have no segment of their own
        cuts = np.round(np.arange(nseg + 1) / max(nseg, 1) * count).astype(int)
        # per segment [lo, a, b, hi]: the untaken indices are lo..a and b..hi
        self.seg = []
        for k in range(nseg):
            lo, hi = int(cuts[k]), int(cuts[k + 1]) - 1
            c = (lo + hi) // 2 if k > 0 else lo
            self.seg.append([lo, c, c + 1, hi])
        self.down = [True] * lanes  # the owner's next side (the first, the middle, is the lower side's top)
        self.thief = [None] * lanes  # (segment, side) a lane works inward from, once its own is used up
        self.last = [None] * lanes

    def _left(self, r):
        lo, a, b, hi = self.seg[r]
        return max(a - lo + 1, 0) + max(hi - b + 1, 0)

    def _own(self, lane):
        lo, a, b, hi = self.seg[lane]
        for _ in range(2):
            if self.down[lane] and a >= lo:
                self.seg[lane][1] -= 1
                self.down[lane] = b > hi  # stays below once the upper side is used up
                return a
            if not self.down[lane] and b <= hi:
                self.seg[lane][2] += 1
                self.down[lane] = a >= lo
                return b
            self.down[lane] = not self.down[lane]
        return None

    def _steal(self, lane):
        ends = []
        for r, (lo, a, b, hi) in enumerate(self.seg):
            if self._left(r) < 2:  # the owner's next index stays with the owner (its factors are near)
                continue
            if lo <= a:
                ends.append((lo, r, "lo"))
            if b <= hi:
                ends.append((hi, r, "hi"))
        if not ends:
            return None
        ref = self.last[lane]
        if ref is None:  # a lane without a segment: the end of the segment with the most left
            i, r, side = max(ends, key=lambda e: (self._left(e[1]), -e[0]))
        else:
            i, r, side = min(ends, key=lambda e: (abs(e[0] - ref), e[0]))
        self.thief[lane] = (r, side)
        return self._inward(lane)

    def _inward(self, lane):
        r, side = self.thief[lane]
        lo, a, b, hi = self.seg[r]
        if side == "lo" and lo <= a:
            self.seg[r][0] += 1
            return lo
        if side == "hi" and b <= hi:
            self.seg[r][3] -= 1
            return hi
        self.thief[lane] = None
        return None

    def take(self, lane):
        i = None
        if self.thief[lane] is not None:
            i = self._inward(lane)
        elif lane < len(self.seg):
            i = self._own(lane)
        if i is None:
            i = self._steal(lane)
        if i is not None:
            self.last[lane] = i
        return i


class RunCursor:
    """LaneCursor's alternative for batches of several meshes, whose points
    are listed mesh after mesh: the range is cut into one
    run per pair of lanes, and the two lanes of a run take its indices from
    the two ends toward the middle (a lone lane takes its run upward).  A
    lane whose run is used up splits the run with the most indices left and
    takes the upper half upward from its middle (a lane that was taking that
    run downward stays with the upper half).  Every index is handed out
    exactly once."""

    def __init__(self, count, lanes):
        sizes = [2] * (lanes // 2) + [1] * (lanes % 2)
        cuts = np.round(np.cumsum([0] + sizes) / lanes * count).astype(int)
        self.lo = [int(c) for c in cuts[:-1]]
        self.hi = [int(c) - 1 for c in cuts[1:]]  # inclusive
        self.run, self.down = [], []
        for r, k in enumerate(sizes):
            self.run += [r] * k
            self.down += [False, True][:k]

    def _split(self, lane):
        left = [h - lo + 1 for lo, h in zip(self.lo, self.hi)]
        r = int(np.argmax(left))
        if left[r] <= 0:
            return False
        mid = (self.lo[r] + self.hi[r] + 1) // 2
        self.lo.append(mid)
        self.hi.append(self.hi[r])
        self.hi[r] = mid - 1
        nr = len(self.lo) - 1
        for other in range(len(self.run)):
            if self.run[other] == r and self.down[other]:
                self.run[other] = nr
        self.run[lane], self.down[lane] = nr, False
        return True

    def take(self, lane):
        r = self.run[lane]
        if self.lo[r] > self.hi[r]:
            if not self._split(lane):
                return None
            r = self.run[lane]
        if self.down[lane]:
            self.hi[r] -= 1
            return self.hi[r] + 1
        self.lo[r] += 1
        return self.lo[r] - 1


_LANE_STREAMS = {}


def _lane_streams(device, lanes):
    """The lanes' streams, kept for the process's lifetime and reused by
    every sweep (the library keeps a look-ahead stream and events per caller
    stream, created on first use)."""
    key = torch.device(device)
    pool = _LANE_STREAMS.setdefault(key, [])
    while len(pool) < lanes:
        pool.append(torch.cuda.Stream(key))
    return pool[:lanes]


def run_lanes(task, count, lanes, device, solvers, graphs=False, cursor=None):
    """Run task(i, lane) for i < count, `lanes` at a time, round robin.
    Each task is a generator that enqueues GPU work on the current stream and
    yields before it waits for it; the lane's stream is made current whenever
    its task resumes, so the GPU work of all lanes overlaps.  Tasks complete
    in order within a lane, and a lane takes its next index (cursor, by
    default LaneCursor: neighbouring indices stay on one lane) as soon as
    its task ends.  The
    solvers bound their LU panels for `lanes` factorisations in flight
    meanwhile."""
    prev = torch.cuda.current_stream(device)
    own = lanes > 1 or graphs  # graphs are captured on a stream of their own
    streams = _lane_streams(device, lanes) if own else [prev]
    for st in streams:
        if st is not prev:
            st.wait_stream(prev)
    for s in solvers:
        s._active_lanes = lanes
    active = [None] * lanes
    cursor = (cursor or LaneCursor)(count, lanes)
    try:
        while True:
            busy = False
            for lane in range(lanes):
                if active[lane] is None:
                    i = cursor.take(lane)
                    if i is not None:
                        active[lane] = task(i, lane)
                if active[lane] is None:
                    continue
                busy = True
                torch.cuda.set_stream(streams[lane])
                try:
                    next(active[lane])
                except StopIteration:
                    active[lane] = None
            if not busy:
                break
    finally:
        torch.cuda.set_stream(prev)
        for st in streams:
            if st is not prev:
                prev.wait_stream(st)
        for s in solvers:
            s._active_lanes = 1


class SweepCheckpoint:
    """The results of a sweep so far in an .npz file, keyed by a digest of
    the frequencies, the mesh and the lumped values: rewritten through a
    temporary file and os.replace after every point, so that a sweep killed
    at any moment leaves the previous complete file (the resume semantics of
    Engine/mom25d/solver.py:1238-1256, 1330-1333)."""

    FIELDS = ("S", "gam", "zl", "zr", "res", "waves")

    def __init__(self, path, solver, freqs):
        import hashlib

        self.path = os.fspath(path)
        self.freqs = np.asarray(freqs, float)
        m = solver.mesh
        h = hashlib.sha256()
        for a in (self.freqs, m.nodes, m.tris, m.tri_layer):
            h.update(np.ascontiguousarray(a).tobytes())
        h.update(repr(sorted((str(k), repr(v)) for k, v in solver.values.items())).encode())
        self.key = h.hexdigest()
        nf, n = len(freqs), len(solver.ports) + len(solver.iports)
        self.arrays = dict(S=np.zeros((nf, n, n), complex), gam=np.zeros((nf, n), complex),
                           zl=np.zeros((nf, n), complex), zr=np.zeros((nf, n)), res=np.zeros((nf, n)),
                           waves=np.zeros((nf, n, n), complex))
        self.done = np.zeros(nf, bool)
        if os.path.exists(self.path):
            try:
                with np.load(self.path) as f:
                    if str(f["key"]) == self.key and f["S"].shape == self.arrays["S"].shape:
                        self.done = f["done"].astype(bool)
                        for k in self.FIELDS:
                            self.arrays[k] = f[k].copy()
                    elif solver.verbose:
                        print("  checkpoint: the file belongs to another sweep; starting afresh", flush=True)
            except (OSError, KeyError, ValueError) as exc:  # unreadable: start afresh
                if solver.verbose:
                    print(f"  checkpoint: unreadable ({exc}); starting afresh", flush=True)
                self.done[:] = False

    def store(self, fi, row):
        """Point fi's (S, gamma, z0_line, residual, z0, waves) as
        sparams_point returns them, then the file."""
        for k, v in zip(("S", "gam", "zl", "res", "zr", "waves"), row):
            self.arrays[k][fi] = v
        self.done[fi] = True
        tmp = self.path + ".tmp.npz"
        np.savez(tmp, key=np.array(self.key), done=self.done, **self.arrays)
        os.replace(tmp, self.path)

    def result(self, names, timing=None):
        a = self.arrays
        return SweepResult(self.freqs, a["S"], a["zr"], a["gam"], a["zl"], a["res"], names, None, a["waves"],
                           timing=timing or {})


def _sweep_many(jobs, lanes, workers, keep_currents=False, graphs=False, on_point=None):
    """Sweeps of (solver, freqs) jobs as one schedule of (job, frequency)
    points on `lanes` streams; wave fits in `workers` processes (0: here).
    on_point(fi, result row) is called for each point of the first job as
    soon as its S is known (only with workers = 0)."""
    out = []
    points = []
    t0 = time.time()
    for ji, (s, freqs, *rest) in enumerate(jobs):
        nf, n = len(freqs), len(s.ports) + len(s.iports)
        o = dict(solver=s, freqs=freqs, S=np.zeros((nf, n, n), complex), gam=np.zeros((nf, n), complex),
                 zl=np.zeros((nf, n), complex), zr=np.zeros((nf, n)), res=np.zeros((nf, n)),
                 waves=np.zeros((nf, n, n), complex), cur=[None] * nf, futures={}, host=0.0)
        o["tabs"] = s.tables(freqs)
        s.geometry().upload(s.device)
        o["B"] = torch.from_numpy(np.ascontiguousarray(s._with_gaps(s.port_excitations()), complex)).to(s.device)
        o["ctx"] = s._fit_context()
        o["interp"] = rest[0] if rest else None
        out.append(o)
        points += [(ji, fi) for fi in range(nf)]
    device = jobs[0][0].device
    n_exc = [len(o["solver"].ports) + len(o["solver"].iports) for o in out]
    for s, *_ in jobs:
        s._anchor = {}  # no factorisation or solution carries over from an earlier call
        s._history = {}
        s._f_poor = 0.0
        s._c64_refs = []
        s._c64_pending = {}
    torch.cuda.synchronize(device)
    t1 = time.time()
    pool = fit_pool(workers)
    done = [0]

    def store(o, fi, r):
        Solver._store_point(r, fi, o["S"], o["gam"], o["zl"], o["zr"], o["res"], o["waves"])

    def point(k, lane):
        ji, fi = points[k]
        o = out[ji]
        s = o["solver"]
        g = o["tabs"][fi]
        x_dev = yield from s._solve_task(g, o["B"], lane, graphs, reuse=True, interp=o["interp"])
        # read without blocking: the other lanes enqueue their work meanwhile
        xh, *status = yield from _host(x_dev, *s._status_tensors(lane))
        f = o["freqs"][fi]
        x = s._apply_lumped(xh, n_exc[ji], f)
        s._check_solve(lane, status)
        if keep_currents:
            o["cur"][fi] = x
        I, V = yield from s._port_iv_task(g, x, lane)
        h0 = time.perf_counter()
        Ii = s._internal_currents(x)
        if pool is None:
            row = sparams_point(o["ctx"], f, I, V, Ii)
            store(o, fi, row)
            if on_point is not None and ji == 0:
                on_point(fi, row)
        else:
            o["futures"][fi] = pool.submit(sparams_point, o["ctx"], f, I, V, Ii)
        o["host"] += time.perf_counter() - h0
        done[0] += 1
        if s.verbose:
            print(f"PROGRESS {done[0]}/{len(points)}", flush=True)

    run_lanes(point, len(points), lanes, device, [o["solver"] for o in out], graphs,
              cursor=RunCursor if len(out) > 1 else LaneCursor)
    for o in out:
        for fi, fut in o["futures"].items():
            store(o, fi, fut.result())
    t2 = time.time()
    return [SweepResult(o["freqs"], o["S"], o["zr"], o["gam"], o["zl"], o["res"], o["solver"].port_names,
                        o["cur"] if keep_currents else None, o["waves"],
                        timing=dict(tables=t1 - t0, solve=t2 - t1, sparams=o["host"])) for o in out]


def sweep_batch(problems, precision="mixed", lanes=None, workers=None, graphs=False, keep_currents=False, **kw):
    """Sweeps of several meshes of uneven size in one schedule.

    problems : list of (mesh, Substrate | Stackup, freqs).
    Every (mesh, frequency) point is a task of one lane scheduler, so the
    latency-bound steps of the small problems (LU panels, triangular solves,
    host de-embedding) overlap the matrix-core work of the others (graphs:
    each lane replays its factorisation from a HIP graph per mesh).  Returns one
    SweepResult per problem (Solver(mesh, stack).sweep(freqs) for each)."""
    solvers = [Solver(m, st, verbose=False, precision=precision, graphs=graphs, **kw) for m, st, _ in problems]
    jobs = [(s, np.atleast_1d(np.asarray(f, float))) for s, (_, _, f) in zip(solvers, problems)]
    npts = sum(len(f) for _, f in jobs)
    if lanes is None:
        lanes = min(s._lane_count() for s in solvers)
    if workers is None:
        workers = solvers[0]._worker_count(npts)
    return _sweep_many(jobs, max(1, min(lanes, npts)), workers, keep_currents, graphs)
