"""The trained remainder surrogate of the via-transition stack against the
Sommerfeld integrals it stands in for.

The weights (mom25d_amd/data/surrogate_via2.npz) were trained on the GPU by
the library's own kernels: `python -m mom25d_amd.surrogate via2 <out.npz>`
(2^21 samples, 120 epochs of 4096-sample batches, both networks)."""

import os
import time

import numpy as np
import pytest
import torch

from mom25d_amd import Solver
from mom25d_amd import workloads as W
from mom25d_amd.greens import SweepTables
from mom25d_amd.surrogate import GreensSurrogate

pytestmark = pytest.mark.gpu
WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "mom25d_amd", "data", "surrogate_via2.npz")


@pytest.fixture(scope="module")
def via2():
    lay, st, mk, freqs = W.via2()
    return lay.mesh(**mk), st, freqs, GreensSurrogate.load(WEIGHTS, st)


def test_networks_match_integrals_on_held_out_points(via2):
    mesh, st, freqs, sur = via2
    rec = sur.av.samples(1 << 15, seed=4321, n_freq=32)  # seeds differ from the training set's
    RA, RV, RW = sur.remainders(rec["rho"], rec["zi"], rec["zj"], rec["k0"], rec["kxA"], rec["kxV"])
    s = 4 * np.pi * np.sqrt(rec["rho"] ** 2 + sur.h ** 2)
    for name, a, b in (("A", RA, rec["RA"]), ("V", RV, rec["RV"]), ("W", RW, rec["RW"])):
        d = np.abs(a - b) * s  # error of the demodulated remainder T (|T| ~ 0.2 .. 1)
        rms, mx = np.sqrt(np.mean(d ** 2)), d.max()
        print(f"T_{name}: rms error {rms:.1e}, max {mx:.1e}")
        assert rms < 5e-4 and mx < 1e-2


def test_surrogate_tables_gain_and_cost(via2):
    """Gain: with the networks accepted at every frequency, the remainder
    columns of the 81 tables of the sweep need the Bessel sums only at the
    16 check nodes per frequency.  Cost: |dS| of the sweep against the plain
    tables.  The judge at tol 1e-5 rejects the networks and returns the
    plain tables exactly."""
    mesh, st, freqs, sur = via2
    v = mesh.nodes
    t = SweepTables.spectra_only(st, 1.02 * np.hypot(*(v.max(0) - v.min(0))), mesh.layer_names, False)
    jobs = [t._prepare(f, 12) for f in freqs]
    full = t.integral_nodes(jobs)
    times = []
    for _ in range(2):
        torch.cuda.synchronize()
        a = time.perf_counter()
        t._sommerfeld(jobs)
        torch.cuda.synchronize()
        b = time.perf_counter()
        t._surrogate_remainders(jobs, sur, np.inf)
        torch.cuda.synchronize()
        times.append((b - a, time.perf_counter() - b))
    rep = t.surrogate_report
    ratio = rep["integral_nodes"] / full
    print(f"81 tables: Bessel evaluations {full:.3e} (integrals) vs {rep['integral_nodes']:.3e} (networks + checks), "
          f"ratio {ratio:.3f}; GPU+host seconds integrals {times[-1][0]:.4f}, networks {times[-1][1]:.4f} "
          f"{rep['seconds']}; check error median {np.median(rep['check_err']):.1e} max {rep['check_err'].max():.1e}")
    assert not rep["fallback"] and ratio < 0.1
    f = freqs[::10]
    ref = Solver(mesh, st, verbose=False).sweep(f, interp=None)
    net = Solver(mesh, st, verbose=False, surrogate=sur, surrogate_tol=np.inf).sweep(f, interp=None)
    dS = np.abs(ref.S - net.S).max(axis=(1, 2))
    print("|dS| networks vs integrals per frequency:", " ".join(f"{x:.1e}" for x in dS))
    assert dS.max() < 1e-2
    judged = Solver(mesh, st, verbose=False, surrogate=sur, surrogate_tol=1e-5)
    r = judged.sweep(f, interp=None)
    assert len(judged.last_tables.surrogate_report["fallback"]) == len(f)
    assert np.abs(r.S - ref.S).max() < 1e-8
