"""End-to-end sweeps on the GPU (fill + LU + de-embedding) against the
oracle and the reference's golden S: |dS| < 1e-8 for the FP64 path and
< 1e-6 for the mixed-precision path (complex64 LU + FP64 refinement)."""

import numpy as np
import pytest

from mom25d_amd import Solver
from mom25d_amd import workloads as W
from oracle.solve_ref import OracleSolver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-6)])
@pytest.mark.parametrize("name", ["tiny", "tee"])
def test_sweep_matches_reference(name, precision, tol, golden):
    g = golden("sweeps.npz")
    lay, sub, mk, _ = getattr(W, name)()
    r = Solver(lay.mesh(**mk), sub, verbose=False, precision=precision).sweep(g[f"{name}_freqs"])
    dS = np.max(np.abs(r.S - g[f"{name}_S"]))
    print(f"{name} {precision} |dS| = {dS:.2e}")
    assert dS < tol
    np.testing.assert_allclose(r.z0_line, g[f"{name}_z0_line"], rtol=100 * tol)


def test_via_transition_sweep_matches_oracle():
    lay, st, mk, _ = W.via2()
    m = lay.mesh(**mk)
    f = [3e9, 11e9]
    r = Solver(m, st, verbose=False).sweep(f)
    o = OracleSolver(m, st, use_c=True).sweep(f)
    assert np.max(np.abs(r.S - o["S"])) < 1e-8
