"""The surrogate's MLP kernels against the oracle's MLP (forward in BF16 and
FP32, the loss gradient), and a short training run on the GPU."""

import numpy as np
import pytest
import torch

from mom25d_amd import workloads as W
from mom25d_amd.surrogate import RemainderSurrogate
from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _model(seed=0):
    lay, st, mk, freqs = W.via2()
    return RemainderSurrogate(st, ["inner", "top"], (freqs[0], freqs[-1]), seed=seed)


def _batch(n, seed=3):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, M.DIN)).astype(np.float32)
    Y = (0.5 * rng.standard_normal((n, M.DOUT))).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("n", [1, 63, 200])
def test_forward_matches_oracle(n):
    m = _model()
    X, _ = _batch(n)
    Xd = torch.from_numpy(X).cuda()
    P = m.P.cpu().numpy().astype(np.float64)
    out32 = m.forward(Xd, 32).cpu().numpy()
    ref32, _ = M.forward(P, X, emulate=False)
    e32 = np.max(np.abs(out32 - ref32)) / np.max(np.abs(ref32))
    out16 = m.forward(Xd, 16).cpu().numpy()
    ref16, _ = M.forward(P, X, emulate=True)
    e16 = np.max(np.abs(out16 - ref16)) / np.max(np.abs(ref16))
    print(f"n={n}: fp32 rel err {e32:.2e}, bf16 (emulated oracle) rel err {e16:.2e}")
    assert e32 < 1e-5
    assert e16 < 2e-3


@pytest.mark.parametrize("n", [64, 300])
def test_gradient_matches_oracle(n):
    m = _model(seed=1)
    X, Y = _batch(n, seed=n)
    w = np.array([1.0, 0.5, 2.0, 1.5])
    loss, g = m.loss_grad(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda(), w, scale=2.0)
    P = m.P.cpu().numpy().astype(np.float64)
    lref, gref = M.loss_grad(P, X, Y, w, scale=2.0, emulate=True)
    _, gexact = M.loss_grad(P, X, Y, w, scale=2.0, emulate=False)
    assert abs(loss - lref) <= 1e-3 * lref
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    for (name, _), a, b in zip(M.SIZES, offs[:-1], offs[1:]):
        e = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        ex = np.linalg.norm(g[a:b] - gexact[a:b]) / np.linalg.norm(gexact[a:b])
        print(f"n={n} {name}: rel err vs emulated {e:.1e}, vs exact {ex:.1e}")
        assert e < 1e-2, name
        assert ex < 5e-2, name


def test_training_reduces_error_against_integrals():
    """A short run (2^18 samples, 12 epochs of 1024-sample batches) on the
    via-transition stack: the mean squared error over the standardised
    outputs must fall 30-fold, and the FP32 network must reproduce held-out
    integrals to a few percent of the largest remainder (the full-length run
    of the committed model is far tighter)."""
    m = _model()
    rec = m.samples(1 << 18, seed=5, n_freq=128)
    m.set_scaling(rec)
    X, Y = m.tensors(rec)
    hist = m.fit(X, Y, epochs=12, batch=1024, lr=2e-3)
    print("mse per epoch", " ".join(f"{h:.2e}" for h in hist))
    assert hist[-1] < hist[0] / 30
    test = m.samples(4096, seed=99, n_freq=16)
    RA, RV = m.remainders(test["rho"], test["zi"], test["zj"], test["k0"], [test["kxA"], test["kxV"]])
    s = 4 * np.pi * np.sqrt(test["rho"] ** 2 + m.h ** 2)
    eA = np.max(np.abs(RA - test["RA"]) * s)
    eV = np.max(np.abs(RV - test["RV"]) * s)
    print(f"held-out max |dT_A| {eA:.2e}, |dT_V| {eV:.2e}")
    assert eA < 0.05 and eV < 0.05
