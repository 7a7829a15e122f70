"""CPU restatement of the remainder-surrogate MLP (test infrastructure only).

The network is 6 -> 128 -> 128 -> 128 -> 4 with SiLU after the hidden
layers; the flat parameter vector is W1[128][6] b1 W2[128][128] b2 W3 b3
W4[4][128] b4.  It has no counterpart in the reference, whose remainder comes
from the Sommerfeld integrals only (Engine/mom25d/layered.py:333-377).

`forward` / `loss_grad` with emulate=True round to BF16 (nearest even) at the
points where the GPU kernels do (hidden pre-activations z, the MFMA operands
silu(z) and the BF16 weights W2 / W3, and the back-propagated dz), and
accumulate in float64 otherwise, so that the GPU gradient can be compared at
FP32 summation-order level.  emulate=False is the exact network in float64,
the one central finite differences pin (tests/test_oracle_mlp.py).

Loss: sum over samples and outputs of wout[o] (out - y)^2 (samples beyond
the valid count excluded); d loss / d out = scale * wout * (out - y), with
scale = 2 for the exact gradient of that sum.
"""

import numpy as np

DIN, DH, DOUT = 6, 128, 4
SIZES = (("W1", (DH, DIN)), ("b1", (DH,)), ("W2", (DH, DH)), ("b2", (DH,)), ("W3", (DH, DH)), ("b3", (DH,)),
         ("W4", (DOUT, DH)), ("b4", (DOUT,)))
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(P):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(P[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(x):
    """Round to BF16 (nearest even) through float32; returns float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, X, emulate=True):
    """Outputs (n, 4) and the cache of the backward pass."""
    p = unpack(P)
    r = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    X = np.asarray(X, np.float64)
    W2, W3 = r(p["W2"]), r(p["W3"])
    z1 = r(X @ p["W1"].T + p["b1"])
    h1 = r(silu(z1))
    z2 = r(h1 @ W2.T + p["b2"])
    h2 = r(silu(z2))
    z3 = r(h2 @ W3.T + p["b3"])
    h3 = silu(z3)
    out = h3 @ p["W4"].T + p["b4"]
    return out, dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, W2=W2, W3=W3, p=p, r=r)


def loss_grad(P, X, Y, wout, scale=2.0, emulate=True):
    """(loss, flat gradient) of the weighted squared error over the batch."""
    out, c = forward(P, X, emulate)
    r, p = c["r"], c["p"]
    wout = np.asarray(wout, np.float64)
    d = out - np.asarray(Y, np.float64)
    loss = float(np.sum(wout * d * d))
    dout = scale * wout * d
    g = {"W4": dout.T @ c["h3"], "b4": dout.sum(0)}
    dz3 = r((dout @ p["W4"]) * dsilu(c["z3"]))
    g["W3"], g["b3"] = dz3.T @ c["h2"], dz3.sum(0)
    dz2 = r((dz3 @ c["W3"]) * dsilu(c["z2"]))
    g["W2"], g["b2"] = dz2.T @ c["h1"], dz2.sum(0)
    dz1 = r((dz2 @ c["W2"]) * dsilu(c["z1"]))
    g["W1"], g["b1"] = dz1.T @ c["X"], dz1.sum(0)
    return loss, pack(g)
