"""The oracle's MLP (forward, loss gradient, BF16 rounding) pinned by
central finite differences and by exact BF16 rounding cases."""

import numpy as np

from oracle import mlp_ref as M


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(M.NPARAM) * scale / np.sqrt(np.repeat(
        [M.DIN, 1, M.DH, 1, M.DH, 1, M.DH, 1],
        [M.DH * M.DIN, M.DH, M.DH * M.DH, M.DH, M.DH * M.DH, M.DH, M.DOUT * M.DH, M.DOUT]))


def test_exact_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    P = _params()
    X = rng.uniform(-1, 1, (5, M.DIN))
    Y = rng.standard_normal((5, M.DOUT))
    w = np.array([1.0, 0.5, 2.0, 1.0])
    _, g = M.loss_grad(P, X, Y, w, emulate=False)
    # a spread of coordinates in every parameter block
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in M.SIZES])
    idx = np.concatenate([rng.integers(a, b, 6) for a, b in zip(offs[:-1], offs[1:])])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (M.loss_grad(Pp, X, Y, w, emulate=False)[0] - M.loss_grad(Pm, X, Y, w, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_round_to_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    cases = {one: one, one + ulp / 2: one, one + 3 * ulp / 2: one + 2 * ulp, one + ulp / 2 + 2 ** -20: one + ulp,
             -3.0: -3.0, 1e-3: 131 / 128 * 2.0 ** -10}
    for x, want in cases.items():
        assert M.bf16(np.array([x]))[0] == want, x


def test_emulated_forward_is_close_to_exact():
    rng = np.random.default_rng(2)
    P = _params()
    X = rng.uniform(-1, 1, (64, M.DIN))
    a, _ = M.forward(P, X, emulate=True)
    b, _ = M.forward(P, X, emulate=False)
    assert np.max(np.abs(a - b)) < 3e-2 * np.max(np.abs(b))
    assert np.max(np.abs(a - b)) > 0
