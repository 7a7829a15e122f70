"""Oracle (test infrastructure only): numpy restatement of the reference's
frequency interpolation of Z, Engine/mom25d/freqinterp.py:62-139.  The product
(mom25d_amd.freqinterp + csrc/interp.hip) never imports it; tests compare the
GPU kernels with it and pin it by closed forms (Lagrange interpolation is exact
on polynomials of degree <= order; the phase round trip is the identity)."""

import numpy as np

C0 = 299792458.0


def window(anchors, order, f):
    """The order+1 anchor indices used at f (freqinterp.py:102-109)."""
    n, k = len(anchors), order + 1
    j = int(np.clip(np.searchsorted(anchors, f, side="right") - 1, 0, n - 2))
    i0 = min(max(j - (k - 1) // 2, 0), n - k)
    return list(range(i0, i0 + k))


def lagrange_weights(anchors, idx, f):
    """Lagrange basis polynomials of the anchors idx evaluated at f
    (freqinterp.py:111-118)."""
    fa = np.asarray(anchors)[idx]
    L = np.ones(len(idx))
    for a in range(len(idx)):
        for b in range(len(idx)):
            if a != b:
                L[a] *= (f - fa[b]) / (fa[a] - fa[b])
    return L


def centre_distance(pos, nb):
    """R_mn between basis centres, 0 on the via-via block (freqinterp.py:70-83)."""
    d = pos[:, None, :] - pos[None, :, :]
    R = np.sqrt((d * d).sum(-1))
    R[nb:, nb:] = 0.0
    return R


def anchor_transform(Z, f, pos, nb, k_ratio):
    """W = j w Z exp(+j k_e R) (freqinterp.py:86-100)."""
    w = 2 * np.pi * f
    ke = k_ratio * w / C0
    return 1j * w * Z * np.exp(1j * ke * centre_distance(pos, nb))


def combine(Ws, L, f, pos, nb, k_ratio):
    """Z = (sum L_i W_i) exp(-j k_e R) / (j w) (freqinterp.py:120-139)."""
    w = 2 * np.pi * f
    ke = k_ratio * w / C0
    acc = sum(Li * Wi for Li, Wi in zip(L, Ws))
    return acc * np.exp(-1j * ke * centre_distance(pos, nb)) / (1j * w)
