"""Triangle quadrature rules and closed-form 1/R potential integrals.

Dunavant degree-5 7-point rule, the interior 3-point rule and the closed-form
integrals of Wilton et al. (IEEE T-AP 1984) over planar triangles, as used by
Engine/mom25d/integrals.py:5-67.  The numeric literals of the 7-point rule are
the published 15-digit values the reference uses (they fix Z to the last
digits, so they are kept as they are).
"""

import numpy as np

_a1, _b1 = 0.059715871789770, 0.470142064105115
_a2, _b2 = 0.797426985353087, 0.101286507323456
TRI7_BARY = np.array([
    [1 / 3, 1 / 3, 1 / 3],
    [_a1, _b1, _b1], [_b1, _a1, _b1], [_b1, _b1, _a1],
    [_a2, _b2, _b2], [_b2, _a2, _b2], [_b2, _b2, _a2],
])
TRI7_W = np.array([0.225] + [0.132394152788506] * 3 + [0.125939180544827] * 3)

TRI3_BARY = np.array([[2 / 3, 1 / 6, 1 / 6], [1 / 6, 2 / 3, 1 / 6], [1 / 6, 1 / 6, 2 / 3]])
TRI3_W = np.full(3, 1 / 3)


def quad_points(verts, bary):
    """verts (N, 3, 2), barycentric rule (K, 3) -> points (N, K, 2)."""
    return np.einsum("kv,nvd->nkd", bary, verts)


def potential_integrals(verts, r, d=0.0):
    """I0 = int_T 1/R dS' (N, K) and I1 = int_T (r' - r)/R dS' (N, K, 2) over
    counter-clockwise triangles verts (N, 3, 2) at observers r (K, 2) a height
    d above the plane, R = sqrt(|r - r'|^2 + d^2)."""
    d = abs(float(d))
    d2 = d * d
    I0 = np.zeros((verts.shape[0], r.shape[0]))
    I1 = np.zeros((verts.shape[0], r.shape[0], 2))
    for e in range(3):
        pm = verts[:, (e + 1) % 3][:, None, :]
        pp = verts[:, (e + 2) % 3][:, None, :]
        edge = pp - pm
        ln = np.sqrt(np.sum(edge * edge, -1, keepdims=True))
        lhat = edge / ln
        uhat = np.stack([lhat[..., 1], -lhat[..., 0]], -1)
        dm = pm - r[None]
        dp = pp - r[None]
        lm = np.sum(dm * lhat, -1)
        lp = np.sum(dp * lhat, -1)
        t = np.sum(dm * uhat, -1)
        Rm = np.sqrt(dm[..., 0] ** 2 + dm[..., 1] ** 2 + d2)
        Rp = np.sqrt(dp[..., 0] ** 2 + dp[..., 1] ** 2 + d2)
        R02 = t * t + d2
        with np.errstate(divide="ignore", invalid="ignore"):
            fp = np.where(lp >= 0, Rp + lp, R02 / (Rp - lp))
            fm = np.where(lm >= 0, Rm + lm, R02 / (Rm - lm))
            logt = np.log(fp / fm)
        logt = np.where(np.sqrt(R02) <= 1e-12 * ln[..., 0], 0.0, logt)
        I0 = I0 + t * logt
        if d:
            I0 = I0 - d * (np.arctan2(t * lp, R02 + d * Rp) - np.arctan2(t * lm, R02 + d * Rm))
        I1 = I1 + 0.5 * uhat * (R02 * logt + lp * Rp - lm * Rm)[..., None]
    return I0, I1
