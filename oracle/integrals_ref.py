"""Oracle: triangle rules and the Wilton closed-form potential integrals
(Engine/mom25d/integrals.py:5-67).  Test infrastructure only."""

import numpy as np

# Dunavant (1985) degree-5 rule: centroid weight 9/40, two orbits of three
# points with the published 15-digit coordinates used by the reference.
_ORBIT1 = (0.059715871789770, 0.470142064105115, 0.132394152788506)
_ORBIT2 = (0.797426985353087, 0.101286507323456, 0.125939180544827)


def _rule7():
    bary = [[1 / 3, 1 / 3, 1 / 3]]
    w = [0.225]
    for a, b, wt in (_ORBIT1, _ORBIT2):
        for k in range(3):
            row = [b, b, b]
            row[k] = a
            bary.append(row)
            w.append(wt)
    return np.array(bary), np.array(w)


TRI7_BARY, TRI7_W = _rule7()
TRI3_BARY = np.full((3, 3), 1 / 6) + np.eye(3) * 0.5
TRI3_W = np.full(3, 1 / 3)


def points(verts, bary):
    """(N, 3, 2) vertices, (K, 3) barycentric -> (N, K, 2)."""
    return np.einsum("kv,nvd->nkd", bary, verts)


def wilton(verts, r, d=0.0):
    """I0 (N, K) = int 1/R, I1 (N, K, 2) = int (r' - r)/R over CCW triangles at
    observers r (K, 2) a height d above the plane (edge-by-edge closed form)."""
    d = abs(float(d))
    I0 = np.zeros((len(verts), len(r)))
    I1 = np.zeros((len(verts), len(r), 2))
    for e in range(3):
        a = verts[:, (e + 1) % 3][:, None, :]
        b = verts[:, (e + 2) % 3][:, None, :]
        edge = b - a
        length = np.linalg.norm(edge, axis=-1, keepdims=True)
        tang = edge / length
        normal = np.stack([tang[..., 1], -tang[..., 0]], -1)   # outward for CCW
        da, db = a - r[None], b - r[None]
        la, lb = (da * tang).sum(-1), (db * tang).sum(-1)
        t = (da * normal).sum(-1)
        Ra = np.sqrt((da ** 2).sum(-1) + d * d)
        Rb = np.sqrt((db ** 2).sum(-1) + d * d)
        R02 = t * t + d * d
        with np.errstate(divide="ignore", invalid="ignore"):
            fb = np.where(lb >= 0, Rb + lb, R02 / (Rb - lb))
            fa = np.where(la >= 0, Ra + la, R02 / (Ra - la))
            lg = np.log(fb / fa)
        lg = np.where(np.sqrt(R02) <= 1e-12 * length[..., 0], 0.0, lg)
        I0 += t * lg
        if d:
            I0 -= d * (np.arctan2(t * lb, R02 + d * Rb) - np.arctan2(t * la, R02 + d * Ra))
        I1 += 0.5 * normal * (R02 * lg + lb * Rb - la * Ra)[..., None]
    return I0, I1
