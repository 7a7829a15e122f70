"""Triangle quadrature rules.

Dunavant degree-5 7-point rule and the interior 3-point rule, as used by
Engine/mom25d/integrals.py:5-21 (the closed-form 1/R integrals of the near
tier are evaluated on the device, csrc/fill.hip).  The numeric literals of the 7-point rule are
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
