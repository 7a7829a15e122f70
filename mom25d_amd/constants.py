"""Physical constants (SI), as used throughout the solve (Engine/mom25d/constants.py:3-6)."""

import math

C0 = 299792458.0
MU0 = 4e-7 * math.pi
EPS0 = 1.0 / (MU0 * C0 * C0)
ETA0 = MU0 * C0
