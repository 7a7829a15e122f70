"""Lumped components across delta gaps (Engine/mom25d/lumped.py).

A component sits across a lumped gap: a short copper bridge between two pad
points (`Layout.add_lumped`) with a delta gap in its middle.  A load Z across
the gap adds Z l l^T to the impedance matrix (l: the signed lengths of the
gap's edges); the solver applies all loads after one solve for the feed and
gap excitations (Woodbury), so the values can change per sweep
(`Solver.sweep(..., values=...)`).

One-port values: `R`, `L`, `C`, `series`, `parallel`, a number [ohm], a
callable f -> Z, or a string such as 'C 1p', 'R 10 + L 0.5n',
'C 1p || R 1k' or a path to a series-mounted part's .s2p.  Multi-port:
`Touchstone(path)` across several gaps (`Layout.add_component`).  The file
and text formats, and the way S becomes Z, are the reference's.
"""

import numpy as np

_SI_PREFIX = {"f": 1e-15, "p": 1e-12, "n": 1e-9, "u": 1e-6, "m": 1e-3, "k": 1e3, "M": 1e6, "G": 1e9}


class Value:
    """A one-port impedance Z(f) [ohm] with a readable description."""

    def __init__(self, fz, desc):
        self.fz, self.desc = fz, desc

    def z(self, f):
        return complex(self.fz(f))

    def __repr__(self):
        return self.desc


def _describe(x, unit):
    for prefix, scale in (("G", 1e9), ("M", 1e6), ("k", 1e3), ("", 1.0), ("m", 1e-3), ("u", 1e-6), ("n", 1e-9),
                          ("p", 1e-12), ("f", 1e-15)):
        if abs(x) >= scale:
            return f"{x / scale:g} {prefix}{unit}"
    return f"{x:g} {unit}"


def R(r):
    """Resistor [ohm]."""
    return Value(lambda f: complex(r), _describe(r, "ohm"))


def L(l):
    """Inductor [H]: Z = j w L."""
    return Value(lambda f: 2j * np.pi * f * l, _describe(l, "H"))


def C(c):
    """Capacitor [F]: Z = 1 / (j w C)."""
    return Value(lambda f: 1 / (2j * np.pi * f * c), _describe(c, "F"))


def as_value(v):
    """A Value from a Value, a number [ohm] or a callable f -> Z."""
    if isinstance(v, Value):
        return v
    if callable(v):
        return Value(v, getattr(v, "__name__", "Z(f)"))
    if np.isreal(v):
        return R(float(np.real(v)))
    zc = complex(v)
    return Value(lambda f: zc, f"{zc} ohm")


def series(*vals):
    parts = [as_value(v) for v in vals]
    return Value(lambda f: sum(p.z(f) for p in parts), "(" + " + ".join(repr(p) for p in parts) + ")")


def parallel(*vals):
    parts = [as_value(v) for v in vals]
    return Value(lambda f: 1 / sum(1 / p.z(f) for p in parts), "(" + " || ".join(repr(p) for p in parts) + ")")


def _number(text):
    text = text.strip()
    scale = _SI_PREFIX.get(text[-1:], None) if text else None
    return float(text[:-1]) * scale if scale is not None else float(text)


def _element(text):
    kind, amount = text.split()
    make = {"R": R, "L": L, "C": C}.get(kind.upper())
    if make is None:
        raise ValueError(f"unknown element {kind!r} (R, L or C)")
    return make(_number(amount))


def parse(text):
    """'R 50', 'C 1p', 'L 2.2n', 'R 10 + L 0.5n', 'C 1p || R 1k' -> Value; a
    path ending in .s2p: series_from_s2p.  '||' binds weaker than '+'."""
    t = text.strip()
    if t.lower().endswith(".s2p"):
        return series_from_s2p(t)
    if "||" in t:
        return parallel(*[parse(p) for p in t.split("||")])
    if "+" in t:
        return series(*[_element(p) for p in t.split("+")])
    return _element(t)


class Touchstone:
    """N-port component from a Touchstone v1 file (.sNp; MA / DB / RI data,
    any frequency unit, reference resistance from the option line).  z(f)
    gives the N x N impedance matrix, with S interpolated linearly (real
    and imaginary parts) between the file's frequencies.  A 2-port file lists
    S11 S21 S12 S22 per frequency."""

    _UNITS = {"HZ": 1.0, "KHZ": 1e3, "MHZ": 1e6, "GHZ": 1e9}

    def __init__(self, path):
        self.path = path
        digits = "".join(ch for ch in path.rsplit(".", 1)[-1] if ch.isdigit())
        if not digits:
            raise ValueError(f"{path}: not a .sNp file name")
        n = int(digits)
        fmt, unit, z0 = "MA", 1e9, 50.0
        numbers = []
        with open(path) as fh:
            for line in fh:
                line = line.split("!", 1)[0].strip()
                if not line:
                    continue
                if line.startswith("#"):
                    opts = line[1:].upper().split()
                    for i, o in enumerate(opts):
                        if o in self._UNITS:
                            unit = self._UNITS[o]
                        elif o in ("MA", "DB", "RI"):
                            fmt = o
                        elif o == "R" and i + 1 < len(opts):
                            z0 = float(opts[i + 1])
                    continue
                numbers.extend(float(x) for x in line.split())
        per_freq = 1 + 2 * n * n
        if len(numbers) % per_freq:
            raise ValueError(f"{path}: {len(numbers)} numbers do not make whole {n}-port records")
        data = np.asarray(numbers).reshape(-1, per_freq)
        self.f = data[:, 0] * unit
        u, v = data[:, 1::2], data[:, 2::2]
        if fmt == "RI":
            s = u + 1j * v
        elif fmt == "MA":
            s = u * np.exp(1j * np.deg2rad(v))
        else:
            s = 10.0 ** (u / 20.0) * np.exp(1j * np.deg2rad(v))
        s = s.reshape(-1, n, n)
        if n == 2:  # S11 S21 S12 S22: column-major per record
            s = np.swapaxes(s, 1, 2)
        self.s_data, self.n, self.z0 = s, n, z0
        self.desc = f"Touchstone {path.rsplit('/', 1)[-1]} ({n}-port)"

    def s(self, f):
        """S at f (n x n)."""
        if not self.f[0] * (1 - 1e-9) <= f <= self.f[-1] * (1 + 1e-9):
            raise ValueError(f"{self.path}: {f / 1e9:g} GHz outside {self.f[0] / 1e9:g}-{self.f[-1] / 1e9:g} GHz")
        flat = self.s_data.reshape(len(self.f), -1)
        out = np.array([np.interp(f, self.f, flat[:, k].real) + 1j * np.interp(f, self.f, flat[:, k].imag)
                        for k in range(flat.shape[1])])
        return out.reshape(self.n, self.n)

    def z(self, f):
        """Z = z0 (I + S)(I - S)^-1."""
        s = self.s(f)
        eye = np.eye(self.n)
        return self.z0 * (eye + s) @ np.linalg.inv(eye - s)

    def __repr__(self):
        return self.desc


def series_from_s2p(path, extrapolate=True, fit_from=None):
    """One-port Value of a series-mounted two-terminal part measured as a
    2-port: Z = 2 z0 (1 - S21) / S21 within the file's band.  Outside it
    (extrapolate=True) a series R(f) + L + C fitted to the data at
    f >= fit_from (default f_max / 6): X = w L - 1 / (w C) by least squares on
    the relative error of X, R linear in f; the fit (R0, R1, L, C) is kept in
    `.fit`.  extrapolate=False: frequencies outside the band raise."""
    t = Touchstone(path)
    name = path.rsplit("/", 1)[-1]

    def z_band(f):
        s21 = t.s(f)[1, 0]
        return 2 * t.z0 * (1 - s21) / s21

    if not extrapolate:
        return Value(z_band, f"series {name}")
    fs = t.f[t.f >= (t.f[-1] / 6 if fit_from is None else fit_from)]
    zs = np.array([z_band(x) for x in fs])
    w = 2 * np.pi * fs
    weight = 1 / np.maximum(np.abs(zs.imag), 1.0)
    # columns scaled to O(1) (w in Grad/s): X = (L 1e9)(w / 1e9) - (1e-9 / C)(1e9 / w)
    coef = np.linalg.lstsq(np.c_[w / 1e9, -1e9 / w] * weight[:, None], zs.imag * weight, rcond=None)[0]
    ind, cap = coef[0] * 1e-9, 1 / (coef[1] * 1e9)
    r1, r0 = np.polyfit(fs, zs.real, 1)

    def z_all(f):
        if t.f[0] <= f <= t.f[-1]:
            return z_band(f)
        wf = 2 * np.pi * f
        return complex(r0 + r1 * f, wf * ind - 1 / (wf * cap))

    v = Value(z_all, f"series {name} (outside {t.f[0] / 1e9:g}-{t.f[-1] / 1e9:g} GHz: R + {ind * 1e9:.3f} nH "
                     f"+ {cap * 1e12:.4f} pF fit)")
    v.fit = (r0, r1, ind, cap)
    return v
