"""Lumped components across delta gaps (the component API of
Engine/mom25d/lumped.py).

A lumped gap is a short copper bridge between two pad points
(`Layout.add_lumped`) cut by a delta gap in its middle.  A load of impedance
Z across the gap adds Z l l^T to the impedance matrix, l being the gap
edges' signed lengths; the solver solves once for the feed and gap
excitations and then applies every load at once (Woodbury), so values can
change per sweep (`Solver.sweep(..., values=...)`).

Values are one-port impedances Z(f) [ohm]: `R`, `L`, `C`, their `series` /
`parallel` combinations, plain numbers, callables f -> Z, or text parsed by
`parse` ('C 1p', 'R 10 + L 0.5n', 'C 1p || R 1k', or a path to the .s2p file
of a series-mounted part).  `Touchstone(path)` is an N-port component across
N gaps (`Layout.add_component`).
"""

import numpy as np

# engineering suffixes of the text form
_SCALE = {"f": 1e-15, "p": 1e-12, "n": 1e-9, "u": 1e-6, "m": 1e-3, "k": 1e3, "M": 1e6, "G": 1e9}


class Value:
    """A one-port impedance: z(f) [ohm] at frequency f [Hz]; repr gives a
    readable description."""

    def __init__(self, fz, desc):
        self.fz, self.desc = fz, desc

    def z(self, f):
        return complex(self.fz(f))

    def __repr__(self):
        return self.desc


def _engineering(x, unit):
    """x with the largest SI prefix that keeps the mantissa >= 1."""
    for p in "GMk_munpf":
        scale = 1.0 if p == "_" else _SCALE[p]
        if abs(x) >= scale:
            return f"{x / scale:g} {'' if p == '_' else p}{unit}"
    return f"{x:g} {unit}"


def R(r):
    """Resistor [ohm]."""
    r = float(r)
    return Value(lambda f: complex(r, 0.0), _engineering(r, "ohm"))


def L(l):
    """Inductor [H]: j omega L."""
    l = float(l)
    return Value(lambda f: complex(0.0, 2 * np.pi * f * l), _engineering(l, "H"))


def C(c):
    """Capacitor [F]: 1 / (j omega C)."""
    c = float(c)
    return Value(lambda f: complex(0.0, -1.0 / (2 * np.pi * f * c)), _engineering(c, "F"))


def as_value(v):
    """A Value from a Value, a real or complex number [ohm] or a callable."""
    if isinstance(v, Value):
        return v
    if callable(v):
        return Value(v, getattr(v, "__name__", "Z(f)"))
    z = complex(v)
    if z.imag == 0:
        return R(z.real)
    return Value(lambda f: z, f"{z} ohm")


def series(*vals):
    """Impedances in series: the sum."""
    parts = [as_value(v) for v in vals]
    return Value(lambda f: sum(p.z(f) for p in parts), "(" + " + ".join(map(repr, parts)) + ")")


def parallel(*vals):
    """Impedances in parallel: the inverse of the summed admittances."""
    parts = [as_value(v) for v in vals]
    return Value(lambda f: 1.0 / sum(1.0 / p.z(f) for p in parts), "(" + " || ".join(map(repr, parts)) + ")")


def parse(text):
    """Value of a text: elements 'R 50', 'C 1p', 'L 2.2n' (engineering
    suffixes f p n u m k M G), joined by '+' (series) and '||' (parallel;
    binds weaker than '+'); a path ending in .s2p is a series-mounted part
    (series_from_s2p)."""
    src = text.strip()
    if src.lower().endswith(".s2p"):
        return series_from_s2p(src)
    return _Parser(src).parse()


class _Parser:
    """Recursive descent over  expr := chain ('||' chain)* ;
    chain := element ('+' element)* ;  element := KIND AMOUNT."""

    def __init__(self, src):
        self.src = src
        self.tokens = src.replace("||", " || ").replace("+", " + ").split()
        self.pos = 0

    def _next(self):
        if self.pos >= len(self.tokens):
            raise ValueError(f"{self.src!r}: incomplete value")
        tok = self.tokens[self.pos]
        self.pos += 1
        return tok

    def parse(self):
        branches = [self._chain()]
        while self.pos < len(self.tokens) and self.tokens[self.pos] == "||":
            self.pos += 1
            branches.append(self._chain())
        if self.pos != len(self.tokens):
            raise ValueError(f"{self.src!r}: unexpected {self.tokens[self.pos]!r}")
        return branches[0] if len(branches) == 1 else parallel(*branches)

    def _chain(self):
        parts = [self._element()]
        while self.pos < len(self.tokens) and self.tokens[self.pos] == "+":
            self.pos += 1
            parts.append(self._element())
        return parts[0] if len(parts) == 1 else series(*parts)

    def _element(self):
        kind = self._next().upper()
        make = {"R": R, "L": L, "C": C}.get(kind)
        if make is None:
            raise ValueError(f"{self.src!r}: element {kind!r} is not R, L or C")
        amount = self._next()
        scale = _SCALE.get(amount[-1])
        return make(float(amount[:-1]) * scale if scale is not None else float(amount))


def _read_touchstone(path):
    """(frequencies [Hz], S (F x n x n), reference resistance) of a
    Touchstone v1 file: '!' comments, one '#' option line (frequency unit,
    MA / DB / RI, R z0), then records of f and n^2 complex numbers; the
    records of a 2-port list S11 S21 S12 S22 (column order)."""
    ext = path.rsplit(".", 1)[-1].lower()
    if not (ext.startswith("s") and ext.endswith("p") and ext[1:-1].isdigit()):
        raise ValueError(f"{path}: not a .sNp file name")
    n = int(ext[1:-1])
    unit, fmt, z0 = 1e9, "MA", 50.0
    values = []
    with open(path) as fh:
        for raw in fh:
            line = raw.partition("!")[0].strip()
            if not line:
                continue
            if line[0] == "#":
                opts = line[1:].upper().split()
                for i, o in enumerate(opts):
                    unit = {"HZ": 1.0, "KHZ": 1e3, "MHZ": 1e6, "GHZ": 1e9}.get(o, unit)
                    fmt = o if o in ("MA", "DB", "RI") else fmt
                    if o == "R" and i + 1 < len(opts):
                        z0 = float(opts[i + 1])
                continue
            values += line.split()
    width = 1 + 2 * n * n
    if len(values) % width:
        raise ValueError(f"{path}: {len(values)} numbers do not make whole {n}-port records")
    rec = np.asarray(values, float).reshape(-1, width)
    first, second = rec[:, 1::2], rec[:, 2::2]
    if fmt == "RI":
        s = first + 1j * second
    else:
        mag = first if fmt == "MA" else 10.0 ** (first / 20.0)
        s = mag * np.exp(1j * np.deg2rad(second))
    s = s.reshape(-1, n, n)
    if n == 2:
        s = s.transpose(0, 2, 1)
    return rec[:, 0] * unit, s, z0


class Touchstone:
    """N-port component from a Touchstone file (.sNp).  s(f) interpolates
    the real and imaginary parts of S linearly between the file's
    frequencies (a frequency outside them raises); z(f) = z0 (I + S)(I - S)^-1."""

    def __init__(self, path):
        self.path = path
        self.f, self.s_data, self.z0 = _read_touchstone(path)
        self.n = self.s_data.shape[1]
        self.desc = f"Touchstone {path.rsplit('/', 1)[-1]} ({self.n}-port)"

    def s(self, f):
        lo, hi = self.f[0], self.f[-1]
        if f < lo * (1 - 1e-9) or f > hi * (1 + 1e-9):
            raise ValueError(f"{self.path}: {f / 1e9:g} GHz is outside the file's {lo / 1e9:g}-{hi / 1e9:g} GHz")
        flat = self.s_data.reshape(len(self.f), -1)
        re = [np.interp(f, self.f, col) for col in flat.real.T]
        im = [np.interp(f, self.f, col) for col in flat.imag.T]
        return (np.array(re) + 1j * np.array(im)).reshape(self.n, self.n)

    def z(self, f):
        s = self.s(f)
        eye = np.eye(self.n)
        return self.z0 * np.linalg.solve((eye - s).T, (eye + s).T).T

    def __repr__(self):
        return self.desc


def series_from_s2p(path, extrapolate=True, fit_from=None):
    """One-port Value of a two-terminal part measured in series between the
    ports of a 2-port: from S21 = 2 z0 / (2 z0 + Z), Z = 2 z0 (1 - S21) / S21,
    within the file's band.  With extrapolate, frequencies outside the band
    take a series model R(f) + j (omega L - 1 / (omega C)) with R linear in f:
    L and 1/C by least squares on the reactance relative to |Z| (columns
    scaled to unit norm), R by a straight line, both over the file's points
    at f >= fit_from (default: the upper half of the band); the model's
    (R0, R1, L, C) is kept in `.fit`.  extrapolate=False: frequencies
    outside the band raise."""
    part = Touchstone(path)
    name = path.rsplit("/", 1)[-1]

    def measured(f):
        s21 = part.s(f)[1, 0]
        return 2 * part.z0 * (1 - s21) / s21

    if not extrapolate:
        return Value(measured, f"series {name}")
    f_lo, f_hi = part.f[0], part.f[-1]
    fit_from = 0.5 * (f_lo + f_hi) if fit_from is None else fit_from
    fs = part.f[part.f >= fit_from]
    zs = np.array([measured(x) for x in fs])
    w = 2 * np.pi * fs
    weight = 1.0 / np.maximum(np.abs(zs), 1e-12)
    cols = np.stack([w, -1.0 / w], 1) * weight[:, None]
    norm = np.linalg.norm(cols, axis=0)
    sol = np.linalg.lstsq(cols / norm, zs.imag * weight, rcond=None)[0] / norm
    ind, cap = sol[0], 1.0 / sol[1]
    r1, r0 = np.polyfit(fs, zs.real, 1)

    def model(f):
        if f_lo <= f <= f_hi:
            return measured(f)
        wf = 2 * np.pi * f
        return complex(r0 + r1 * f, wf * ind - 1.0 / (wf * cap))

    v = Value(model, f"series {name} (outside {f_lo / 1e9:g}-{f_hi / 1e9:g} GHz: "
                     f"{_engineering(ind, 'H')} + {_engineering(cap, 'F')} + R(f))")
    v.fit = (r0, r1, ind, cap)
    return v
