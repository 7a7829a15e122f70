"""MLP surrogate of the smooth layered-Green remainders R_A, R_V.

The tables of a stack (Engine/mom25d/layered.py:333-405) hold, per pair of
metal levels, S(rho) = c exp(-j kx R) / (4 pi R) + R(rho), where the
remainder R is a Sommerfeld integral (layered.py:360-377) -- a Bessel sum over
thousands of spectral nodes per table node.  The surrogate maps

    (log rho, z, z', k0, eps_r, h)  ->  T_A, T_V (complex, 4 reals)

with an MLP 6 -> 128 -> 128 -> 128 -> 4 (SiLU), trained on the GPU by the
library's own BF16 kernels (mom_mlp_train_step / mom_mlp_adam) on samples
of the Sommerfeld integrals, and evaluated in FP32 (mom_mlp_forward) to
build the A / V columns of the tables without the integrals.  T is the
remainder demodulated by the phase of the leading term and scaled by its
distance law,

    T_m = R_m 4 pi sqrt(rho^2 + h^2) exp(+j kx_m rho),

which is smooth in log rho (R itself oscillates at large k0 rho); outputs
are standardised per channel.  The Sommerfeld integrals stay the judge:
tables built with a surrogate check it at a few nodes per frequency and fall
back to the integrals where it misses the tolerance.
"""

import math

import numpy as np
import torch

from . import _lib
from .constants import C0

DIN, DH, DOUT = 6, 128, 4


def _init_params(rng):
    """He-style uniform initialisation of the flat parameter vector."""
    def U(fan_in, shape):
        b = math.sqrt(6.0 / fan_in) * 0.5
        return rng.uniform(-b, b, shape)

    parts = [U(DIN, (DH, DIN)), np.zeros(DH), U(DH, (DH, DH)), np.zeros(DH), U(DH, (DH, DH)), np.zeros(DH),
             U(DH, (DOUT, DH)) * 0.5, np.zeros(DOUT)]
    return np.concatenate([p.ravel() for p in parts])


class RemainderSurrogate:
    """Network for the remainders of one stack over a frequency band.

    stack   : the Stackup whose tables it serves
    levels  : metal level names (the pairs (z, z') it covers)
    f_range : (f_min, f_max) in Hz
    rho_range : (rho_min, rho_max) in m
    kind    : "AV" -- outputs T_A, T_V (the vector and scalar potential
              kernels of the fill); "W" -- outputs T_W (the line-voltage
              kernel) in the first two outputs, the last two unused (zero
              loss weight).
    """

    def __init__(self, stack, levels, f_range, rho_range=(1e-6, 0.05), device="cuda", seed=0, kind="AV"):
        if kind not in ("AV", "W"):
            raise ValueError("kind must be 'AV' or 'W'")
        self.kind = kind
        self.wout = np.array([1.0, 1.0, 1.0, 1.0]) if kind == "AV" else np.array([1.0, 1.0, 0.0, 0.0])
        self.stack = stack
        self.levels = list(levels)
        self.f_range = (float(f_range[0]), float(f_range[1]))
        self.rho_range = (float(rho_range[0]), float(rho_range[1]))
        self.device = torch.device(device)
        self.h = float(stack.z[-1])
        self.zl = np.array([stack.level_z(n) for n in self.levels])
        self.eps_r = float(np.mean([d.eps_r for d in stack.layers]))
        n = int(_lib.lib().mom_mlp_param_count())
        assert n == DH * DIN + DH + 2 * (DH * DH + DH) + DOUT * DH + DOUT
        self.nparam = n
        self.P = torch.from_numpy(_init_params(np.random.default_rng(seed)).astype(np.float32)).to(self.device)
        self.Wbf = torch.empty(2 * DH * DH, dtype=torch.int16, device=self.device)
        self.mean = np.zeros(DOUT)
        self.std = np.ones(DOUT)
        self.history = []
        self.train_info = {}
        self._pack()

    # ------------------------------------------------------------ mapping
    def features(self, rho, zi, zj, k0):
        """Network inputs (n, 6) float32 in about [-1, 1].  rho below the
        training range is clamped to it (the remainder is flat there: its
        relative change from rho_min to 0 is below 1e-6 for rho_min = 1 um)."""
        lr0, lr1 = np.log(self.rho_range[0]), np.log(self.rho_range[1])
        k_lo, k_hi = (2 * np.pi * f / C0 for f in self.f_range)
        rho = np.asarray(rho, float)
        n = rho.size
        X = np.empty((n, DIN), np.float32)
        X[:, 0] = 2 * (np.log(np.maximum(rho, self.rho_range[0])) - lr0) / (lr1 - lr0) - 1
        X[:, 1] = 2 * np.asarray(zi, float) / self.h - 1
        X[:, 2] = 2 * np.asarray(zj, float) / self.h - 1
        X[:, 3] = 2 * (np.asarray(k0, float) - k_lo) / (k_hi - k_lo) - 1
        X[:, 4] = (self.eps_r - 1.0) / 10.0
        X[:, 5] = self.h / 1e-3 - 1.0
        return X

    def demodulate(self, R, rho, kx):
        """T = R 4 pi sqrt(rho^2 + h^2) exp(+j kx rho)."""
        return R * (4 * np.pi * np.sqrt(rho ** 2 + self.h ** 2)) * np.exp(1j * kx * rho)

    def modulate(self, T, rho, kx):
        return T / (4 * np.pi * np.sqrt(rho ** 2 + self.h ** 2)) * np.exp(-1j * kx * rho)

    def _channels(self, rec):
        if self.kind == "AV":
            return [(rec["RA"], rec["kxA"]), (rec["RV"], rec["kxV"])]
        return [(rec["RW"], rec["kxV"])]

    def _raw_targets(self, rec):
        Y = np.zeros((rec["rho"].size, DOUT))
        for c, (R, kx) in enumerate(self._channels(rec)):
            T = self.demodulate(R, rec["rho"], kx)
            Y[:, 2 * c], Y[:, 2 * c + 1] = T.real, T.imag
        return Y

    def decode(self, Y, rho, kxs):
        """Standardised outputs (n, 4) -> the remainders of this network's
        kernels (list of complex arrays), kxs their leading wavenumbers."""
        Y = np.asarray(Y, float) * self.std + self.mean
        return [self.modulate(Y[:, 2 * c] + 1j * Y[:, 2 * c + 1], rho, kx) for c, kx in enumerate(kxs)]

    # ------------------------------------------------------------ training data
    def samples(self, n, seed=1, n_freq=256):
        """About n samples of the Sommerfeld integrals (GPU): n_freq
        frequencies uniform in the band, log-uniform rho, every level pair.
        Returns the record (k0, rho, zi, zj, RA, RV, RW, kxA, kxV) as arrays."""
        from .greens import sommerfeld_remainders

        rng = np.random.default_rng(seed)
        L = len(self.levels)
        freqs = np.sort(rng.uniform(*self.f_range, n_freq))
        per = max(1, -(-n // (n_freq * L * L)))
        lr0, lr1 = np.log(self.rho_range[0]), np.log(self.rho_range[1])
        rhos = [np.exp(rng.uniform(lr0, lr1, per)) for _ in freqs]
        res, kx = sommerfeld_remainders(self.stack, self.levels, freqs, rhos, self.rho_range[1], self.device)
        raw = []
        for f, rho, R, kf in zip(freqs, rhos, res, kx):
            k0 = 2 * np.pi * f / C0
            m = rho.size
            for i in range(L):
                for j in range(L):
                    raw.append((np.full(m, k0), rho, np.full(m, self.zl[i]), np.full(m, self.zl[j]), R[:, 0, i, j],
                                R[:, 1, i, j], R[:, 2, i, j], np.full(m, kf[0, i, j]), np.full(m, kf[1, i, j])))
        cols = [np.concatenate(c) for c in zip(*raw)]
        return dict(zip(("k0", "rho", "zi", "zj", "RA", "RV", "RW", "kxA", "kxV"), cols))

    def set_scaling(self, rec):
        Y = self._raw_targets(rec)
        self.mean = Y.mean(0)
        self.std = np.maximum(Y.std(0), 1e-12)
        if self.kind == "W":
            self.mean[2:], self.std[2:] = 0.0, 1.0

    def tensors(self, rec):
        X = self.features(rec["rho"], rec["zi"], rec["zj"], rec["k0"])
        Y = ((self._raw_targets(rec) - self.mean) / self.std).astype(np.float32)
        return torch.from_numpy(X).to(self.device), torch.from_numpy(Y).to(self.device)

    # ------------------------------------------------------------ kernels
    def _pack(self):
        _lib.call("mom_mlp_pack", _lib.ptr(self.P), _lib.ptr(self.Wbf), _lib.stream_ptr())

    def loss_grad(self, X, Y, wout=None, scale=2.0):
        """(loss, gradient) of sum wout (out - y)^2 over the batch, on the GPU
        (BF16 training arithmetic); no update."""
        n = X.shape[0]
        nslots = -(-n // 64)
        stride = int(_lib.lib().mom_mlp_partial_stride())
        part = torch.empty(nslots * stride, dtype=torch.float32, device=self.device)
        w = torch.tensor(np.ones(DOUT) if wout is None else wout, dtype=torch.float32, device=self.device)
        grad = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.device)
        P = _lib.ptr
        s = _lib.stream_ptr()
        _lib.call("mom_mlp_train_step", n, P(X), P(Y), P(w), float(scale), P(self.P), P(self.Wbf), P(part), nslots, s)
        _lib.call("mom_mlp_adam", nslots, P(part), P(self.P), None, None, P(self.Wbf), P(grad), P(loss), 0.9, 0.999,
                  1e-8, 0.0, s)
        return float(loss.item()), grad.cpu().numpy()

    def fit(self, X, Y, epochs=20, batch=8192, lr=2e-3, lr_end=2e-5, beta1=0.9, beta2=0.999, eps=1e-8, seed=2,
            log_every=0):
        """Adam over shuffled mini-batches with a cosine learning-rate decay;
        returns the mean loss per epoch (loss = mean squared error over the
        standardised outputs)."""
        n = X.shape[0]
        dev = self.device
        nb = -(-n // batch)
        stride = int(_lib.lib().mom_mlp_partial_stride())
        part = torch.empty(-(-batch // 64) * stride, dtype=torch.float32, device=dev)
        m1 = torch.zeros(self.nparam, dtype=torch.float32, device=dev)
        m2 = torch.zeros_like(m1)
        w = torch.tensor(self.wout, dtype=torch.float32, device=dev)
        losses = torch.zeros(nb, dtype=torch.float32, device=dev)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        P = _lib.ptr
        s = _lib.stream_ptr()
        total = epochs * nb
        t = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=dev, generator=gen)
            for b in range(nb):
                idx = perm[b * batch:(b + 1) * batch]
                xb = X.index_select(0, idx)
                yb = Y.index_select(0, idx)
                m = xb.shape[0]
                nslots = -(-m // 64)
                t += 1
                frac = (t - 1) / max(total - 1, 1)
                lr_t = lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * frac))
                step = lr_t * math.sqrt(1 - beta2 ** t) / (1 - beta1 ** t)
                _lib.call("mom_mlp_train_step", m, P(xb), P(yb), P(w), 2.0 / (m * self.wout.sum()), P(self.P), P(self.Wbf),
                          P(part), nslots, s)
                _lib.call("mom_mlp_adam", nslots, P(part), P(self.P), P(m1), P(m2), P(self.Wbf), None,
                          P(losses[b:b + 1]), beta1, beta2, eps, step, s)
            ep_loss = float(losses.sum().item()) / (n * self.wout.sum())
            self.history.append(ep_loss)
            if log_every and (ep + 1) % log_every == 0:
                print(f"  epoch {ep + 1}/{epochs}: mse {ep_loss:.3e}", flush=True)
        return list(self.history)

    def forward(self, X, precision=32):
        """Standardised outputs (n, 4) on the device: precision 32 (FP32) or
        16 (the BF16 training arithmetic)."""
        n = X.shape[0]
        out = torch.empty((n, DOUT), dtype=torch.float32, device=self.device)
        _lib.call("mom_mlp_forward", precision, n, _lib.ptr(X), _lib.ptr(self.P), _lib.ptr(self.Wbf), _lib.ptr(out),
                  _lib.stream_ptr())
        return out

    def remainders(self, rho, zi, zj, k0, kxs, precision=32):
        """The remainders of this network's kernels (complex128 numpy) at
        the given points; kxs: leading wavenumbers per kernel."""
        X = torch.from_numpy(self.features(rho, zi, zj, k0)).to(self.device)
        Y = self.forward(X, precision).cpu().numpy()
        return self.decode(Y, np.asarray(rho, float), kxs)

    # ------------------------------------------------------------ files
    def state(self, prefix):
        return {f"{prefix}P": self.P.cpu().numpy(), f"{prefix}mean": self.mean, f"{prefix}std": self.std,
                f"{prefix}history": np.array(self.history)}

    def set_state(self, d, prefix):
        self.P.copy_(torch.from_numpy(d[f"{prefix}P"].astype(np.float32)))
        self.mean, self.std = d[f"{prefix}mean"], d[f"{prefix}std"]
        self.history = list(d[f"{prefix}history"])
        self._pack()


class GreensSurrogate:
    """The pair of networks that stands in for all three remainder kernels
    of a stack's tables: `av` (R_A, R_V) and `w` (R_W)."""

    def __init__(self, av, w):
        self.av, self.w = av, w
        self.f_range, self.rho_range = av.f_range, av.rho_range
        self.h = av.h

    def remainders(self, rho, zi, zj, k0, kxA, kxV):
        RA, RV = self.av.remainders(rho, zi, zj, k0, [kxA, kxV])
        (RW,) = self.w.remainders(rho, zi, zj, k0, [kxV])
        return RA, RV, RW

    def demodulate(self, R, rho, kx):
        return self.av.demodulate(R, rho, kx)

    def modulate(self, T, rho, kx):
        return self.av.modulate(T, rho, kx)

    def save(self, path, **info):
        a = self.av
        np.savez(path, f_range=a.f_range, rho_range=a.rho_range, h=a.h, eps_r=a.eps_r, zl=a.zl,
                 levels=np.array(a.levels), **a.state("av_"), **self.w.state("w_"),
                 **{f"info_{k}": v for k, v in info.items()})

    @classmethod
    def load(cls, path, stack, device="cuda"):
        d = np.load(path, allow_pickle=False)
        levels = [str(x) for x in d["levels"]]
        nets = [RemainderSurrogate(stack, levels, tuple(d["f_range"]), tuple(d["rho_range"]), device, kind=k)
                for k in ("AV", "W")]
        m = nets[0]
        if not (np.isclose(float(d["h"]), m.h) and np.allclose(d["zl"], m.zl) and np.isclose(float(d["eps_r"]),
                                                                                           m.eps_r)):
            raise ValueError(f"{path}: surrogate was trained for another stack")
        nets[0].set_state(d, "av_")
        nets[1].set_state(d, "w_")
        return cls(*nets)


def train_for_workload(name, out=None, n_samples=1 << 21, epochs=120, batch=4096, lr=2e-3, seed=0, log_every=0):
    """Train both networks of a workload's stack (its metal levels, its
    frequency band) on one sample set; save them to `out` (npz) if given."""
    import time

    from . import workloads as W

    lay, st, mk, freqs = W.ALL[name]()
    st = st.to_stackup() if hasattr(st, "to_stackup") else st
    levels = sorted(st.metals, key=st.level_z)
    nets = [RemainderSurrogate(st, levels, (freqs[0], freqs[-1]), seed=seed + k, kind=kind)
            for k, kind in enumerate(("AV", "W"))]
    t0 = time.perf_counter()
    rec = nets[0].samples(n_samples, seed=seed + 1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for m in nets:
        m.set_scaling(rec)
        X, Y = m.tensors(rec)
        m.fit(X, Y, epochs=epochs, batch=batch, lr=lr, log_every=log_every)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    sur = GreensSurrogate(*nets)
    sur.train_info = dict(workload=name, samples=int(rec["rho"].size), epochs=epochs, batch=batch, lr=lr,
                          sample_seconds=t1 - t0, train_seconds=t2 - t1, mse_av=nets[0].history[-1],
                          mse_w=nets[1].history[-1])
    if out:
        sur.save(out, **{k: v for k, v in sur.train_info.items() if k != "workload"})
    return sur


__all__ = ["RemainderSurrogate", "GreensSurrogate", "train_for_workload"]


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="train the remainder surrogate of a workload's stack")
    ap.add_argument("workload")
    ap.add_argument("out")
    ap.add_argument("--samples", type=int, default=1 << 21)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    a = ap.parse_args()
    mdl = train_for_workload(a.workload, a.out, a.samples, a.epochs, a.batch, a.lr, log_every=10)
    print(mdl.train_info, flush=True)
