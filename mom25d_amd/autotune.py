"""Autotuner of the LU's launch shapes on the device at hand.

For a matrix size and precision it times mom_getrf on a synthetic complex
matrix of that size (random entries plus a dominant diagonal, so that the
pivoting does the same work as on an impedance matrix) for every candidate
panel width, update depth (block width) and trailing-update tile, and keeps
the fastest.  Results are
cached per (precision, size rounded up to a multiple of 512) in memory and,
when MOM25D_TUNE_CACHE names a file, in that JSON file.
"""

import json
import os
import time

import numpy as np
import torch

from . import _lib

PANELS = (32, 64, 128)
BLOCKS = (1, 2, 4)  # update depth nb = panel width x this (up to 256)
TILES = ((64, 64), (64, 128))
_CACHE = {}


def _key(prec, n):
    return f"{prec}:{-(-n // 512) * 512}"


def _load():
    path = os.environ.get("MOM25D_TUNE_CACHE")
    if path and os.path.exists(path) and not _CACHE.get("_loaded"):
        with open(path) as fh:
            _CACHE.update(json.load(fh))
        _CACHE["_loaded"] = True


def _save():
    path = os.environ.get("MOM25D_TUNE_CACHE")
    if path:
        with open(path, "w") as fh:
            json.dump({k: v for k, v in _CACHE.items() if not k.startswith("_")}, fh, indent=1)


def set_tile(prec, tile):
    _lib.call("mom_set_gemm_tile", prec, int(tile[0]), int(tile[1]))


def set_block(prec, nb):
    _lib.call("mom_set_lu_block", prec, int(nb))


def get_block(prec):
    return int(_lib.lib().mom_get_lu_block(prec))


def get_tile(prec):
    import ctypes

    bm, bn = ctypes.c_int(), ctypes.c_int()
    _lib.call("mom_get_gemm_tile", prec, ctypes.byref(bm), ctypes.byref(bn))
    return bm.value, bn.value


def time_getrf(n, prec, jb, tile, device="cuda", reps=2, seed=0, nb=0, pivot=True, symmetric=False):
    """Median milliseconds of mom_getrf (n, prec, panel jb, block nb, tile)
    on a synthetic matrix; pivot=False: mom_getrf_nopiv (the mixed path's
    preconditioner factorisation); symmetric: mom_sytrf_nopiv on a
    complex-symmetric matrix."""
    dev = torch.device(device)
    dt = torch.complex128 if prec == 64 else torch.complex64
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    A0 = torch.complex(torch.randn(n, n, device=dev, generator=g, dtype=torch.float64),
                       torch.randn(n, n, device=dev, generator=g, dtype=torch.float64)).to(dt)
    if symmetric:
        A0 = A0 + A0.T
    A0.diagonal().add_(float(np.sqrt(n)) * 4)
    A = torch.empty_like(A0)
    ipiv = torch.empty(n, dtype=torch.int32, device=dev)
    info = torch.zeros(4, dtype=torch.int32, device=dev)
    wb = int(_lib.lib().mom_getrf_work_bytes(prec, n, jb))
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    old, old_nb = get_tile(prec), get_block(prec)
    set_tile(prec, tile)
    set_block(prec, nb)
    P = _lib.ptr
    s = _lib.stream_ptr()
    times = []
    try:
        for _ in range(reps + 1):
            A.copy_(A0)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            if pivot:
                _lib.call("mom_getrf", prec, n, P(A), n, P(ipiv), P(info), P(work), wb, jb, 0, s)
            else:
                _lib.call("mom_sytrf_nopiv" if symmetric else "mom_getrf_nopiv", prec, n, P(A), n, P(ipiv), P(info),
                          P(work), wb, jb, s)
            e1.record()
            torch.cuda.synchronize(dev)
            times.append(e0.elapsed_time(e1))
    finally:
        set_tile(prec, old)
        set_block(prec, old_nb)
    return float(np.median(times[1:]))


def tune_lu(n, prec, device="cuda", panels=PANELS, blocks=BLOCKS, tiles=TILES, reps=2, verbose=False, pivot=True,
            symmetric=False):
    """The fastest (panel width, block width, tile) for an n x n
    factorisation in prec (64: complex128, 32: complex64), with row
    interchanges (mom_getrf) or without (pivot=False, mom_getrf_nopiv; and
    symmetric: mom_sytrf_nopiv): dict(jb, nb, tile, ms, table)."""
    _load()
    key = _key(prec, n) + ("" if pivot else ":sym" if symmetric else ":nopiv")
    if key in _CACHE:
        return _CACHE[key]
    m = -(-n // 512) * 512
    t0 = time.perf_counter()
    table = []
    for jb in panels:
        for mult in blocks:
            nb = jb * mult
            if nb > 256 or nb > m:
                continue
            for tile in tiles:
                try:
                    ms = time_getrf(m, prec, jb, tile, device, reps, nb=nb, pivot=pivot, symmetric=symmetric)
                except RuntimeError as e:  # e.g. a panel too large for the LDS at this width
                    if verbose:
                        print(f"  autotune n={m} prec={prec} jb={jb} nb={nb} tile={tile}: {e}", flush=True)
                    continue
                table.append((jb, nb, list(tile), ms))
                if verbose:
                    print(f"  autotune n={m} prec={prec} jb={jb} nb={nb} tile={tile}: {ms:.2f} ms", flush=True)
    if not table:
        raise RuntimeError(f"autotune: no panel width / block / tile runs for n={m}")
    jb, nb, tile, ms = min(table, key=lambda r: r[3])
    best = dict(jb=jb, nb=nb, tile=tile, ms=ms, n=m, seconds=time.perf_counter() - t0, table=table)
    _CACHE[key] = best
    _save()
    return best


def main(argv=None):
    """Print the autotuner's timing table for one size and factorisation:
    python -m mom25d_amd.autotune --n 12428 --prec 32 --kind sym."""
    import argparse

    ap = argparse.ArgumentParser(description=main.__doc__.split(":")[0])
    ap.add_argument("--n", type=int, default=12428)
    ap.add_argument("--prec", type=int, default=32, choices=(32, 64))
    ap.add_argument("--kind", default="sym", choices=("pivot", "nopiv", "sym"),
                    help="pivot: mom_getrf; nopiv: mom_getrf_nopiv; sym: mom_sytrf_nopiv (L D L^T)")
    a = ap.parse_args(argv)
    best = tune_lu(a.n, a.prec, pivot=a.kind == "pivot", symmetric=a.kind == "sym", verbose=True)
    print(f"best: panel {best['jb']}, block {best['nb']}, tile {best['tile']}: {best['ms']:.2f} ms (n = {best['n']})")


if __name__ == "__main__":
    main()
