"""Benchmark: frequency points per second of the per-frequency solve
(Green's tables + FP64 fill + LU + port de-embedding / S) on one MI355X.

Workload (default): the 40 x 20 mm via-stitched grounded-coplanar pour
(`workloads.gcpw40`, 12 104 RWG + 324 via unknowns = 12 428, 51 frequencies
1-20 GHz).  A step is one frequency point of the sweep: Green's tables, FP64
fill, LU + port solve, de-embedding to S; by default the whole sweep is
timed (--steps K: K points spread evenly over it).  The headline path is
chosen with --precision: "mixed" (default: FP64 fill, complex64 LU on the
matrix cores as the preconditioner of FP64 GMRES refinement, each lane's LU
reused for its neighbouring frequencies while that pays) or "fp64" (the
plain path, a complex128 LU per point); the other path is timed beside it on
the same frequencies and |dS| between the two is reported.  The headline
path takes the Green's remainders of its tables from the trained networks of
the workload's stack when their weights exist (mom25d_amd/data/
surrogate_<workload>.npz; corrected and judged by the integrals at a few
nodes per frequency, --surrogate off for the plain tables); the other path
always uses the plain tables.  Inputs (mesh, fill plan) are resident on the
GPU before the timed region; the Green's tables of the timed frequencies are
built inside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--precision fp64|mixed]

Prints one JSON line (rank 0).  More than one GPU is not part of this build.
"""

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64_mfma": 78.6e12, "fp32_mfma": 157.3e12, "fp64_valu": 78.6e12, "hbm": 8e12}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# kernels timed with HIP events inside the timed sweep (library kinds of mom_timing_begin):
# kind -> (kernel, bound, peak, unit of achieved)
TIMED = {
    0: ("cgemm_sub_kernel<float, {bm}, {bn}> (complex64 LU trailing update)", "mfma", 157.3e12, "TFLOP/s"),
    4: ("cgemm_sub_kernel<double, {bm}, {bn}> (complex128 LU trailing update)", "mfma", 78.6e12, "TFLOP/s"),
    1: ("symv_strip_kernel<2, 2, 8> (GMRES product D Z D x, symmetric half of Z)", "hbm", 8e12, "GB/s"),
    2: ("sb_chain_kernel<float, 2, *> (complex64 preconditioner, one triangular chain)", "hbm", 8e12, "GB/s"),
    3: ("sb_chain_kernel<double, 2, *> (complex128 preconditioner, one triangular chain)", "hbm", 8e12, "GB/s"),
}
FILL_KERNEL = "fill_kernel<{t}> (FP64 fill, {name} tier)"


def kernel_table(launch_stats, fill_events, tile):
    """Per timed kernel: launches, summed and average duration, algorithmic
    work per launch, achieved rate and fraction of its peak (fill tiers
    against the FP64 vector peak)."""
    rows = []
    for kind, (n, ms, work) in launch_stats.items():
        if n <= 0:
            continue
        name, bound, peak, unit = TIMED[kind]
        rows.append(dict(kernel=name.format(bm=tile[0], bn=tile[1]), bound=bound, launches=int(n), total_ms=ms,
                         avg_launch_ms=ms / n, work_per_launch=work / n, achieved=work / (ms * 1e-3), peak=peak,
                         unit=unit))
    by_tier = {}
    for tier, e0, e1, flops in fill_events:
        b = by_tier.setdefault(tier, [0, 0.0, 0.0])
        b[0] += 1
        b[1] += e0.elapsed_time(e1)
        b[2] += flops
    for tier, (n, ms, work) in by_tier.items():
        rows.append(dict(kernel=FILL_KERNEL.format(t=tier, name=("near", "mid", "far")[tier]), bound="valu (fp64)",
                         launches=n, total_ms=ms, avg_launch_ms=ms / n, work_per_launch=work / n,
                         achieved=work / (ms * 1e-3), peak=PEAK["fp64_valu"], unit="TFLOP/s"))
    for r in rows:
        r["frac"] = r["achieved"] / r["peak"]
    return sorted(rows, key=lambda r: -r["total_ms"])


def time_path(mesh, sub, freqs, precision, warmup, steps, torch, lanes=None, autotune=True, panel=None, tile=None,
              block=None, interp=None, timing=False, **kw):
    """Solver set-up, warm-up sweep, then the timed sweep; timing: HIP
    events around the launches of the kernels in TIMED and of the fill tiers
    during the timed sweep (solver.kernel_stats, solver.fill_events)."""
    from mom25d_amd import Solver, _lib

    s = Solver(mesh, sub, verbose=False, precision=precision, lanes=lanes, autotune=autotune and tile is None,
               panel=panel, **kw)
    if tile is not None:
        s.tile = tuple(tile)
        s.block = block
    s.geometry().upload(s.device)
    if warmup:
        # untimed points spread over the band like the timed ones, so that every kind of point
        # (complex64 and complex128 preconditioners, their lane buffers) has run once before the clock starts
        s.sweep(sweep_points(freqs, warmup), interp=None)
    s.lu_stats = {k: 0 for k in s.lu_stats}
    s.point_log = []
    if timing:
        for kind in TIMED:
            _lib.call("mom_timing_begin", kind, 65536)
        s.fill_events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = s.sweep(sweep_points(freqs, steps), interp=interp)  # None: exact fill at every point
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    if timing:
        s.kernel_stats = {}
        out = np.zeros(3)
        for kind in TIMED:
            _lib.call("mom_timing_end", kind, out.ctypes.data_as(__import__("ctypes").c_void_p))
            s.kernel_stats[kind] = tuple(out)
    return s, r, t1 - t0


def sweep_points(freqs, steps):
    """The timed frequency points: the whole sweep when steps is its length,
    else `steps` points spread evenly over it (low-frequency points cost more
    refinement in the mixed path, so a contiguous slice would be biased)."""
    nf = len(freqs)
    if steps >= nf:
        return freqs[np.arange(steps) % nf]
    return freqs[np.round(np.linspace(0, nf - 1, steps)).astype(int)]


def fill_time(solver, freq, torch, reps=3):
    g = solver.tables([freq])[0]
    n = solver.mesh.n_unknowns
    Z = solver._buf("Z", (n, n), torch.complex128)
    solver._fill(g, Z)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        solver._fill(g, Z)
    ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) * 1e-3 / reps


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (its name up to the first " (") in
    the timed sweep of `workload`, from the committed counter summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by
    the same counters on a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        table = json.load(fh)
    return table.get("kernels", {}).get(f"{kernel.split(' (')[0]} {workload}", {}).get("bytes_per_launch")


def batch_random16(torch, precision):
    """The 16 seeded random layouts (N 601-2965, 64 frequencies each) as one
    batched sweep (sweep_batch); the first two problems warm it up."""
    from mom25d_amd import sweep_batch
    from mom25d_amd import workloads as W

    problems = [(lay.mesh(**mk), st, f) for lay, st, mk, f in W.random16()]
    sweep_batch(problems[:2], precision=precision)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rs = sweep_batch(problems, precision=precision)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    npts = sum(len(f) for _, _, f in problems)
    return dict(points=npts, seconds=dt, points_per_s=npts / dt, precision=precision,
                N=[m.n_unknowns for m, _, _ in problems], host_fit_seconds=sum(r.timing["sparams"] for r in rs))


def cpu_baseline(mesh, sub, freqs, npts, reference_defaults=False):
    """The oracle on the host: npts points with all host threads; with
    reference_defaults also the same points at the reference's own thread
    defaults (8 OpenMP threads for the fill, Engine/mom25d/_fast.py:49-50;
    4 BLAS threads, Engine/mom25d/__init__.py:5-17)."""
    from threadpoolctl import threadpool_limits

    from oracle.solve_ref import OracleSolver

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    o = OracleSolver(mesh, sub, use_c=True)
    o.geometry()
    with threadpool_limits(cores):
        t0 = time.perf_counter()
        o.sweep(freqs[:npts])
        dt = time.perf_counter() - t0
    split = {k: round(v, 3) for k, v in o.timing.items()}
    extra = {}
    if reference_defaults:
        with threadpool_limits(limits=8, user_api="openmp"), threadpool_limits(limits=4, user_api="blas"):
            t1 = time.perf_counter()
            o.sweep(freqs[:npts])
            extra["reference_defaults"] = dict(value=npts / (time.perf_counter() - t1), fill_threads=8, blas_threads=4,
                                               split={k: round(v, 3) for k, v in o.timing.items()})
    return dict(value=npts / dt, unit="freq points/s", cores=cores, kind="port", **extra,
                sample=f"{npts} frequency point(s) ({', '.join(f'{f / 1e9:g}' for f in freqs[:npts])} GHz) of the same "
                       f"workload: oracle tables (scipy J0), C/OpenMP pair kernel + serial scatter, "
                       f"LAPACK zgetrf/zgetrs (OpenBLAS), numpy de-embedding; per-mesh set-up untimed; "
                       f"CPU {_cpu_model()}",
                split=split)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="frequency points timed, spread over the sweep (default: the whole sweep)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="gcpw40")
    ap.add_argument("--cpu-sample", type=int, default=1, help="frequency points timed on the CPU oracle (0: skip)")
    ap.add_argument("--precision", default="mixed", choices=("fp64", "mixed"), help="headline path")
    ap.add_argument("--no-alt", action="store_true", help="skip timing the other path")
    ap.add_argument("--lanes", type=int, default=None, help="frequency points in flight (default: automatic)")
    ap.add_argument("--no-tune", action="store_true", help="default LU panel width / tiles instead of the autotuner's")
    ap.add_argument("--panel", type=int, default=None, help="LU panel width (with --tile: instead of the autotuner)")
    ap.add_argument("--tile", default=None, help="update-GEMM tile BM,BN of the headline path (skips the autotuner)")
    ap.add_argument("--block", type=int, default=None, help="LU update depth (with --tile)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with a plain 1 GiB device copy (reference for the HBM byte counters)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched sweep of the 16 random layouts")
    ap.add_argument("--no-interp", action="store_true",
                    help="skip timing the headline path with the interpolated impedance (interp='auto')")
    ap.add_argument("--cpu-reference-defaults", action="store_true",
                    help="also time the CPU sample at the reference's thread defaults (8 fill, 4 BLAS threads)")
    ap.add_argument("--lu-equiv", type=float, default=None,
                    help="mixed path: cost of a factorisation in GMRES steps for the LU reuse policy (default: model)")
    ap.add_argument("--lu64-steps", type=int, default=20,
                    help="mixed path: GMRES steps after which a complex64 solve gives way to a complex128 anchor")
    ap.add_argument("--solve-sb", type=int, default=512,
                    help="mixed path: super-block rows of the preconditioner solves (0: the 64-row chain)")
    ap.add_argument("--surrogate", default="auto", choices=("auto", "off"),
                    help="headline path: Green's remainders from the trained networks of the workload's stack "
                         "(mom25d_amd/data/surrogate_<workload>.npz, corrected and judged by the integrals); "
                         "auto: when the weights exist")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="no HIP events around the timed sweep's launches (then no roofline object)")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES of this process (HIP's default is 4)")
    args = ap.parse_args()
    # One hardware queue per stream: each lane of a sweep has a stream, and its factorisations a look-ahead
    # stream.  With HIP's 4 queues per process, streams share queues, and a kernel waits behind the kernels
    # (and event waits) of another lane queued on the same hardware queue.  Read by HIP when it initialises,
    # so it is set before torch touches the GPU.
    os.environ["GPU_MAX_HW_QUEUES"] = str(max(1, min(32, args.hw_queues)))
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build runs on one GPU only (frequency sharding across GPUs is a design note, "
              "not implemented); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from mom25d_amd import workloads as W

    lay, sub, mk, freqs = W.ALL[args.workload]()
    if args.steps is None:
        args.steps = len(freqs)
    mesh = lay.mesh(**mk)
    N = mesh.n_unknowns
    print(f"workload {args.workload}: N={N} ({mesh.n_basis} RWG + {mesh.n_vias} vias), {len(freqs)} f", file=sys.stderr,
          flush=True)

    prec = args.precision
    alt = "mixed" if prec == "fp64" else "fp64"
    tile = tuple(int(x) for x in args.tile.split(",")) if args.tile else None
    mixed_kw = dict(lu_equiv=args.lu_equiv, lu64_steps=args.lu64_steps if args.lu64_steps >= 0 else None,
                    solve_sb=args.solve_sb)
    head_kw = dict(mixed_kw if prec == "mixed" else {})
    sur_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mom25d_amd", "data",
                            f"surrogate_{args.workload}.npz")
    if args.surrogate == "auto" and os.path.exists(sur_path):
        from mom25d_amd.surrogate import GreensSurrogate

        head_kw["surrogate"] = GreensSurrogate.load(sur_path, sub.to_stackup() if hasattr(sub, "to_stackup") else sub)
    sm, rm, tm = time_path(mesh, sub, freqs, prec, args.warmup, args.steps, torch, args.lanes, not args.no_tune,
                           args.panel, tile, args.block, timing=not args.no_kernel_timing, **head_kw)
    sur_line = None
    if "surrogate" in head_kw:
        rep = sm.last_tables.surrogate_report
        sur_line = dict(weights=os.path.relpath(sur_path, os.path.dirname(os.path.abspath(__file__))),
                        accepted=int(len(rep["check_err"]) - len(rep["fallback"])), tables=int(len(rep["check_err"])),
                        tol=rep["tol"], check_err_max=float(np.max(rep["check_err"])),
                        bessel_ratio=float(rep["integral_nodes"] / sm.last_tables.integral_nodes_total))
        print(f"surrogate remainders: {sur_line}", file=sys.stderr, flush=True)
    for rec in sm.point_log:
        print(f"  point {rec['freq'] / 1e9:6.3f} GHz lane {rec['lane']}: " +
              ", ".join(f"{k} {v}" for k, v in rec["solves"]), file=sys.stderr)
    lanes = sm._lane_count()
    sweep_lu_stats = dict(sm.lu_stats)  # the timed sweep's (later solves add to sm.lu_stats)
    print(f"{prec}: {args.steps} points in {tm:.3f} s (LU: {sm.lu_stats})", file=sys.stderr, flush=True)
    other = None
    if not args.no_alt:
        sa, ra, ta = time_path(mesh, sub, freqs, alt, args.warmup, args.steps, torch, args.lanes, not args.no_tune,
                               **(mixed_kw if alt == "mixed" else {}))
        other = dict(precision=alt, value=args.steps / ta, ms_per_step=1e3 * ta / args.steps,
                     dS_max_vs_headline=float(np.max(np.abs(rm.S - ra.S))), lu_stats=dict(sa.lu_stats))
        print(f"{alt}: {args.steps} points in {ta:.3f} s; |dS| {other['dS_max_vs_headline']:.2e}", file=sys.stderr,
              flush=True)
        del sa

    interp_line = None
    if not args.no_interp and args.steps == len(freqs):
        si, rI, tI = time_path(mesh, sub, freqs, prec, args.warmup, args.steps, torch, args.lanes, not args.no_tune,
                               args.panel, tile, args.block, interp="auto", **(mixed_kw if prec == "mixed" else {}))
        fi = si.last_interp
        interp_line = dict(value=args.steps / tI, ms_per_step=1e3 * tI / args.steps,
                           anchors=None if fi is None else len(fi.anchors),
                           dS_max_vs_headline=float(np.max(np.abs(rI.S - rm.S))),
                           dS_max_vs_fp64=None if other is None else float(np.max(np.abs(rI.S - ra.S))))
        print(f"{prec} with interp='auto': {args.steps} points in {tI:.3f} s; {interp_line}", file=sys.stderr,
              flush=True)
        del si

    f_probe = freqs[len(freqs) // 2]
    from mom25d_amd.autotune import get_tile

    bm, bn = sm.tile or get_tile(64 if prec == "fp64" else 32)
    # the roofline of the kernel with the largest summed time in the timed sweep (HIP events on
    # each launch's stream); the other timed kernels beside it
    table = kernel_table(sm.kernel_stats, sm.fill_events, (bm, bn)) if not args.no_kernel_timing else []
    sm.fill_events = None
    roof = None
    if table:
        top = table[0]
        scale = 1e12 if top["unit"] == "TFLOP/s" else 1e9
        roof = dict(bound=top["bound"], achieved=top["achieved"] / scale, peak=top["peak"] / scale, unit=top["unit"],
                    frac=top["frac"], traffic=measured_traffic(top["kernel"], args.workload), kernel=top["kernel"],
                    launches=top["launches"], avg_launch_ms=top["avg_launch_ms"], summed_ms=top["total_ms"],
                    work_per_launch=top["work_per_launch"],
                    work_unit="flop" if top["unit"] == "TFLOP/s" else "algorithmic bytes")
    for r in table:
        print(f"  timed kernel {r['kernel']}: {r['launches']} launches, {r['total_ms']:.1f} ms summed, "
              f"{r['avg_launch_ms'] * 1e3:.1f} us avg, {r['frac']:.3f} of peak", file=sys.stderr)
    tf = fill_time(sm, f_probe, torch)
    fill_flops = sm.geometry().algorithmic_flops()
    extra = dict(fill=dict(seconds=tf, algorithmic_flops=fill_flops, achieved_tflops=fill_flops / tf / 1e12,
                           frac_fp64_valu=fill_flops / tf / PEAK["fp64_valu"]))
    extra["lu_config"] = dict(panel=sm.panel, block=sm.block or sm.panel, tile=[bm, bn],
                              autotuned=not args.no_tune and tile is None)
    extra["timed_kernels"] = [{k: (round(v, 6) if isinstance(v, float) else v) for k, v in r.items()} for r in table]
    extra["lu_stats"] = sweep_lu_stats  # factorisations / reused LUs / GMRES steps in the timed sweep
    if prec == "mixed":
        extra["mixed_policy"] = dict(lu_equiv=sm._lu_equiv(), lu64_steps=sm.lu64_steps, solve_sb=sm.solve_sb)
    if sur_line is not None:
        extra["surrogate"] = sur_line  # the headline's tables: networks corrected and judged per frequency
    if interp_line is not None:
        extra["interp_auto"] = interp_line  # the reference's default sweep mode (interpolated Z between anchors)
    if not args.no_batch:
        extra["batch_random16"] = batch_random16(torch, prec)
        b = extra["batch_random16"]
        print(f"batch random16 ({prec}): {b['points']} points in {b['seconds']:.2f} s = {b['points_per_s']:.1f} points/s",
              file=sys.stderr, flush=True)
    cpu = None
    if args.cpu_sample > 0:
        cpu = cpu_baseline(mesh, sub, freqs, args.cpu_sample, args.cpu_reference_defaults)
        print(f"cpu baseline: {cpu['value']:.4f} points/s {cpu['split']}", file=sys.stderr, flush=True)

    line = {
        "metric": "freq points/s (fill+LU+S) at N~12k RWG, |dS|<1e-4 vs FP64; fill FLOP/s",
        "value": args.steps / tm,
        "unit": "freq points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * tm / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64 (complex128 fill and LU)" if prec == "fp64" else "f64 fill / c64 LU + f64 GMRES refinement")
                 + ("; Green's remainders from the surrogate networks" if sur_line else ""),
        "data": "synthetic (seeded / deterministic layouts)",
        "config": {"workload": f"{args.workload}: 40x20 mm via-stitched GCPW pour, N={N} unknowns, "
                               f"{len(freqs)} f 1-20 GHz" if args.workload == "gcpw40" else f"{args.workload}: N={N}",
                   "N": N, "precision": prec, "lanes": lanes, "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "other_path": other,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    if args.calibrate_copy:
        n = 1 << 28  # 2^28 float32 = 1 GiB read + 1 GiB written
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"calibration copy: {4 * n} bytes read, {4 * n} bytes written", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
