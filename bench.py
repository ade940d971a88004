#!/usr/bin/env python3
"""Benchmark of the MI355X direct kernel summation (one JSON line on stdout).

Workload (one step): GenericKernel<Laplace3D_FxU>::Eval of 1,048,576 sources
uniform by area on the unit sphere onto 1,048,576 targets uniform in the shell
1.05 < r < 2, densities N(0,1), seed 0 -- 1.0995e12 interactions -- with the
inputs already resident in HBM.  The headline value is the mixed-precision
path (FP64 tile-centred differencing, FP32 pairwise, Kahan FP64 accumulation)
in G interactions/s; the plain FP64 path is timed on the same inputs and the
mixed result is checked against it (normwise max and RMS relative error).
A step includes everything Eval does on the device (Morton ordering, tile
packing, the summation kernel).

roofline: the dominant kernel (the pairwise summation) is VALU-bound;
achieved = 11 flop/interaction (3 sub + 5 for r^2 + 1 rsqrt + 2 for the
density FMA) x interactions per launch / mean launch time, the launch time
taken from HIP events recorded around that kernel on its own stream.

cpu_baseline: the reference's own Eval<double,true> (compiled from its sources
into oracle/_ref, AVX-512 + OpenMP on all host cores) timed on a bounded
sample of the same workload (a subset of the targets against all sources).
"""
import os

os.environ.setdefault("OMP_PLACES", "cores")
os.environ.setdefault("OMP_PROC_BIND", "close")

import argparse  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402
import ctypes  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import sctl_amd as S  # noqa: E402
from sctl_amd import workloads as W  # noqa: E402
from sctl_amd.hiptime import Event  # noqa: E402

METRIC = "G interactions/s mixed vs FP64 at ≤1e-6 rel err; % of VALU peak"
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6   # MI355X published FP64 vector peak
FLOP_PER_INT = {0: 11, 1: 19, 2: 29, 3: 30}  # algorithmic flop per interaction


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed_evals(kernel, dev, out, precision, digits, steps, warmup, stream):
    xt, xs, n, f = dev
    for _ in range(warmup):
        S.eval(kernel, xt, xs, n, f, out, precision=precision, digits=digits, stream=stream)
    S.synchronize(stream)
    S.profile_reset()
    S.profile_enable(True)
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        S.eval(kernel, xt, xs, n, f, out, precision=precision, digits=digits, stream=stream)
    e1.record(stream)
    S.synchronize(stream)
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    prof = S.profile_read()
    S.profile_enable(False)
    return dict(total_ms=ms, wall_s=wall, kernel_ms=prof["kernel_ms"], launches=prof["timed_launches"])


def cpu_baseline(kernel, xt, xs, n, f, seconds):
    """The reference's own Eval<double,true> on a target subset (all sources)."""
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    kind = "reference"
    if oracle.Reference.available("v4") and oracle.have_avx512():
        ref = oracle.Reference("v4")
        run = lambda m: ref.eval(kernel, xt[:3 * m], xs, n, f, omp=True)
        what = "reference GenericKernel<Laplace3D_FxU>::Eval<double,true> (AVX-512 + OpenMP)"
    else:
        kind = "port"
        mode = oracle.AVX512 if oracle.have_avx512() else oracle.EXACT
        run = lambda m: oracle.eval(kernel, xt[:3 * m], xs, n, f, mode=mode, nthreads=cores)
        what = "oracle restatement of Eval"
    ns = xs.size // 3
    m0 = 256
    t = time.perf_counter()
    run(m0)
    rate = m0 * ns / max(time.perf_counter() - t, 1e-6)
    m = int(min(xt.size // 3, max(m0, seconds * rate / ns / 3)))
    m = max(64, m // 64 * 64)
    times = []
    for _ in range(3):
        t = time.perf_counter()
        run(m)
        times.append(time.perf_counter() - t)
    tmed = float(np.median(times))
    return {"value": round(m * ns / tmed / 1e9, 3), "unit": "G interactions/s", "cores": cores, "kind": kind,
            "sample": "%s: %d of the %d targets x all %d sources, median of 3 runs (%.2f s each)" % (
                what, m, xt.size // 3, ns, tmed)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20, help="sources = targets")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work per baseline run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    N = args.n
    K = S.LAPLACE3D_FXU
    inter = N * N
    log("generating sphere/shell inputs, N = %d" % N)
    xt, xs, nrm, f = W.sphere(N, N, seed=0)
    dev = [S.DeviceArray.from_host(a) for a in (xt, xs, nrm, f)]
    dev[2] = None  # Laplace FxU takes no normals
    stream = ctypes.c_void_p()
    S.check(S.lib().sctl_amd_stream_create(ctypes.byref(stream)))
    out = S.DeviceArray.from_host(np.zeros(N))

    log("mixed precision: %d warmup + %d timed steps" % (args.warmup, args.steps))
    mx = timed_evals(K, dev, out, S.MIXED, -1, args.steps, args.warmup, stream)
    log("fp64 (digits 15): timed steps")
    f64 = timed_evals(K, dev, out, S.FP64, -1, args.steps, min(args.warmup, 1), stream)
    log("fp64 (digits 14): timed steps")
    f64b = timed_evals(K, dev, out, S.FP64, 14, args.steps, min(args.warmup, 1), stream)

    log("accuracy: mixed vs fp64 on the full output")
    o64 = S.DeviceArray.from_host(np.zeros(N))
    omx = S.DeviceArray.from_host(np.zeros(N))
    S.eval(K, dev[0], dev[1], None, dev[3], o64, precision=S.FP64, stream=stream)
    S.eval(K, dev[0], dev[1], None, dev[3], omx, precision=S.MIXED, stream=stream)
    S.synchronize(stream)
    e_inf, e_rms = W.rel_errors(omx.download(), o64.download())

    def rate(r):
        return inter * args.steps / (r["total_ms"] * 1e-3) / 1e9

    v_mx, v_64, v_64b = rate(mx), rate(f64), rate(f64b)
    kern_ms = mx["kernel_ms"] / max(mx["launches"], 1)
    achieved = FLOP_PER_INT[K] * inter / (kern_ms * 1e-3) / 1e12
    step_tflops = FLOP_PER_INT[K] * v_mx * 1e9 / 1e12
    k64_ms = f64["kernel_ms"] / max(f64["launches"], 1)
    achieved64 = FLOP_PER_INT[K] * inter / (k64_ms * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (~%.0f s per run)" % args.cpu_seconds)
        try:
            cpu = cpu_baseline(K, xt, xs, None, f, args.cpu_seconds)
        except Exception as ex:  # reported, never fatal to the GPU measurement
            log("cpu baseline failed:", repr(ex))

    line = {
        "metric": METRIC,
        "value": round(v_mx, 2),
        "unit": "G interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mx["total_ms"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pairwise / f64 accumulate (mixed)",
        "data": "synthetic",
        "config": {
            "workload": "Laplace3D_FxU Eval, %d sphere sources x %d shell targets (1.05<r<2), N(0,1) densities, "
                        "seed 0" % (N, N),
            "kernel": "Laplace3D_FxU", "n_src": N, "n_trg": N, "interactions_per_step": inter,
            "precision": "mixed", "parallelism": "single GPU",
        },
        "max_rel_err_vs_fp64": float("%.3g" % e_inf),
        "rms_rel_err_vs_fp64": float("%.3g" % e_rms),
        "fp64": {"value": round(v_64, 2), "ms_per_step": round(f64["total_ms"] / args.steps, 3), "digits": 15,
                 "pct_fp64_valu_peak": round(100 * achieved64 / PEAK_FP64_TFLOPS, 1),
                 "digits14_value": round(v_64b, 2)},
        "speedup_vs_fp64": round(v_mx / v_64, 3),
        "speedup_vs_fp64_digits14": round(v_mx / v_64b, 3),
        "pct_fp32_valu_peak": round(100 * step_tflops / PEAK_FP32_TFLOPS, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                     "kernel": "k_eval_mixed<Laplace3D_FxU>", "mean_launch_ms": round(kern_ms, 3),
                     "flop_per_interaction": FLOP_PER_INT[K]},
        "cpu_baseline": cpu,
    }
    log("wall %.2fs for %d mixed steps (events %.1f ms)" % (mx["wall_s"], args.steps, mx["total_ms"]))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
