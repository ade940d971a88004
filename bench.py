#!/usr/bin/env python3
"""Benchmark of the MI355X direct kernel summation (one JSON line on stdout).

Workload (one step): GenericKernel<Laplace3D_FxU>::Eval of 1,048,576 sources
uniform by area on the unit sphere onto 1,048,576 targets uniform in the shell
1.05 < r < 2, densities N(0,1), seed 0 -- 1.0995e12 interactions -- with the
inputs already resident in HBM.  The headline value is the mixed-precision
path (FP64 tile-centred differencing, FP32 pairwise, Kahan FP64 accumulation)
in G interactions/s; the plain FP64 path is timed on the same inputs and the
mixed result is checked against it (normwise max and RMS relative error).
A step includes everything Eval does on the device (Hilbert ordering, tile
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_eval_mixed_laplace_fxu_sphere_1M.json")


def measured_traffic(n):
    """HBM bytes per launch of the headline kernel from the committed
    FETCH_SIZE / WRITE_SIZE counter runs (copy-calibrated), for the full-size
    workload only."""
    path = TRAFFIC_FILE
    if n != 1 << 20 or not os.path.exists(path):
        return None
    with open(path) as fh:
        return float(json.load(fh)["traffic_bytes"])


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed_evals(kernel, dev, out, precision, digits, steps, warmup, stream, flags=0):
    xt, xs, n, f = dev
    for _ in range(warmup):
        S.eval(kernel, xt, xs, n, f, out, precision=precision, digits=digits, stream=stream, flags=flags)
    S.synchronize(stream)
    S.profile_reset()
    S.profile_enable(True)
    e0, e1 = Event(), Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        S.eval(kernel, xt, xs, n, f, out, precision=precision, digits=digits, stream=stream, flags=flags)
    e1.record(stream)
    S.synchronize(stream)
    wall = time.perf_counter() - t0
    ms = e0.elapsed_ms(e1)
    prof = S.profile_read()
    S.profile_enable(False)
    return dict(total_ms=ms, wall_s=wall, kernel_ms=prof["kernel_ms"], launches=prof["timed_launches"])


def host_cpu(avx512):
    """The host the CPU baseline ran on: model, logical CPUs, ISA class, OpenMP threads."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "isa": "avx512" if avx512 else "below avx512",
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(kernel, xt, xs, n, f, seconds):
    """The reference's own Eval<double,true> on a target subset (all sources)."""
    import oracle
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    kind = "reference"
    kname = S.kernel_info(kernel)[6].replace("-", "_")
    if oracle.Reference.available("v4") and oracle.have_avx512():
        ref = oracle.Reference("v4")
        run = lambda m: ref.eval(kernel, xt[:3 * m], xs, n, f, omp=True)
        what = "reference GenericKernel<%s>::Eval<double,true> (AVX-512 + OpenMP)" % kname
    else:
        kind = "port"
        mode = oracle.AVX512 if oracle.have_avx512() else oracle.EXACT
        run = lambda m: oracle.eval(kernel, xt[:3 * m], xs, n, f, mode=mode, nthreads=cores)
        what = "oracle restatement of Eval"
    ns = xs.size // 3
    m0 = min(256, xt.size // 3)
    run(m0)  # warm-up: the OpenMP thread pool
    t = time.perf_counter()
    run(m0)
    rate = m0 * ns / max(time.perf_counter() - t, 1e-6)
    m = int(min(xt.size // 3, max(m0, seconds * rate / ns / 3)))
    m = max(min(64, xt.size // 3), m // 64 * 64)
    # a sample shorter than a third of `seconds` is repeated (small configurations)
    reps = max(1, int(seconds / 3 * rate / (m * ns)))
    times = []
    for _ in range(3):
        t = time.perf_counter()
        for _ in range(reps):
            run(m)
        times.append((time.perf_counter() - t) / reps)
    tmed = float(np.median(times))
    return {"value": round(m * ns / tmed / 1e9, 3), "unit": "G interactions/s", "cores": cores, "kind": kind,
            "sample": "%s: %d of the %d targets x all %d sources, median of 3 runs of %d repeats (%.4f s each)" % (
                what, m, xt.size // 3, ns, reps, tmed), "host": host_cpu(oracle.have_avx512())}


def cpu_blocks_baseline(kernel, B, seconds, mode=0):
    """BoundaryIntegralOp's per-block loop of the reference, compiled from its
    sources (OpenMP over the blocks, boundary_integral.txx:1165-1211): each
    block by GenericKernel::Eval<double,false> (mode 0) or
    KernelMatrix<double,false> (mode 1), on the first m blocks, m grown until
    one run takes about `seconds`; rate = the sample's Ns*Nt / time."""
    import oracle
    if not (oracle.Reference.available("v4") and oracle.have_avx512()):
        return None
    ref = oracle.Reference("v4")
    cnt = B["trg_cnt"] * B["src_cnt"]
    total = int(cnt.sum())

    def run(m):
        args = (B["trg_dsp"][:m], B["trg_cnt"][:m], B["src_dsp"][:m], B["src_cnt"][:m], B["r_trg"], B["r_src"],
                B["n_src"])
        if mode == 0:
            ref.blocks(kernel, 0, *args, v_src=B["v_src"])
        else:
            ref.blocks(kernel, 1, *args)

    m = 16
    t = time.perf_counter()
    run(m)
    dt = max(time.perf_counter() - t, 1e-6)
    inter = int(cnt[:m].sum())
    while dt < seconds / 3 and m < cnt.size:
        m = int(min(cnt.size, max(2 * m, m * seconds / dt)))
        t = time.perf_counter()
        run(m)
        dt = max(time.perf_counter() - t, 1e-6)
        inter = int(cnt[:m].sum())
    times = []
    for _ in range(3):
        t = time.perf_counter()
        run(m)
        times.append(time.perf_counter() - t)
    tmed = float(np.median(times))
    what = "Eval<double,false>" if mode == 0 else "KernelMatrix<double,false>"
    return {"value": round(inter / tmed / 1e9, 3), "unit": "G interactions/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1), "kind": "reference",
            "sample": "reference per-block %s loop (OpenMP over blocks, AVX-512): first %d of %d blocks (%.3g of "
                      "%.3g interactions), median of 3 runs (%.2f s each)" % (what, m, cnt.size, inter, total, tmed),
            "host": host_cpu(oracle.have_avx512())}


def kernel_matrix_end_to_end(kernel, B, nblocks, threads=8):
    """The drop-in's KernelMatrix as BoundaryIntegralOp calls it: per block,
    host arrays in, the (Ns*K0) x (Nt*K1) matrix back in host memory
    (staging, kernel and device-to-host copy included), from `threads` host
    threads at once over the first `nblocks` blocks."""
    from concurrent.futures import ThreadPoolExecutor
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    L = S.lib()
    sizes = [(int(B["trg_cnt"][b]) * K1, int(B["src_cnt"][b]) * K0) for b in range(nblocks)]
    biggest = max(a * c for a, c in sizes)
    bufs = [np.empty(biggest) for _ in range(threads)]
    xt, xs, xn = B["r_trg"], B["r_src"], B["n_src"]

    def one(b, buf):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        S.check(L.sctl_amd_kernel_matrix(kernel, -1, xt[3 * t0:].ctypes.data, tc, xs[3 * s0:].ctypes.data,
                                         xn[3 * s0:].ctypes.data if ND else None, sc, buf.ctypes.data,
                                         S.HOST_PTRS, None))

    def worker(w):
        for b in range(w, nblocks, threads):
            one(b, bufs[w])

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(worker, range(threads)))  # warm-up pass: per-thread streams, device pool, pinned staging
        t = time.perf_counter()
        list(ex.map(worker, range(threads)))
        dt = time.perf_counter() - t
    entries = int(sum(B["trg_cnt"][:nblocks] * B["src_cnt"][:nblocks])) * K0 * K1
    inter = int(sum(B["trg_cnt"][:nblocks] * B["src_cnt"][:nblocks]))
    return {"blocks": nblocks, "host_threads": threads, "s": round(dt, 4),
            "G_int_per_s": round(inter / dt / 1e9, 2), "matrix_GB": round(entries * 8 / 1e9, 3),
            "matrix_GB_per_s_to_host": round(entries * 8 / dt / 1e9, 2)}


def time_once(fn, stream, reps=1):
    e0, e1 = Event(), Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    S.synchronize(stream)
    return e0.elapsed_ms(e1) / reps


def dense_pair(kernel, xt, xs, n, f, stream, reps, digits=-1, ctx=None, warm=1, blocks=1):
    """FP64 and mixed Eval of one dense configuration: rates and mixed-vs-FP64
    errors.  Each precision is warmed with `warm` calls; with blocks > 1 the
    two precisions are timed in alternating blocks of `reps` calls and each
    reports its median block (calls of a fraction of a millisecond are timed
    at the device's steady clock rather than its ramp)."""
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    nt, ns = xt.size // 3, xs.size // 3
    d = [S.DeviceArray.from_host(a) for a in (xt, xs, n if ND else np.zeros(3), f)]
    dn = d[2] if ND else None
    res = {}
    outs = {}
    calls = {}
    for name, prec in (("fp64", S.FP64), ("mixed", S.MIXED)):
        o = S.DeviceArray.from_host(np.zeros(nt * K1))
        S.eval(kernel, d[0], d[1], dn, d[3], o, precision=prec, digits=digits, stream=stream, ctx=ctx)  # warm-up
        S.synchronize(stream)
        outs[name] = o.download()
        calls[name] = (lambda o=o, prec=prec: S.eval(kernel, d[0], d[1], dn, d[3], o, precision=prec, digits=digits,
                                                     stream=stream, ctx=ctx))
        for _ in range(warm - 1):
            calls[name]()
    S.synchronize(stream)
    times = {name: [] for name in calls}
    for _ in range(blocks):
        for name, fn in calls.items():
            times[name].append(time_once(fn, stream, reps))
    for name in calls:
        ms = float(np.median(times[name]))
        res[name] = {"ms": round(ms, 4 if ms < 1 else 3), "G_int_per_s": round(nt * ns / ms / 1e6, 2)}
        if blocks > 1:
            res[name]["timing"] = "median of %d alternating blocks of %d calls after %d warm-up calls" % (blocks, reps,
                                                                                                       warm)
    e_inf, e_rms = W.rel_errors(outs["mixed"], outs["fp64"])
    res["mixed_vs_fp64"] = {"max_rel": float("%.3g" % e_inf), "rms_rel": float("%.3g" % e_rms)}
    res["speedup"] = round(res["fp64"]["ms"] / res["mixed"]["ms"], 3)
    return res, outs


def dense_plan_rates(kernel, xt, xs, n, f, stream, reps, blocks=10):
    """Repeated applies of a dense plan (geometry resident, set-up once):
    rates per precision (median of `blocks` alternating blocks of `reps`
    applies after `reps` warm-up applies) and the setup time."""
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    nt, ns = xt.size // 3, xs.size // 3
    d = [S.DeviceArray.from_host(a) for a in (xt, xs, n if ND else np.zeros(3), f)]
    res, apply = {}, {}
    for name, prec in (("fp64", S.FP64), ("mixed", S.MIXED)):
        S.synchronize(stream)  # the setup clock must not include the other plan's warm-up applies
        t = time.perf_counter()
        P = S.DensePlan(kernel, d[0], d[1], d[2] if ND else None, precision=prec, stream=stream)
        setup_ms = (time.perf_counter() - t) * 1e3
        o = S.DeviceArray.from_host(np.zeros(nt * K1))
        apply[name] = (lambda P=P, o=o: P.apply(d[3], o, stream=stream))
        for _ in range(reps):
            apply[name]()
        res[name] = {"setup_ms": round(setup_ms, 3)}
    S.synchronize(stream)
    times = {name: [] for name in apply}
    for _ in range(blocks):
        for name, fn in apply.items():
            times[name].append(time_once(fn, stream, reps))
    for name in apply:
        ms = float(np.median(times[name]))
        res[name].update({"ms_per_apply": round(ms, 4), "G_int_per_s": round(nt * ns / ms / 1e6, 2)})
    res["speedup"] = round(res["fp64"]["ms_per_apply"] / res["mixed"]["ms_per_apply"], 3)
    return res


def plan_pair(kernel, B, stream, reps):
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    dv = S.DeviceArray.from_host(B["v_src"])
    geo = [S.DeviceArray.from_host(B[a]) for a in ("r_trg", "r_src", "n_src")]
    res, outs = {}, {}
    for name, prec in (("fp64", S.FP64), ("mixed", S.MIXED)):
        P = S.Plan(kernel, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
        P.set_geometry(geo[0], geo[1], geo[2] if ND else None, stream=stream)
        o = S.DeviceArray.from_host(np.zeros(nt * K1))
        spl, tpl, cap, _ = P.tune(dv, o, stream=stream)
        res[name + "_shape"] = {"sources_per_lane": spl, "targets_per_lane": tpl, "wg_per_cu_cap": cap}
        o.upload(np.zeros(nt * K1))
        P.apply(dv, o, stream=stream)
        S.synchronize(stream)
        outs[name] = o.download()
        ms = time_once(lambda: P.apply(dv, o, stream=stream), stream, reps)
        res[name] = {"ms_per_apply": round(ms, 3), "G_int_per_s": round(P.interactions() / ms / 1e6, 2)}
        res["interactions_per_apply"] = P.interactions()
        P.destroy()
    e_inf, e_rms = W.rel_errors(outs["mixed"], outs["fp64"])
    res["mixed_vs_fp64"] = {"max_rel": float("%.3g" % e_inf), "rms_rel": float("%.3g" % e_rms)}
    res["speedup"] = round(res["fp64"]["ms_per_apply"] / res["mixed"]["ms_per_apply"], 3)
    return res, outs


def graph_pair(kernel, B, stream, nsteps, seed=0, tune=True):
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    geo = [S.DeviceArray.from_host(B[a]) for a in ("r_trg", "r_src", "n_src")]
    res, outs = {}, {}
    for name, prec in (("fp64", S.FP64), ("mixed", S.MIXED)):
        P = S.Plan(kernel, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
        P.set_geometry(geo[0], geo[1], geo[2] if ND else None, stream=stream)
        vs = S.DeviceArray(ns * K0)
        vt = S.DeviceArray.from_host(np.zeros(nt * K1))
        S.gen_normal(seed, 25, vs, stream=stream)
        if tune:
            P.tune(vs, vt, stream=stream)
        spl, tpl, cap = P.shape()
        res[name + "_shape"] = {"sources_per_lane": spl, "targets_per_lane": tpl, "wg_per_cu_cap": cap}
        vt.upload(np.zeros(nt * K1))
        G = P.graph(nsteps, seed, 25, vs, vt)
        G.launch(stream)  # warm-up replay
        S.synchronize(stream)
        vt.upload(np.zeros(nt * K1))
        ms = time_once(lambda: G.launch(stream), stream, 1)
        outs[name] = vt.download()
        res[name] = {"ms_per_graph": round(ms, 3), "ms_per_apply": round(ms / nsteps, 4),
                     "G_int_per_s": round(G.interactions() / ms / 1e6, 2)}
        res["interactions_per_apply"] = P.interactions()
        G.destroy()
        P.destroy()
    e_inf, e_rms = W.rel_errors(outs["mixed"], outs["fp64"])
    res["mixed_vs_fp64"] = {"max_rel": float("%.3g" % e_inf), "rms_rel": float("%.3g" % e_rms)}
    res["applies"] = nsteps
    res["speedup"] = round(res["fp64"]["ms_per_graph"] / res["mixed"]["ms_per_graph"], 3)
    return res


def calibration_copy(mb, stream):
    """One device-to-device copy of mb MiB: the reference point that scales
    the HBM byte counters of a profiled run."""
    nb = mb << 20
    a, b = S.DeviceArray(nb // 8), S.DeviceArray(nb // 8)
    S.check(S.lib().sctl_amd_memcpy(b.ptr, a.ptr, nb, 2, stream))
    S.synchronize(stream)
    log("calibration copy: %d bytes device to device" % nb)
    a.free()
    b.free()


def near_ops(kernel, B, stream, reps):
    """BoundaryIntegralOp's near-field operator over one config's blocks, on
    the device: the block matrices K_near (target-major, FP64) written
    (store), updated (KNearSubtract) and applied (ComputeNearInterac's GEMV).
    HBM-bound: bytes are 8 per matrix entry written (store), read + written
    (subtract) and read (GEMV, plus its F and U_near).  Parity: the GEMV of
    the stored matrices equals the plan's direct apply of w*F."""
    _, K0, K1, ND, _, _, _ = S.kernel_info(kernel)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = S.Plan(kernel, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    geo = [S.DeviceArray.from_host(B[a]) for a in ("r_trg", "r_src", "n_src")]
    P.set_geometry(geo[0], geo[1], geo[2] if ND else None, stream=stream)
    k_len = P.near_layout()
    w = W.uniform(0, 40, ns, 0.5, 1.5)
    dw = S.DeviceArray.from_host(w)
    dK = S.DeviceArray(k_len)
    dF = S.DeviceArray.from_host(B["v_src"])
    dU = S.DeviceArray.from_host(np.zeros(nt * K1))
    peak = 8000.0
    res = {"matrix_entries": int(k_len), "matrix_GB": round(k_len * 8 / 1e9, 2)}
    P.near_matrix(dK, dw, None, S.NEAR_STORE, stream=stream)  # warm-up, and the matrices the GEMV uses
    S.synchronize(stream)
    for name, mode, bpe in (("store", S.NEAR_STORE, 8), ("subtract", S.NEAR_SUBTRACT, 16)):
        ms = time_once(lambda: P.near_matrix(dK, dw, None, mode, stream=stream), stream, reps)
        gbs = k_len * bpe / ms / 1e6
        res[name] = {"ms": round(ms, 3), "GB_per_s": round(gbs, 1), "frac_hbm_peak": round(gbs / peak, 3)}
    P.near_matrix(dK, dw, None, S.NEAR_STORE, stream=stream)
    P.near_gemv(dK, dF, dU, stream=stream)
    S.synchronize(stream)
    via_matrix = dU.download()
    ms = time_once(lambda: P.near_gemv(dK, dF, dU, stream=stream), stream, reps)
    gbs = (k_len * 8 + ns * K0 * 8 + nt * K1 * 8) / ms / 1e6
    res["gemv"] = {"ms": round(ms, 3), "GB_per_s": round(gbs, 1), "frac_hbm_peak": round(gbs / peak, 3)}
    o = S.DeviceArray.from_host(np.zeros(nt * K1))
    dwf = S.DeviceArray.from_host(np.repeat(w, K0) * B["v_src"])
    P.apply(dwf, o, stream=stream)
    S.synchronize(stream)
    e_inf, e_rms = W.rel_errors(via_matrix, o.download())
    res["gemv_of_stored_vs_direct_apply"] = {"max_rel": float("%.3g" % e_inf), "rms_rel": float("%.3g" % e_rms)}
    for a in (dK, dF, dU, dw, dwf, o) + tuple(geo):
        a.free()
    P.destroy()
    return res


def _near_list_ms(xt, xs, radius, cnt_h, dsp_h, stream, reps):
    """(median ms, pairs) of device near-list builds on device-resident inputs
    (the build sizes its output, so it synchronises: host clock)."""
    ne = cnt_h.size
    d = [S.DeviceArray.from_host(a) for a in (xt, xs)]
    rad = S.DeviceArray.from_host(np.full(xs.size // 3, radius))
    cnt = S.DeviceArray.from_host(cnt_h.astype(np.int64), dtype=np.int64)
    dsp = S.DeviceArray.from_host(dsp_h.astype(np.int64), dtype=np.int64)
    L = S.lib()
    nt, ns = xt.size // 3, xs.size // 3
    times, n_near = [], 0
    for it in range(reps + 1):
        h = ctypes.c_void_p()
        S.synchronize(stream)
        t0 = time.perf_counter()
        S.check(L.sctl_amd_near_list_build(ctypes.byref(h), d[0].ptr, nt, None, d[1].ptr, ns, rad.ptr, ne, cnt.ptr,
                                           dsp.ptr, 0, stream))
        t1 = time.perf_counter()
        n = ctypes.c_int64()
        S.check(L.sctl_amd_near_list_size(h, ctypes.byref(n)))
        n_near = n.value
        L.sctl_amd_near_list_destroy(h)
        if it:
            times.append((t1 - t0) * 1e3)
    return float(np.median(times)), n_near


def near_list_build(B, radius, stream, reps=3, cpu_elements=256, cpu=True):
    """BuildNearList on the device for one config's surface: the elements are
    the blocks' source patches, the targets the blocks' targets, every node's
    radius `radius`.  Beside it, the reference's own BuildNearList (compiled
    from its sources, on the host cores) on a bounded sample -- the first
    `cpu_elements` elements against all targets -- and the device on the
    same sample."""
    ms, n_near = _near_list_ms(B["r_trg"], B["r_src"], radius, B["src_cnt"], B["src_dsp"], stream, reps)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    res = {"elements": int(B["src_cnt"].size), "nodes": ns, "targets": nt, "radius": radius,
           "near_pairs": n_near, "ms": round(ms, 3), "M_pairs_per_s": round(n_near / ms / 1e3, 1)}
    import oracle
    if cpu and oracle.ReferenceNearList.available():
        cnt, dsp = B["src_cnt"][:cpu_elements], B["src_dsp"][:cpu_elements]
        nss = int(dsp[-1] + cnt[-1])
        xs = B["r_src"][:3 * nss]
        sec, pairs = oracle.ReferenceNearList().build_seconds(B["r_trg"], xs, np.full(nss, radius), cnt, dsp)
        gms, gpairs = _near_list_ms(B["r_trg"], xs, radius, cnt, dsp, stream, reps)
        res["cpu_reference_sample"] = {
            "sample": "first %d elements (%d nodes) x all %d targets" % (cpu_elements, nss, nt),
            "kind": "reference", "cores": os.cpu_count(), "s": round(sec, 3), "pairs": pairs,
            "device_ms_same_sample": round(gms, 3), "device_pairs_same_sample": gpairs}
    return res


def hbm_copy_reference(stream, gib=4):
    """Achievable HBM bandwidth for context: the runtime's device-to-device
    copy of `gib` GiB (read + write bytes / time) and its fill of the same
    buffer (write bytes / time)."""
    from sctl_amd.hiptime import memset_async
    n = gib << 30
    a, b = S.DeviceArray(n // 8), S.DeviceArray(n // 8)
    L = S.lib()
    S.check(L.sctl_amd_memcpy(b.ptr, a.ptr, n, 2, stream))
    S.synchronize(stream)
    ms = time_once(lambda: S.check(L.sctl_amd_memcpy(b.ptr, a.ptr, n, 2, stream)), stream, 3)
    memset_async(b.ptr, 0, n, stream)
    S.synchronize(stream)
    ms_fill = time_once(lambda: memset_async(b.ptr, 0, n, stream), stream, 3)
    a.free()
    b.free()
    return {"copy": {"bytes": 2 * n, "ms": round(ms, 3), "GB_per_s": round(2 * n / ms / 1e6, 1),
                     "frac_hbm_peak": round(2 * n / ms / 1e6 / 8000.0, 3)},
            "fill": {"bytes": n, "ms": round(ms_fill, 3), "GB_per_s": round(n / ms_fill / 1e6, 1),
                     "frac_hbm_peak": round(n / ms_fill / 1e6 / 8000.0, 3)}}


def add_roofline(r, kernel, key_ms="ms", inter=None):
    """Per-path roofline of one configuration line: algorithmic TFLOP/s
    (flop per interaction x rate) against the FP64 (78.6) or FP32 (157.3)
    VALU peak."""
    flop = FLOP_PER_INT.get(kernel, S.kernel_info(kernel)[4])
    out = {"flop_per_interaction": flop}
    for name, peak in (("fp64", PEAK_FP64_TFLOPS), ("mixed", PEAK_FP32_TFLOPS)):
        if name in r and "G_int_per_s" in r[name]:
            tf = r[name]["G_int_per_s"] * flop / 1e3
            out[name] = {"achieved_TFLOP_per_s": round(tf, 2), "peak": peak, "frac": round(tf / peak, 4)}
    r["roofline"] = out
    return r


def other_configs(stream, quick=False, cpu=True, cpu_seconds=3.0):
    """The other configurations: measured, not the headline.  With cpu, the
    reference's own CPU code is timed beside each on a bounded sample."""
    import oracle
    out = {}
    log("config 1: Laplace FxU FP64, 20000 x 20000 uniform cube, parity vs the oracle")
    xt, xs, f = W.cube(20000, 20000, seed=0)
    r, o = dense_pair(S.LAPLACE3D_FXU, xt, xs, None, f, stream, reps=20, warm=20, blocks=10)
    add_roofline(r, S.LAPLACE3D_FXU)
    if cpu:
        r["cpu_reference"] = cpu_baseline(S.LAPLACE3D_FXU, xt, xs, None, f, 1.0)
    ref = oracle.eval(S.LAPLACE3D_FXU, xt, xs, None, f)
    r["fp64_vs_oracle"] = dict(zip(("max_rel", "rms_rel"), (float("%.3g" % e) for e in W.rel_errors(o["fp64"], ref))))
    r["mixed_vs_oracle"] = dict(zip(("max_rel", "rms_rel"), (float("%.3g" % e) for e in W.rel_errors(o["mixed"], ref))))
    r["dense_plan"] = dense_plan_rates(S.LAPLACE3D_FXU, xt, xs, None, f, stream, reps=20)
    out["c1_laplace_fxu_cube_20k"] = r
    N = 1 << 20
    log("config 2 (DxU leg): Laplace DxU, 2^20 sphere sources x 2^20 shell targets")
    xt, xs, nrm, f = W.sphere(N, N, seed=0)
    r, _ = dense_pair(S.LAPLACE3D_DXU, xt, xs, nrm, f, stream, reps=1 if quick else 2)
    add_roofline(r, S.LAPLACE3D_DXU)
    if cpu:
        log("config 2: reference CPU Eval on a target sample")
        r["cpu_reference"] = cpu_baseline(S.LAPLACE3D_DXU, xt, xs, nrm, f, cpu_seconds)
    out["c2_laplace_dxu_sphere_1M"] = r
    log("config 3: Stokes FxU, 524288 points in 64 Gaussian clusters")
    M = 524288
    xt, xs, f = W.clusters(M, M, seed=0, k0=3)
    r, o = dense_pair(S.STOKES3D_FXU, xt, xs, None, f, stream, reps=1 if quick else 2)
    sub = 512
    ref = oracle.eval(S.STOKES3D_FXU, xt[:3 * sub], xs, None, f)
    q = np.arange(3 * sub)
    r["mixed_vs_oracle_subset"] = dict(zip(("max_rel", "rms_rel"), (float("%.3g" % e) for e in W.rel_errors(
        o["mixed"][q], ref))))
    r["fp64_vs_oracle_subset"] = dict(zip(("max_rel", "rms_rel"), (float("%.3g" % e) for e in W.rel_errors(
        o["fp64"][q], ref))))
    r["oracle_subset_targets"] = sub
    add_roofline(r, S.STOKES3D_FXU)
    if cpu:
        log("config 3: reference CPU Eval on a target sample")
        r["cpu_reference"] = cpu_baseline(S.STOKES3D_FXU, xt, xs, None, f, cpu_seconds)
    out["c3_stokes_fxu_clusters_512k"] = r
    log("config 4: batched Laplace DxU, 4096 torus blocks")
    B = W.torus_blocks(64, 64, seed=0, k0=1)
    r, _ = plan_pair(S.LAPLACE3D_DXU, B, stream, reps=5)
    add_roofline(r, S.LAPLACE3D_DXU)
    if cpu:
        log("config 4: reference CPU per-block Eval loop on a block sample")
        r["cpu_reference"] = cpu_blocks_baseline(S.LAPLACE3D_DXU, B, cpu_seconds, mode=0)
    out["c4_batched_laplace_dxu_4096"] = r
    log("config 4 blocks: drop-in KernelMatrix per block, host arrays in and out, 8 host threads")
    km = kernel_matrix_end_to_end(S.LAPLACE3D_DXU, B, 512)
    if cpu:
        km["cpu_reference"] = cpu_blocks_baseline(S.LAPLACE3D_DXU, B, cpu_seconds, mode=1)
    out["c4_kernel_matrix_end_to_end"] = km
    log("HBM reference: device-to-device copy and fill")
    out["hbm_copy_4GiB"] = hbm_copy_reference(stream)
    log("config 4 blocks: near-field matrices (KNearSubtract, GEMV), Laplace DxU")
    out["c4_near_matrix_laplace_dxu_4096"] = near_ops(S.LAPLACE3D_DXU, B, stream, reps=3)
    log("config 4 surface: near lists (BuildNearList)")
    out["c4_near_list_4096"] = near_list_build(B, 0.05, stream)
    log("config 5: batched Stokes DxU, 2048 torus blocks, 100 applies in one hipGraph")
    B = W.torus_blocks(64, 32, seed=0, k0=3)
    r = graph_pair(S.STOKES3D_DXU, B, stream, nsteps=20 if quick else 100)
    add_roofline(r, S.STOKES3D_DXU)
    if cpu:
        log("config 5: reference CPU per-block Eval loop on a block sample")
        r["cpu_reference"] = cpu_blocks_baseline(S.STOKES3D_DXU, B, cpu_seconds, mode=0)
    out["c5_stokes_dxu_graph_2048x100"] = r
    log("config 5 blocks: near-field matrices (KNearSubtract, GEMV), Stokes DxU")
    out["c5_near_matrix_stokes_dxu_2048"] = near_ops(S.STOKES3D_DXU, B, stream, reps=2)
    log("other micro-kernels: FP64 and mixed Eval, 2^17 x 2^17 uniform cube")
    out["other_kernels_cube_128k"] = other_kernels(stream, reps=1 if quick else 2)
    return out


def config_only(name, stream):
    """One configuration's FP64 and mixed legs at default launch shapes (for
    traces and counters): c1-c3 one timed Eval each after a warm-up, c5 a
    20-apply graph per precision."""
    if name == "c1":
        xt, xs, f = W.cube(20000, 20000, seed=0)
        r, _ = dense_pair(S.LAPLACE3D_FXU, xt, xs, None, f, stream, reps=1)
    elif name == "c2":
        xt, xs, nrm, f = W.sphere(1 << 20, 1 << 20, seed=0)
        r, _ = dense_pair(S.LAPLACE3D_DXU, xt, xs, nrm, f, stream, reps=1)
    elif name == "c3":
        xt, xs, f = W.clusters(524288, 524288, seed=0, k0=3)
        r, _ = dense_pair(S.STOKES3D_FXU, xt, xs, None, f, stream, reps=1)
    else:
        B = W.torus_blocks(64, 32, seed=0, k0=3)
        r = graph_pair(S.STOKES3D_DXU, B, stream, nsteps=20, tune=False)
    return r


HELMHOLTZ_MU = 3.0  # wavenumber of the Helmholtz lines (about half a wavelength across the unit cube)


def other_kernels(stream, reps):
    """FP64 and mixed Eval of the twelve other micro-kernels on a
    2^17-target x 2^17-source uniform cube (Helmholtz at wavenumber
    HELMHOLTZ_MU): rates, the reference's FLOPS() count per interaction times
    the rate, and mixed-vs-FP64 errors."""
    N = 1 << 17
    out = {}
    for k in range(4, S.NUM_KERNELS):
        _, K0, K1, ND, flops, _, name = S.kernel_info(k)
        xt, xs, f = W.cube(N, N, seed=k, k0=K0)
        nrm = W.normal(k, 3, N * 3) if ND else None
        ctx = HELMHOLTZ_MU if k in S.CTX_KERNELS else None
        r, _ = dense_pair(k, xt, xs, nrm, f, stream, reps=reps, ctx=ctx)
        if ctx is not None:
            r["wavenumber"] = ctx
        for prec in ("fp64", "mixed"):
            r[prec]["ref_TFLOP_per_s"] = round(r[prec]["G_int_per_s"] * flops / 1e3, 2)
        r["ref_flops_per_interaction"] = flops
        out[name] = r
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20, help="sources = targets")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU work per baseline run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations")
    ap.add_argument("--no-tune", action="store_true", help="use default launch shapes (no autotuning)")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the mixed headline steps (for hardware-counter runs): no FP64 legs, configs, CPU")
    ap.add_argument("--headline-precision", choices=["mixed", "fp64"], default="mixed",
                    help="with --headline-only: which path to run")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    ap.add_argument("--near-only", action="store_true",
                    help="only the near-field operator and near-list lines of configs 4 and 5 (for kernel traces)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the other micro-kernels' FP64 and mixed Eval lines (for kernel traces)")
    ap.add_argument("--batched-only", choices=["mixed", "fp64"], default=None,
                    help="only applies of the config-4 batched plan (4096 torus blocks, Laplace DxU) in this "
                         "precision, at --plan-shape or the default shape, without tuning (for counters)")
    ap.add_argument("--plan-shape", default=None, metavar="SPL,TPL,CAP",
                    help="launch shape of the --batched-only plan (sources per lane, targets per lane, cap)")
    ap.add_argument("--config-only", choices=["c1", "c2", "c3", "c5"], default=None,
                    help="only the FP64 and mixed legs of one other configuration, untuned, without CPU or "
                         "oracle (for kernel traces and counters)")
    ap.add_argument("--shape", default=None, metavar="TPL,CAP",
                    help="pin the mixed headline launch shape (targets per lane, workgroups-per-CU cap) "
                         "instead of tuning it")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.config_only:
        stream = ctypes.c_void_p()
        S.check(S.lib().sctl_amd_stream_create(ctypes.byref(stream)))
        if args.copy_mb:
            calibration_copy(args.copy_mb, stream)
        print(json.dumps({"config_only": args.config_only, "result": config_only(args.config_only, stream)}),
              flush=True)
        return

    if args.near_only:
        stream = ctypes.c_void_p()
        S.check(S.lib().sctl_amd_stream_create(ctypes.byref(stream)))
        if args.copy_mb:
            calibration_copy(args.copy_mb, stream)
        log("config 4 blocks: near-field matrices and near lists")
        B = W.torus_blocks(64, 64, seed=0, k0=1)
        out = {"c4_near_matrix_laplace_dxu_4096": near_ops(S.LAPLACE3D_DXU, B, stream, reps=3),
               "c4_near_list_4096": near_list_build(B, 0.05, stream, cpu=False)}
        log("config 5 blocks: near-field matrices, Stokes DxU")
        out["c5_near_matrix_stokes_dxu_2048"] = near_ops(S.STOKES3D_DXU, W.torus_blocks(64, 32, seed=0, k0=3), stream,
                                                         reps=2)
        print(json.dumps({"near_only": True, "configs": out}), flush=True)
        return

    if args.batched_only:
        stream = ctypes.c_void_p()
        S.check(S.lib().sctl_amd_stream_create(ctypes.byref(stream)))
        B = W.torus_blocks(64, 64, seed=0, k0=1)
        nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
        prec = S.MIXED if args.batched_only == "mixed" else S.FP64
        P = S.Plan(S.LAPLACE3D_DXU, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
        P.set_geometry(*[S.DeviceArray.from_host(B[a]) for a in ("r_trg", "r_src", "n_src")], stream=stream)
        if args.plan_shape:
            P.set_shape(*(int(x) for x in args.plan_shape.split(",")))
        dv = S.DeviceArray.from_host(B["v_src"])
        o = S.DeviceArray.from_host(np.zeros(nt))
        log("config 4 batched plan, %s, shape %s: 3 applies" % (args.batched_only, P.shape()))
        P.apply(dv, o, stream=stream)
        S.synchronize(stream)
        ms = time_once(lambda: P.apply(dv, o, stream=stream), stream, 2)
        print(json.dumps({"batched_only": args.batched_only, "shape": list(P.shape()),
                          "interactions_per_apply": int(P.interactions()), "ms_per_apply": round(ms, 4)}), flush=True)
        return

    if args.kernels_only:
        stream = ctypes.c_void_p()
        S.check(S.lib().sctl_amd_stream_create(ctypes.byref(stream)))
        log("other micro-kernels: FP64 and mixed Eval, 2^17 x 2^17 uniform cube")
        print(json.dumps({"kernels_only": True, "configs": {"other_kernels_cube_128k": other_kernels(stream, 2)}}),
              flush=True)
        return

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

    if args.copy_mb:
        calibration_copy(args.copy_mb, stream)
    tuned = {}
    if args.shape:
        tpl, cap = (int(x) for x in args.shape.split(","))
        S.set_dense_shape(K, N, N, tpl, cap, precision=S.MIXED)
        tuned["mixed"] = {"targets_per_lane": tpl, "wg_per_cu_cap": cap, "pinned": True}
    if not args.headline_only and not args.no_tune:
        log("autotuning the launch shapes (not timed)")
        for name, prec, dig in (("mixed", S.MIXED, -1), ("fp64", S.FP64, -1), ("fp64_digits14", S.FP64, 14)):
            if name in tuned:
                continue
            tpl, cap, ms = S.tune_eval(K, dev[0], dev[1], None, dev[3], out, precision=prec, digits=dig,
                                       stream=stream)
            tuned[name] = {"targets_per_lane": tpl, "wg_per_cu_cap": cap, "ms": round(ms, 3)}
        out.upload(np.zeros(N))
    hs = S.dense_shape(K, N, N, precision=S.MIXED)
    headline_kernel = "k_eval_mixed<Laplace3D_FxU, TPL %d>" % hs[0]
    log("mixed precision: %d warmup + %d timed steps (%s)" % (args.warmup, args.steps, headline_kernel))
    hp = S.FP64 if (args.headline_only and args.headline_precision == "fp64") else S.MIXED
    mx = timed_evals(K, dev, out, hp, -1, args.steps, args.warmup, stream)
    if args.headline_only:
        v = inter * args.steps / (mx["total_ms"] * 1e-3) / 1e9
        print(json.dumps({"metric": METRIC, "value": round(v, 2), "unit": "G interactions/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(mx["total_ms"] / args.steps, 3), "headline_only": True,
                          "precision": args.headline_precision}), flush=True)
        return
    log("fp64 (digits 15): timed steps")
    f64 = timed_evals(K, dev, out, S.FP64, -1, args.steps, min(args.warmup, 1), stream)
    log("fp64 (digits 14): timed steps")
    f64b = timed_evals(K, dev, out, S.FP64, 14, args.steps, min(args.warmup, 1), stream)

    log("mixed with one FP32 Newton step after v_rsq_f32 (SCTL_AMD_MIXED_NEWTON): timed steps")
    mxn = timed_evals(K, dev, out, S.MIXED, -1, args.steps, min(args.warmup, 1), stream, flags=S.MIXED_NEWTON)

    log("accuracy: mixed vs fp64 on the full output")
    o64 = S.DeviceArray.from_host(np.zeros(N))
    omx = S.DeviceArray.from_host(np.zeros(N))
    S.eval(K, dev[0], dev[1], None, dev[3], o64, precision=S.FP64, stream=stream)
    S.eval(K, dev[0], dev[1], None, dev[3], omx, precision=S.MIXED, stream=stream)
    S.synchronize(stream)
    e_inf, e_rms = W.rel_errors(omx.download(), o64.download())
    omn = S.DeviceArray.from_host(np.zeros(N))
    S.eval(K, dev[0], dev[1], None, dev[3], omn, precision=S.MIXED, stream=stream, flags=S.MIXED_NEWTON)
    S.synchronize(stream)
    n_inf, n_rms = W.rel_errors(omn.download(), o64.download())

    def rate(r):
        return inter * args.steps / (r["total_ms"] * 1e-3) / 1e9

    v_mx, v_64, v_64b, v_mxn = rate(mx), rate(f64), rate(f64b), rate(mxn)
    kern_ms = mx["kernel_ms"] / max(mx["launches"], 1)
    achieved = FLOP_PER_INT[K] * inter / (kern_ms * 1e-3) / 1e12
    step_tflops = FLOP_PER_INT[K] * v_mx * 1e9 / 1e12
    k64_ms = f64["kernel_ms"] / max(f64["launches"], 1)
    achieved64 = FLOP_PER_INT[K] * inter / (k64_ms * 1e-3) / 1e12

    configs = None
    if not args.no_configs:
        try:
            configs = other_configs(stream, cpu=not args.no_cpu_baseline)
        except Exception as ex:
            log("other configurations failed:", repr(ex))
            configs = {"error": repr(ex)}

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
        "mixed_newton": {"value": round(v_mxn, 2), "ms_per_step": round(mxn["total_ms"] / args.steps, 3),
                         "max_rel_err_vs_fp64": float("%.3g" % n_inf), "rms_rel_err_vs_fp64": float("%.3g" % n_rms),
                         "speedup_vs_fp64": round(v_mxn / v_64, 3),
                         "pct_fp32_valu_peak": round(100 * FLOP_PER_INT[K] * v_mxn / 1e3 / PEAK_FP32_TFLOPS, 1),
                         "form": "v_rsq_f32 and one FP32 Newton step per pair (SCTL_AMD_MIXED_NEWTON)"},
        "speedup_vs_fp64": round(v_mx / v_64, 3),
        "speedup_vs_fp64_digits14": round(v_mx / v_64b, 3),
        "pct_fp32_valu_peak": round(100 * step_tflops / PEAK_FP32_TFLOPS, 1),
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic(N),
                     "traffic_source": "not measured in this run: FETCH_SIZE/WRITE_SIZE of a separate rocprofv3 "
                                       "run of the same launch, copy-calibrated, profiles/" +
                                       os.path.basename(TRAFFIC_FILE),
                     "kernel": headline_kernel, "mean_launch_ms": round(kern_ms, 3),
                     "flop_per_interaction": FLOP_PER_INT[K]},
        "autotuned_shapes": tuned or None,
        "cpu_baseline": cpu,
        "configs": configs,
    }
    log("wall %.2fs for %d mixed steps (events %.1f ms)" % (mx["wall_s"], args.steps, mx["total_ms"]))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
