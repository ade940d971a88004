"""Parity at the full sizes of the benchmark configurations, through
size-independent properties and sampled oracle checks (the oracle cannot run
the whole sums in test time):

  headline  Laplace FxU, 2^20 sphere sources x 2^20 shell targets, mixed:
            against the device's own FP64 result at 1e-6 (max and RMS) and a
            2048-target sample against the oracle;
  c3        Stokes FxU, 524 288 points in 64 Gaussian clusters, mixed and
            FP64: a 512-target sample against the oracle, mixed vs FP64, and
            linearity Eval(f) + Eval(g) = Eval(f + g);
  c5        Stokes DxU on 2048 torus blocks, 100 applies captured as one
            hipGraph with densities regenerated per step: the graph's sum
            equals the oracle's blocks on the summed densities (sampled
            blocks) and the eager sequence of applies."""
import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_headline_mixed_full_size(gpu):
    N = 1 << 20
    xt, xs, _, f = W.sphere(N, N, seed=0)
    d = [gpu.DeviceArray.from_host(a) for a in (xt, xs, f)]
    outs = {}
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        o = gpu.DeviceArray.from_host(np.zeros(N))
        gpu.eval(0, d[0], d[1], None, d[2], o, precision=prec)
        outs[name] = o.download()
    e_inf, e_rms = W.rel_errors(outs["mixed"], outs["fp64"])
    assert e_inf < 1e-6 and e_rms < 1e-6, (e_inf, e_rms)
    sub = 2048
    ref = oracle.eval(0, xt[:3 * sub], xs, None, f)
    assert W.rel_errors(outs["fp64"][:sub], ref)[0] < 1e-13
    assert W.rel_errors(outs["mixed"][:sub], ref)[0] < 1e-6


def test_c3_stokes_clusters_full_size(gpu):
    M = 524288
    xt, xs, f = W.clusters(M, M, seed=0, k0=3)
    g = W.normal(7, 3, M * 3)
    d = {k: gpu.DeviceArray.from_host(a) for k, a in (("xt", xt), ("xs", xs), ("f", f), ("g", g), ("fg", f + g))}
    outs = {}
    for name, prec in (("fp64", gpu.FP64), ("mixed", gpu.MIXED)):
        for dens in ("f", "g", "fg"):
            o = gpu.DeviceArray.from_host(np.zeros(M * 3))
            gpu.eval(2, d["xt"], d["xs"], None, d[dens], o, precision=prec)
            outs[name, dens] = o.download()
    e_inf, e_rms = W.rel_errors(outs["mixed", "f"], outs["fp64", "f"])
    assert e_inf < 1e-6 and e_rms < 1e-6, (e_inf, e_rms)
    for name, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
        assert W.rel_errors(outs[name, "f"] + outs[name, "g"], outs[name, "fg"])[0] < tol, name
    sub = 512
    ref = oracle.eval(2, xt[:3 * sub], xs, None, f)
    assert W.rel_errors(outs["fp64", "f"][:3 * sub], ref)[0] < 1e-13
    assert W.rel_errors(outs["mixed", "f"][:3 * sub], ref)[0] < 1e-6


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_c5_stokes_dxu_graph_full_size(gpu, prec, tol):
    k, nsteps, seed, stream_id = 3, 100, 0, 25
    B = W.torus_blocks(64, 32, seed=0, k0=3)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    assert B["trg_cnt"].size == 2048
    P = gpu.Plan(k, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
    P.set_geometry(gpu.DeviceArray.from_host(B["r_trg"]), gpu.DeviceArray.from_host(B["r_src"]),
                   gpu.DeviceArray.from_host(B["n_src"]))
    v_src = gpu.DeviceArray(ns * 3)
    v_trg = gpu.DeviceArray.from_host(np.zeros(nt * 3))
    G = P.graph(nsteps, seed, stream_id, v_src, v_trg)
    G.launch()
    gpu.synchronize()
    got = v_trg.download()
    assert np.isfinite(got).all()
    # eager sequence of the same applies
    eager = gpu.DeviceArray.from_host(np.zeros(nt * 3))
    for s in range(nsteps):
        gpu.gen_normal(seed + s, stream_id, v_src)
        P.apply(v_src, eager)
    gpu.synchronize()
    assert W.rel_errors(got, eager.download())[0] < (1e-15 if prec == 0 else 1e-12)
    # sampled blocks against the oracle on the summed densities (linearity)
    dens = sum(W.normal(seed + s, stream_id, ns * 3) for s in range(nsteps))
    for b in range(0, 2048, 97):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        ref = oracle.eval(k, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
                          B["n_src"][3 * s0:3 * (s0 + sc)], dens[3 * s0:3 * (s0 + sc)])
        assert W.rel_errors(got[3 * t0:3 * (t0 + tc)], ref)[0] < tol, b
