"""Repeated plan applications captured as one hipGraph, and the device
generator.  Every replay is checked against the same sequence run eagerly,
with eager work and synchronisations between the replays."""
import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

DENS_STREAM = 25


def test_device_generator_matches_host(gpu):
    for seed, stream, n in ((0, 25, 1), (3, 25, 1001), (17, 5, 4096)):
        d = gpu.DeviceArray(n)
        gpu.gen_normal(seed, stream, d)
        got = d.download()
        want = W.normal(seed, stream, n)
        assert np.allclose(got, want, rtol=0, atol=1e-13 * max(1.0, np.abs(want).max()))


@pytest.mark.parametrize("k", [3, 8], ids=["StokesDxU", "StokesFSxU"])
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_graph_replays_match_eager(gpu, k, prec, tol):
    """Stokes DxU (3-vector densities, normals) and Stokes FSxU (4 densities
    per source)."""
    _, K0, K1, ND, _ = oracle.dims(k)
    B = W.torus_blocks(8, 8, seed=1, k0=K0, ns_range=(64, 1024), nt_range=(16, 300))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = gpu.Plan(k, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"] if ND else None)
    nsteps, seed = 5, 100
    # eager: the same steps one call at a time
    v_src = gpu.DeviceArray(ns * K0)
    eager = gpu.DeviceArray.from_host(np.zeros(nt * K1))
    for s in range(nsteps):
        gpu.gen_normal(seed + s, DENS_STREAM, v_src)
        P.apply(v_src, eager)
    ref = eager.download()
    # the oracle on the summed densities (linearity), per block
    dens = sum(W.normal(seed + s, DENS_STREAM, ns * K0) for s in range(nsteps))
    want = np.zeros(nt * K1)
    for b in range(B["trg_cnt"].size):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        want[K1 * t0:K1 * (t0 + tc)] = oracle.eval(
            k, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
            B["n_src"][3 * s0:3 * (s0 + sc)] if ND else None, dens[K0 * s0:K0 * (s0 + sc)])
    assert W.rel_errors(ref, want)[0] < tol
    # graph: replay several times, eager work in between
    v_trg = gpu.DeviceArray.from_host(np.zeros(nt * K1))
    G = P.graph(nsteps, seed, DENS_STREAM, v_src, v_trg)
    assert G.interactions() == nsteps * P.interactions()
    for rep in range(3):
        v_trg.upload(np.zeros(nt * K1))
        G.launch()
        gpu.synchronize()
        got = v_trg.download()
        assert W.rel_errors(got, ref)[0] < (1e-15 if prec == 0 else 1e-12), "replay %d" % rep
        # eager work between replays
        tmp = gpu.DeviceArray.from_host(np.zeros(nt * K1))
        P.apply(v_src, tmp)
        gpu.synchronize()
    # accumulation across two launches without clearing
    v_trg.upload(np.zeros(nt * 3))
    G.launch()
    G.launch()
    gpu.synchronize()
    assert W.rel_errors(v_trg.download(), 2 * ref)[0] < (1e-15 if prec == 0 else 1e-12)


def test_graph_refused_after_plan_changes(gpu):
    """A graph bakes in its plan's work items, launch shape and context: after
    set_shape, tune or set_ctx on the plan it is refused, not replayed with
    freed or stale state; a graph captured again replays correctly."""
    B = W.torus_blocks(4, 4, seed=2, k0=1, ns_range=(64, 512), nt_range=(16, 200))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = gpu.Plan(1, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=0)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
    P.set_shape(1, 1, 0)
    v_src = gpu.DeviceArray(ns)
    v_trg = gpu.DeviceArray.from_host(np.zeros(nt))
    G = P.graph(2, 11, DENS_STREAM, v_src, v_trg)
    G.launch()
    gpu.synchronize()
    first = v_trg.download()
    P.set_shape(2, 2, 0)  # two targets per lane: new work items
    with pytest.raises(gpu.SctlAmdError, match="capture it again"):
        G.launch()
    scratch = gpu.DeviceArray(nt)
    P.tune(v_src, scratch)
    with pytest.raises(gpu.SctlAmdError, match="capture it again"):
        G.launch()
    G2 = P.graph(2, 11, DENS_STREAM, v_src, v_trg)
    v_trg.upload(np.zeros(nt))
    G2.launch()
    gpu.synchronize()
    assert W.rel_errors(v_trg.download(), first)[0] < 1e-13
    H = gpu.Plan(12, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=0)
    H.set_geometry(B["r_trg"], B["r_src"], None)
    H.set_ctx(1.5)
    h_src, h_trg = gpu.DeviceArray(ns * 2), gpu.DeviceArray.from_host(np.zeros(nt * 2))
    GH = H.graph(1, 3, DENS_STREAM, h_src, h_trg)
    GH.launch()
    H.set_ctx(2.5)
    with pytest.raises(gpu.SctlAmdError, match="capture it again"):
        GH.launch()
    gpu.synchronize()
