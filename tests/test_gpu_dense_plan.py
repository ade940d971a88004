"""Dense plans (resident geometry, repeated applies): the same values as
sctl_amd_eval, the oracle's, accumulation, device pointers, and a hipGraph of
applies with densities regenerated on the device against the eager sequence."""
import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [0, 1, 2, 3, 8, 14]
IDS = ["LaplaceFxU", "LaplaceDxU", "StokesFxU", "StokesDxU", "StokesFSxU", "HelmholtzFxdU"]


def case(k, nt, ns, seed):
    _, K0, K1, ND, _ = oracle.dims(k)
    xt, xs, v = W.cube(ns, nt, seed=seed, k0=K0)  # cube(n_src, n_trg)
    n = W.normal(seed, 9, ns * 3) if ND else None
    return xt, xs, n, v, K1


@pytest.mark.parametrize("k", CASES, ids=IDS)
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
@pytest.mark.parametrize("nt,ns", [(3000, 2500), (70000, 60000)])
def test_dense_plan_equals_eval(gpu, k, prec, tol, nt, ns):
    xt, xs, n, v, K1 = case(k, nt, ns, seed=nt % 97)
    mu = 2.5 if k >= 12 else None
    P = gpu.DensePlan(k, xt, xs, n, precision=prec, ctx=mu)
    assert P.shape()[0] == prec
    out = P.apply(v)
    ref = gpu.eval(k, xt, xs, n, v, precision=prec, ctx=mu)
    assert np.array_equal(out, ref)
    sub = 300
    want = oracle.eval(k, xt[:3 * sub], xs, n, v, ctx=mu)
    assert W.rel_errors(out[:sub * K1], want)[0] < tol
    # a second apply accumulates; new densities reuse the resident geometry
    v2 = W.normal(3, 4, v.size)
    out2 = P.apply(v2, out.copy())
    assert W.rel_errors(out2 - out, gpu.eval(k, xt, xs, n, v2, precision=prec, ctx=mu))[0] < 1e-12


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_dense_plan_device_pointers_and_graph(gpu, prec):
    k, nt, ns, nsteps, seed, sid = 3, 40000, 30000, 6, 11, 25
    xt, xs, n, _, K1 = case(k, nt, ns, seed=5)
    d = [gpu.DeviceArray.from_host(a) for a in (xt, xs, n)]
    P = gpu.DensePlan(k, d[0], d[1], d[2], precision=prec)
    v_src = gpu.DeviceArray(ns * 3)
    eager = gpu.DeviceArray.from_host(np.zeros(nt * K1))
    for s in range(nsteps):
        gpu.gen_normal(seed + s, sid, v_src)
        P.apply(v_src, eager)
    gpu.synchronize()
    ref = eager.download()
    v_trg = gpu.DeviceArray.from_host(np.zeros(nt * K1))
    G = P.graph(nsteps, seed, sid, v_src, v_trg)
    assert G.interactions() == nsteps * nt * ns
    for rep in range(3):
        v_trg.upload(np.zeros(nt * K1))
        G.launch()
        gpu.synchronize()
        assert np.array_equal(v_trg.download(), ref), "replay %d" % rep
        P.apply(v_src, gpu.DeviceArray.from_host(np.zeros(nt * K1)))  # eager work between replays
        gpu.synchronize()
    dens = sum(W.normal(seed + s, sid, ns * 3) for s in range(nsteps))
    want = oracle.eval(k, xt[:3 * 200], xs, n, dens)
    assert W.rel_errors(ref[:200 * K1], want)[0] < (1e-13 if prec == 0 else 1e-6)


def test_dense_plan_errors(gpu):
    xt, xs, n, v, _ = case(12, 10, 10, 1)
    with pytest.raises(gpu.SctlAmdError):
        gpu.DensePlan(12, xt, xs)  # Helmholtz without its wavenumber
    with pytest.raises(gpu.SctlAmdError):
        gpu.DensePlan(16, xt, xs)
    with pytest.raises(gpu.SctlAmdError):
        gpu.DensePlan(1, xt, xs)  # a double layer without normals
    P = gpu.DensePlan(0, xt[:0], xs)  # no targets: applies do nothing
    assert P.apply(v[:10]).size == 0
