"""GPU parity of the batched near-field plan (one persistent launch over
uneven blocks) against a per-block oracle loop -- the way BoundaryIntegralOp
evaluates its near blocks one by one (boundary_integral.txx:1169-1238,
1342-1364)."""
import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu


MU = 2.5  # wavenumber of the Helmholtz kernels


def ctx_of(k):
    return MU if k >= 12 else None


def per_block_oracle(k, B, v_src, v_trg0=None):
    _, K0, K1, ND, _ = oracle.dims(k)
    out = np.zeros(B["r_trg"].size // 3 * K1) if v_trg0 is None else v_trg0.copy()
    for b in range(B["trg_cnt"].size):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        if not tc or not sc:
            continue
        out[t0 * K1:(t0 + tc) * K1] = oracle.eval(
            k, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
            B["n_src"][3 * s0:3 * (s0 + sc)] if ND else None, v_src[s0 * K0:(s0 + sc) * K0],
            v_trg=out[t0 * K1:(t0 + tc) * K1], ctx=ctx_of(k))
    return out


def small_torus(k, nu=8, nv=8, seed=0):
    _, K0, _, _, _ = oracle.dims(k)
    return W.torus_blocks(nu, nv, seed=seed, k0=K0, ns_range=(64, 1024), nt_range=(1, 300))


KERNELS = list(range(16))
IDS = ["LaplaceFxU", "LaplaceDxU", "StokesFxU", "StokesDxU", "LaplaceFxdU", "LaplaceDxdU", "LaplaceFxd2U",
       "StokesFxT", "StokesFSxU", "StokesFxUP", "BiotSavartFxU", "BiotSavartFxdU", "HelmholtzFxU", "HelmholtzDxU",
       "HelmholtzFxdU", "HelmholtzDiffFxdU"]


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_plan_matches_per_block_oracle(gpu, k, prec, tol):
    B = small_torus(k)
    _, K0, K1, ND, _ = oracle.dims(k)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = gpu.Plan(k, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec, ctx=ctx_of(k))
    assert P.interactions() == W.interactions(B)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"] if ND else None)
    v0 = np.random.default_rng(2).standard_normal(nt * K1)
    out = P.apply(B["v_src"], v0.copy())
    ref = per_block_oracle(k, B, B["v_src"], v0)
    assert W.rel_errors(out, ref)[0] < tol
    # a second apply accumulates again
    out2 = P.apply(B["v_src"], out.copy())
    assert W.rel_errors(out2 - v0, 2 * (ref - v0))[0] < tol


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["fp64", "mixed"])
def test_plan_ragged_and_empty_blocks(gpu, prec, tol):
    """Empty blocks, single-target and single-source blocks, blocks listed out
    of order and targets not covered by any block (left untouched)."""
    rng = np.random.default_rng(9)
    ns, nt = 2000, 1200
    xs, xt, n = rng.random(ns * 3), rng.random(nt * 3), rng.standard_normal(ns * 3)
    xt[:30] = xs[:30]  # coincident pairs inside block 0
    trg_dsp = np.array([0, 700, 100, 150, 300, 301, 1000], dtype=np.int64)
    trg_cnt = np.array([100, 64, 50, 0, 1, 17, 150], dtype=np.int64)
    src_dsp = np.array([0, 100, 200, 0, 900, 1000, 1500], dtype=np.int64)
    src_cnt = np.array([100, 1, 700, 10, 33, 1000, 500], dtype=np.int64)
    B = dict(r_trg=xt, r_src=xs, n_src=n, trg_dsp=trg_dsp, trg_cnt=trg_cnt, src_dsp=src_dsp, src_cnt=src_cnt)
    for k in (0, 1, 2, 5, 8, 11, 13):
        _, K0, K1, ND, _ = oracle.dims(k)
        v = rng.uniform(-1, 1, ns * K0)
        P = gpu.Plan(k, trg_dsp, trg_cnt, src_dsp, src_cnt, nt, ns, precision=prec, ctx=ctx_of(k))
        P.set_geometry(xt, xs, n if ND else None)
        out = P.apply(v, np.zeros(nt * K1))
        ref = per_block_oracle(k, B, v)
        assert np.isfinite(out).all()
        assert W.rel_errors(out, ref)[0] < tol
        covered = np.zeros(nt, bool)
        for d, c in zip(trg_dsp, trg_cnt):
            covered[d:d + c] = True
        assert np.all(out.reshape(nt, K1)[~covered] == 0)


def test_plan_device_pointers_and_torus_config(gpu):
    """The 64x64-patch torus configuration (4096 blocks) through device
    pointers: FP64 against mixed precision, and a sample of blocks against the
    oracle."""
    B = W.torus_blocks(64, 64, seed=0, k0=1)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    out = {}
    for prec in (0, 1):
        P = gpu.Plan(1, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=prec)
        P.set_geometry(gpu.DeviceArray.from_host(B["r_trg"]), gpu.DeviceArray.from_host(B["r_src"]),
                       gpu.DeviceArray.from_host(B["n_src"]))
        o = gpu.DeviceArray.from_host(np.zeros(nt))
        P.apply(gpu.DeviceArray.from_host(B["v_src"]), o)
        out[prec] = o.download()
    assert W.rel_errors(out[1], out[0])[0] < 1e-6
    for b in range(0, 4096, 257):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        ref = oracle.eval(1, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
                          B["n_src"][3 * s0:3 * (s0 + sc)], B["v_src"][s0:s0 + sc])
        assert W.rel_errors(out[0][t0:t0 + tc], ref)[0] < 1e-13


def test_plan_helmholtz_needs_its_wavenumber(gpu):
    B = small_torus(12, nu=2, nv=2)
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = gpu.Plan(12, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    P.set_geometry(B["r_trg"], B["r_src"])
    with pytest.raises(gpu.SctlAmdError):
        P.apply(B["v_src"], np.zeros(nt * 2))
    P.set_ctx(MU)
    out = P.apply(B["v_src"], np.zeros(nt * 2))
    assert W.rel_errors(out, per_block_oracle(12, B, B["v_src"]))[0] < 1e-13
