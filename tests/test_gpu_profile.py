"""The profile counters (sctl_amd_profile_*): the counterpart of the
reference profiler's FLOP counter, which Eval and KernelMatrix increment by
Ns*Nt*FLOPS() on every call, zero-size calls included
(generic-kernel.txx:210,224; profile.txx:466-473).  Over a mix of Eval (FP64
and mixed, all kernels), KernelMatrix, batched and dense plan applies and
graph launches, the counters must equal the sums worked out here from the
call sizes and the reference's FLOPS() table."""
import numpy as np
import pytest

from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

# FLOPS() of kernel_functions.hpp, kernel ids 0..15
REF_FLOPS = [6, 14, 23, 26, 11, 22, 25, 39, 26, 26, 15, 51, 20, 29, 35, 41]


def test_profile_counts_interactions_and_reference_flops(gpu):
    gpu.synchronize()
    gpu.profile_reset()
    want_int = want_flops = 0

    def add(k, n):
        nonlocal want_int, want_flops
        want_int += n
        want_flops += n * REF_FLOPS[k]

    rng = np.random.default_rng(5)
    for k in range(16):
        _, K0, K1, ND, _, _, _ = gpu.kernel_info(k)
        ctx = 1.5 if k >= 12 else None
        for nt, ns in ((37, 211), (0, 50), (12, 0), (300, 1)):
            xt, xs = rng.random(3 * nt), rng.random(3 * ns) + 2.0
            n = rng.standard_normal(3 * ns) if ND else None
            v = rng.standard_normal(K0 * ns)
            for prec in (gpu.FP64, gpu.MIXED):
                gpu.eval(k, xt, xs, n, v, precision=prec, ctx=ctx)
                add(k, nt * ns)
            gpu.kernel_matrix(k, xt, xs, n, ctx=ctx)
            add(k, nt * ns)
    # a batched plan: one launch counts the plan's interactions
    B = W.torus_blocks(4, 4, seed=3, k0=3, ns_range=(64, 256), nt_range=(16, 64))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = gpu.Plan(3, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns, precision=1)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
    assert P.interactions() == int(np.sum(B["trg_cnt"] * B["src_cnt"]))
    v_src = gpu.DeviceArray.from_host(B["v_src"])
    v_trg = gpu.DeviceArray.from_host(np.zeros(3 * nt))
    for _ in range(2):
        P.apply(v_src, v_trg)
        add(3, P.interactions())
    G = P.graph(3, 1, 25, v_src, v_trg)  # capture itself counts nothing
    G.launch()
    add(3, 3 * P.interactions())
    # a dense plan
    xt, xs, f = W.cube(500, 700, seed=6)
    D = gpu.DensePlan(0, gpu.DeviceArray.from_host(xt), gpu.DeviceArray.from_host(xs), None, precision=1)
    out = gpu.DeviceArray.from_host(np.zeros(700))
    D.apply(gpu.DeviceArray.from_host(f), out)
    add(0, 500 * 700)
    gpu.synchronize()
    got = gpu.profile_read()
    assert got["interactions"] == want_int
    assert got["flops"] == want_flops
