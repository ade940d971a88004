"""GPU near lists (BuildNearList, boundary_integral.txx:50-520, one process)
against the reference's own output (golden fixtures written by the compiled
reference, tests/golden/make_golden_near_list.py) and against the oracle's
brute-force restatement: integer and index outputs bit-identical, target
coordinates copied exactly.  The order of one target's pairs in the scatter
index is the oracle's (element order); against the reference it is compared
as a set (the reference sorts by target alone)."""
import numpy as np
import pytest

import oracle
import pins
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

KEYS = ("Xtrg_near", "Xn_trg_near", "near_elem_cnt", "near_elem_dsp", "near_scatter_index", "near_trg_cnt",
        "near_trg_dsp")


def _same(a, b):
    for k in KEYS:
        if (a[k] is None) != (b[k] is None):
            return k
        if a[k] is not None and not np.array_equal(a[k], b[k]):
            return k
    return None


@pytest.mark.parametrize("case", pins.NEAR_LIST_CASES)
def test_near_list_matches_reference_golden(gpu, case):
    inp, want = pins.near_list_golden_case(case)
    got = gpu.near_list(**inp)
    assert not pins.near_list_mismatches(got, want)
    orc = oracle.default().near_list(**inp)
    assert _same(got, orc) is None, _same(got, orc)


def _torus_nodes(n_u, n_v, seed):
    c = W.near_field_case(n_u=n_u, n_v=n_v, n_trg=10, seed=seed)
    return c["X_far"], c["elem_nds_cnt"], c["elem_nds_dsp"]


@pytest.mark.parametrize("radii", ["constant", "multiscale"])
def test_near_list_matches_oracle_torus(gpu, radii):
    xs, cnt, dsp = _torus_nodes(16, 12, seed=3)
    ns = xs.size // 3
    rad = np.full(ns, 0.12) if radii == "constant" else 0.002 * 100.0 ** W.uniform(3, 51, ns)
    u, v = W.uniform(3, 52, 3000) * 2 * np.pi, W.uniform(3, 53, 3000) * 2 * np.pi
    X, N = W.torus_point(u, v)
    xt = (X + W.uniform(3, 54, 3000, -0.1, 0.1)[:, None] * N).ravel()
    got = gpu.near_list(xt, xs, rad, cnt, dsp, Xn_trg=N.ravel())
    want = oracle.default().near_list(xt, xs, rad, cnt, dsp, Xn_trg=N.ravel())
    assert want["near_scatter_index"].size > 1000
    assert _same(got, want) is None, _same(got, want)


def test_near_list_device_pointers(gpu):
    inp, want = pins.near_list_golden_case("cloud")
    d = {k: gpu.DeviceArray.from_host(v) for k, v in inp.items() if k not in ("elem_nds_cnt", "elem_nds_dsp")}
    for k in ("elem_nds_cnt", "elem_nds_dsp"):
        d[k] = gpu.DeviceArray.from_host(inp[k], dtype=np.int64)
    got = gpu.near_list(d["Xtrg"], d["Xsrc"], d["src_radius"], d["elem_nds_cnt"], d["elem_nds_dsp"],
                        Xn_trg=d["Xn_trg"])
    assert not pins.near_list_mismatches(got, want)


def test_near_list_empty_inputs(gpu):
    r = gpu.near_list(np.zeros(0), np.zeros(6), np.ones(2), [2], [0])
    assert r["near_scatter_index"].size == 0 and list(r["near_elem_cnt"]) == [0]
    r = gpu.near_list(np.zeros(9), np.zeros(0), np.zeros(0), [0, 0], [0, 0])
    assert list(r["near_trg_cnt"]) == [0, 0, 0] and list(r["near_elem_dsp"]) == [0, 0]
    r = gpu.near_list(np.zeros(9), np.zeros(6), np.ones(2), [], [])
    assert r["near_scatter_index"].size == 0 and r["near_trg_cnt"].size == 3
    with pytest.raises(gpu.SctlAmdError):
        gpu.near_list(np.zeros(9), np.zeros(6), np.ones(2), [3], [0])  # node range outside the sources


def test_near_list_large_properties(gpu):
    """A config-4-sized surface (4096 patches, 8..24^2 nodes each, 200000
    targets): every listed pair is near (checked exactly on a sample) and
    complete for a sample of elements (numpy brute force over all targets)."""
    B = W.torus_blocks(64, 64, seed=1, k0=1, ns_range=(64, 576), nt_range=(1, 2))
    xs, cnt = B["r_src"], B["src_cnt"].astype(np.int64)
    dsp = B["src_dsp"].astype(np.int64)
    ns = xs.size // 3
    rad = np.full(ns, 0.04)
    u, v = W.uniform(5, 60, 200000) * 2 * np.pi, W.uniform(5, 61, 200000) * 2 * np.pi
    X, N = W.torus_point(u, v)
    xt = X + W.uniform(5, 62, 200000, -0.06, 0.06)[:, None] * N
    r = gpu.near_list(xt.ravel(), xs, rad, cnt, dsp)
    n = r["near_scatter_index"].size
    assert n > 100000
    assert r["near_elem_cnt"].sum() == n == r["near_trg_cnt"].sum()
    assert np.array_equal(np.sort(r["near_scatter_index"]), np.arange(n))
    xsr = xs.reshape(-1, 3)
    rng = np.random.default_rng(0)
    pair_elem = np.repeat(np.arange(cnt.size), r["near_elem_cnt"])
    xtn = r["Xtrg_near"].reshape(-1, 3)
    for p in rng.choice(n, 300, replace=False):
        e = pair_elem[p]
        d2 = ((xsr[dsp[e]:dsp[e] + cnt[e]] - xtn[p]) ** 2).sum(1)
        assert (d2 < rad[dsp[e]:dsp[e] + cnt[e]] ** 2).any()
    for e in rng.choice(cnt.size, 40, replace=False):
        nodes = xsr[dsp[e]:dsp[e] + cnt[e]]
        near = np.zeros(xt.shape[0], dtype=bool)
        for j in range(0, cnt[e], 64):
            d2 = ((xt[:, None, :] - nodes[None, j:j + 64, :]) ** 2).sum(-1)
            near |= (d2 < 0.04 ** 2).any(1)
        e0, ec = r["near_elem_dsp"][e], r["near_elem_cnt"][e]
        assert ec == near.sum()
        assert np.array_equal(xtn[e0:e0 + ec], xt[np.nonzero(near)[0]])
