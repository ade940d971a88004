"""Near-field operator, host side: the block partition of the C-ABI
(sctl_amd_near_blocks, host-only) and the K_near block offsets against the
oracle's restatement of BoundaryIntegralOp::SetupNear
(boundary_integral.txx:1009-1057)."""
import numpy as np
import pytest

import oracle
import sctl_amd as S
from sctl_amd import workloads as W


def _near_lists(rng, nelem, max_cnt):
    cnt = rng.integers(0, max_cnt + 1, nelem)
    cnt[rng.random(nelem) < 0.2] = 0
    dsp = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    return cnt, dsp


@pytest.mark.parametrize("nthreads", [1, 3, 8, 64, 1000])
def test_near_blocks_match_oracle(nthreads):
    rng = np.random.default_rng(nthreads)
    orc = oracle.default()
    for nelem, max_cnt in ((1, 5), (17, 40), (300, 900), (2000, 3)):
        cnt, dsp = _near_lists(rng, nelem, max_cnt)
        got = S.near_blocks(cnt, dsp, nthreads)
        want = orc.near_blocks(cnt, dsp, nthreads)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def test_near_blocks_cover_every_near_target_once():
    rng = np.random.default_rng(3)
    cnt, dsp = _near_lists(rng, 500, 200)
    e, t0, c = S.near_blocks(cnt, dsp, 16)
    hits = np.zeros(int(cnt.sum()), dtype=np.int64)
    for ee, tt, cc in zip(e, t0, c):
        assert cc >= 1 and tt + cc <= cnt[ee]
        hits[dsp[ee] + tt:dsp[ee] + tt + cc] += 1
    assert np.all(hits == 1)
    assert np.all(np.diff(e) >= 0)


def test_near_blocks_empty_and_errors():
    assert all(a.size == 0 for a in S.near_blocks([], [], 4))
    assert all(a.size == 0 for a in S.near_blocks([0, 0], [0, 0], 4))
    with pytest.raises(S.SctlAmdError):
        S.near_blocks([3], [0], 0)
    with pytest.raises(S.SctlAmdError):
        S.near_blocks([3, -1], [0, 3], 2)


@pytest.mark.parametrize("kernel,ndot", [(0, False), (1, False), (2, False), (3, False), (2, True), (3, True),
                                         (5, True), (8, False), (9, False), (11, False)])
def test_near_offsets_match_reference_layout(kernel, ndot):
    """Block offsets of the product's NearField against K_near_dsp of the
    oracle (boundary_integral.txx:1039-1057, :1187-1189)."""
    orc = oracle.default()
    _, K0, K1, _, _ = orc.dims(kernel)
    dof = K1 // 3 if ndot else K1
    c = W.near_field_case(n_u=5, n_v=4, n_trg=120, seed=2, k0=K0)
    mf = np.zeros(20, dtype=np.int32)
    mf[[3, 11]] = 1
    blocks = orc.near_blocks(c["near_elem_cnt"], c["near_elem_dsp"], 7)
    kc, kd, n = orc.near_layout(kernel, ndot, c["near_elem_cnt"], c["elem_nds_cnt"], mf)
    off, k_len = S.near_offsets(K0, dof, blocks, c["near_elem_cnt"], c["elem_nds_cnt"], mf)
    assert k_len == n
    for b, (e, t0, cnt) in enumerate(zip(*blocks)):
        if mf[e]:
            assert off[b] == -1
        else:
            assert off[b] == kd[e] * K0 * dof + t0 * dof * c["elem_nds_cnt"][e] * K0


def test_near_field_case_is_consistent():
    c = W.near_field_case(seed=4)
    n_pairs = int(c["near_elem_cnt"].sum())
    assert n_pairs == c["near_trg_cnt"].sum() == c["near_scatter_index"].size
    assert np.array_equal(np.sort(c["near_scatter_index"]), np.arange(n_pairs))
    assert c["Xtrg_near"].size == 3 * n_pairs
    # the quadrature weights integrate the torus area 4 pi^2 R r
    assert abs(c["wts_far"].sum() - 4 * np.pi ** 2 * 0.3) < 1e-10
    # every target-major pair points at an element-major pair of that target
    tm_trg = np.repeat(np.arange(c["near_trg_cnt"].size), c["near_trg_cnt"])
    xt = c["Xtrg"].reshape(-1, 3)
    xn = c["Xtrg_near"].reshape(-1, 3)
    assert np.array_equal(xn[c["near_scatter_index"]], xt[tm_trg])
