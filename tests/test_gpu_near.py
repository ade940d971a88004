"""GPU parity of BoundaryIntegralOp's near-field operator on the device:
KNearSubtract (boundary_integral.txx:1165-1240), the GEMV route of
ComputeNearInterac (:1324-1340) and ScatterForward + the per-target sum
(:1368-1379), against the oracle's restatement of the same steps on the
reference's own arrays (oracle/near_field.c).

Tolerances: the block matrices are FP64 KernelMatrix entries (hardware rsqrt
plus two Newton steps on the device, exact 1/sqrt in the oracle), so the
subtract is compared normwise at 1e-13 of the subtracted part; the GEMV sums
in another order than mat::gemm (1e-13 normwise); the scatter-accumulate adds
in the reference's order and must be bit-identical."""
import numpy as np
import pytest

import oracle
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

KERNEL_CASES = [(0, False), (1, False), (2, False), (3, False), (2, True), (3, True),
                (4, False), (5, False), (5, True), (6, False), (6, True), (7, False), (7, True), (8, False),
                (9, False), (10, False), (11, True), (12, False), (13, False), (14, False), (15, True)]
IDS = ["LaplaceFxU", "LaplaceDxU", "StokesFxU", "StokesDxU", "StokesFxU-ndot", "StokesDxU-ndot",
       "LaplaceFxdU", "LaplaceDxdU", "LaplaceDxdU-ndot", "LaplaceFxd2U", "LaplaceFxd2U-ndot", "StokesFxT",
       "StokesFxT-ndot", "StokesFSxU", "StokesFxUP", "BiotSavartFxU", "BiotSavartFxdU-ndot", "HelmholtzFxU",
       "HelmholtzDxU", "HelmholtzFxdU", "HelmholtzDiffFxdU-ndot"]
MU = 3.5  # wavenumber of the Helmholtz kernels


def ctx_of(kernel):
    return MU if kernel >= 12 else None


def _case(kernel, ndot, seed=1, n_u=6, n_v=5, n_trg=200, matrix_free=None, nthreads=5):
    orc = oracle.default()
    _, K0, K1, ND, _ = orc.dims(kernel)
    c = W.near_field_case(n_u=n_u, n_v=n_v, n_trg=n_trg, seed=seed, k0=K0)
    blocks = orc.near_blocks(c["near_elem_cnt"], c["near_elem_dsp"], nthreads)
    kc, kd, n = orc.near_layout(kernel, ndot, c["near_elem_cnt"], c["elem_nds_cnt"], matrix_free)
    return orc, c, blocks, kc, kd, n


def _near_field(S, kernel, ndot, c, blocks, matrix_free=None):
    return S.NearField(kernel, c["near_elem_cnt"], c["near_elem_dsp"], c["elem_nds_cnt_far"], c["elem_nds_dsp_far"],
                       c["Xtrg_near"], c["X_far"], c["Xn_far"], c["wts_far"], c["Xn_trg_near"],
                       matrix_free=matrix_free, trg_normal_dot=ndot, blocks=blocks, ctx=ctx_of(kernel))


def _oracle_subtract(orc, kernel, ndot, c, blocks, kc, kd, K, matrix_free=None):
    return orc.near_subtract(kernel, ndot, blocks, c["near_elem_dsp"], c["elem_nds_cnt"], c["elem_nds_cnt_far"],
                             c["elem_nds_dsp_far"], kc, kd, c["Xtrg_near"], c["Xn_trg_near"], c["X_far"],
                             c["Xn_far"], c["wts_far"], K, matrix_free=matrix_free, ctx=ctx_of(kernel))


@pytest.mark.parametrize("kernel,ndot", KERNEL_CASES, ids=IDS)
def test_k_near_subtract_matches_oracle(gpu, kernel, ndot):
    orc, c, blocks, kc, kd, n = _case(kernel, ndot)
    F = _near_field(gpu, kernel, ndot, c, blocks)
    assert F.k_len == n
    K0m = np.random.default_rng(kernel).standard_normal(n)
    ref = _oracle_subtract(orc, kernel, ndot, c, blocks, kc, kd, K0m)
    got = F.k_near_subtract(K0m.copy())
    assert W.rel_errors(got - K0m, ref - K0m)[0] < 1e-13
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref - K0m))


@pytest.mark.parametrize("kernel,ndot", [(1, False), (3, True)], ids=["LaplaceDxU", "StokesDxU-ndot"])
def test_k_near_subtract_device_pointers_and_matrix_free(gpu, kernel, ndot):
    """Device-resident K_near; two matrix-free elements (no matrix, untouched)."""
    mf = np.zeros(30, dtype=np.int32)
    mf[[4, 17]] = 1
    orc, c, blocks, kc, kd, n = _case(kernel, ndot, seed=3, matrix_free=mf)
    F = _near_field(gpu, kernel, ndot, c, blocks, matrix_free=mf)
    K0m = np.random.default_rng(7).standard_normal(n)
    ref = _oracle_subtract(orc, kernel, ndot, c, blocks, kc, kd, K0m, matrix_free=mf)
    dK = gpu.DeviceArray.from_host(K0m)
    F.k_near_subtract(dK)
    gpu.synchronize()
    got = dK.download()
    assert W.rel_errors(got - K0m, ref - K0m)[0] < 1e-13


@pytest.mark.parametrize("kernel,ndot", [(0, False), (2, True)], ids=["LaplaceFxU", "StokesFxU-ndot"])
def test_near_matrix_store_is_minus_subtract_from_zero(gpu, kernel, ndot):
    orc, c, blocks, kc, kd, n = _case(kernel, ndot, seed=4)
    F = _near_field(gpu, kernel, ndot, c, blocks)
    sub = F.plan.near_matrix(np.zeros(n), c["wts_far"], c["Xn_trg_near"], gpu.NEAR_SUBTRACT)
    st = F.plan.near_matrix(np.full(n, 7.0), c["wts_far"], c["Xn_trg_near"], gpu.NEAR_STORE)
    assert np.array_equal(st, -sub)


def test_near_matrix_leaves_gaps_untouched(gpu):
    """Explicit offsets with gaps between the block matrices and one block
    without a matrix (-1): entries outside the blocks keep their values."""
    S = gpu
    B = W.torus_blocks(3, 3, seed=2, k0=3, ns_range=(16, 200), nt_range=(1, 40))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = S.Plan(3, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"])
    sizes = B["trg_cnt"] * 3 * B["src_cnt"] * 3
    off = np.concatenate([[5], 5 + np.cumsum(sizes + 11)[:-1]])
    off[4] = -1
    k_len = P.near_layout(off)
    assert k_len == off[-1] + sizes[-1]
    K = np.full(k_len, 3.5)
    P.near_matrix(K, None, None, S.NEAR_STORE)
    inside = np.zeros(k_len, dtype=bool)
    for b in range(off.size):
        if off[b] >= 0:
            inside[off[b]:off[b] + sizes[b]] = True
    assert np.all(K[~inside] == 3.5)
    # each stored block equals the block's KernelMatrix, transposed to target-major
    orc = oracle.default()
    for b in (0, 8):
        t0, tc, s0, sc = (int(B[a][b]) for a in ("trg_dsp", "trg_cnt", "src_dsp", "src_cnt"))
        M = orc.kernel_matrix(3, B["r_trg"][3 * t0:3 * (t0 + tc)], B["r_src"][3 * s0:3 * (s0 + sc)],
                              B["n_src"][3 * s0:3 * (s0 + sc)])
        got = K[off[b]:off[b] + sizes[b]].reshape(tc * 3, sc * 3)
        assert W.rel_errors(got, M.T)[0] < 1e-13


@pytest.mark.parametrize("kernel,ndot", KERNEL_CASES, ids=IDS)
def test_near_gemv_matches_oracle(gpu, kernel, ndot):
    mf = np.zeros(30, dtype=np.int32)
    mf[9] = 1
    orc, c, blocks, kc, kd, n = _case(kernel, ndot, seed=5, matrix_free=mf)
    F = _near_field(gpu, kernel, ndot, c, blocks, matrix_free=mf)
    K = np.random.default_rng(8).standard_normal(n)
    n_near = int(c["near_elem_dsp"][-1] + c["near_elem_cnt"][-1])
    ref = orc.near_gemv(kernel, ndot, blocks, c["near_elem_dsp"], c["elem_nds_cnt"], c["elem_nds_dsp"], kc, kd, K,
                        c["F"], n_near)
    got = F.plan.near_gemv(K, c["F"], np.full(n_near * F.dof, np.nan))
    assert W.rel_errors(got, ref)[0] < 1e-13
    # the matrix-free element's rows are zero, as the reference writes them
    e0, e1 = c["near_elem_dsp"][9] * F.dof, (c["near_elem_dsp"][9] + c["near_elem_cnt"][9]) * F.dof
    assert np.all(got[e0:e1] == 0)


@pytest.mark.parametrize("dof", [1, 3, 4, 9])
@pytest.mark.parametrize("device", [False, True], ids=["host", "device"])
def test_near_accumulate_bit_identical(gpu, dof, device):
    S = gpu
    orc = oracle.default()
    c = W.near_field_case(n_u=7, n_v=5, n_trg=333, seed=6)
    n_pairs = c["near_scatter_index"].size
    rng = np.random.default_rng(dof)
    Un = rng.standard_normal(n_pairs * dof) * np.exp(rng.uniform(-20, 20, n_pairs * dof))
    U0 = rng.standard_normal(c["near_trg_cnt"].size * dof)
    ref = orc.near_accumulate(U0, dof, c["near_trg_cnt"], c["near_trg_dsp"], c["near_scatter_index"], Un)
    if device:
        d = [S.DeviceArray.from_host(a, dtype=np.int64)
             for a in (c["near_trg_cnt"], c["near_trg_dsp"], c["near_scatter_index"])]
        dU, dUn = S.DeviceArray.from_host(U0), S.DeviceArray.from_host(Un)
        S.near_accumulate(dU, dof, d[0], d[1], d[2], dUn)
        S.synchronize()
        got = dU.download()
    else:
        got = S.near_accumulate(U0.copy(), dof, c["near_trg_cnt"], c["near_trg_dsp"], c["near_scatter_index"], Un)
    assert np.array_equal(got, ref)


def test_near_accumulate_rejects_bad_indices(gpu):
    S = gpu
    with pytest.raises(S.SctlAmdError):
        S.near_accumulate(np.zeros(2), 1, [1, 1], [0, 1], [0, 5], np.zeros(2))
    with pytest.raises(S.SctlAmdError):
        S.near_accumulate(np.zeros(2), 1, [1, 2], [0, 1], [0, 1], np.zeros(2))
    with pytest.raises(S.SctlAmdError):
        S.near_accumulate(np.zeros(4), 2, [1, 1], [0, 1], [0, 1], np.zeros(4))


@pytest.mark.parametrize("kernel,ndot", [(1, False), (2, True), (6, True), (9, False)],
                         ids=["LaplaceDxU", "StokesFxU-ndot", "LaplaceFxd2U-ndot", "StokesFxUP"])
def test_compute_near_interac_end_to_end(gpu, kernel, ndot):
    """SetupNear's subtract then ComputeNearInterac (GEMV, ScatterForward,
    per-target sum) on the device against the oracle's chain of the same
    steps, starting from the same stand-in for NearInterac's matrix."""
    orc, c, blocks, kc, kd, n = _case(kernel, ndot, seed=9)
    F = _near_field(gpu, kernel, ndot, c, blocks)
    Kq = np.random.default_rng(10).standard_normal(n)  # stands in for the quadrature matrices
    K_ref = _oracle_subtract(orc, kernel, ndot, c, blocks, kc, kd, Kq)
    n_near = int(c["near_elem_dsp"][-1] + c["near_elem_cnt"][-1])
    Un_ref = orc.near_gemv(kernel, ndot, blocks, c["near_elem_dsp"], c["elem_nds_cnt"], c["elem_nds_dsp"], kc, kd,
                           K_ref, c["F"], n_near)
    ntrg = c["near_trg_cnt"].size
    U_ref = orc.near_accumulate(np.zeros(ntrg * F.dof), F.dof, c["near_trg_cnt"], c["near_trg_dsp"],
                                c["near_scatter_index"], Un_ref)
    K = F.k_near_subtract(Kq.copy())
    U = F.near_interac(None, c["F"], K, c["near_trg_cnt"], c["near_trg_dsp"], c["near_scatter_index"])
    assert U.shape == U_ref.shape
    assert W.rel_errors(U, U_ref)[0] < 1e-12


@pytest.mark.parametrize("kernel", [0, 3, 7], ids=["LaplaceFxU", "StokesDxU", "StokesFxT"])
def test_gemv_of_stored_matrix_equals_weighted_apply(gpu, kernel):
    """A size-independent property that ties the two routes together:
    near_gemv(K = w*M^T, F) == plan.apply(w*F), block by block."""
    S = gpu
    _, K0, K1, ND, _ = oracle.dims(kernel)
    B = W.torus_blocks(12, 12, seed=3, k0=K0, ns_range=(64, 1024), nt_range=(16, 512))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = S.Plan(kernel, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    P.set_geometry(B["r_trg"], B["r_src"], B["n_src"] if ND else None)
    w = W.uniform(3, 40, ns, 0.5, 1.5)
    k_len = P.near_layout()
    dK = S.DeviceArray(k_len)
    dw = S.DeviceArray.from_host(w)
    P.near_matrix(dK, dw, None, S.NEAR_STORE)
    dF = S.DeviceArray.from_host(B["v_src"])
    dU = S.DeviceArray(nt * K1)
    P.near_gemv(dK, dF, dU)
    S.synchronize()
    via_matrix = dU.download()
    direct = P.apply(np.repeat(w, K0) * B["v_src"], np.zeros(nt * K1))
    assert W.rel_errors(via_matrix, direct)[0] < 1e-13


def test_near_layout_errors(gpu):
    S = gpu
    B = W.torus_blocks(2, 2, seed=1, k0=1, ns_range=(16, 64), nt_range=(4, 16))
    nt, ns = B["r_trg"].size // 3, B["r_src"].size // 3
    P = S.Plan(0, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    P.set_geometry(B["r_trg"], B["r_src"])
    with pytest.raises(S.SctlAmdError):
        P.near_matrix(np.zeros(10))  # no layout yet
    with pytest.raises(S.SctlAmdError):
        P.near_layout(None, trg_normal_dot=True)  # Laplace has one output per target
    with pytest.raises(S.SctlAmdError):
        P.near_layout(np.zeros(4, dtype=np.int64))  # overlapping block matrices
    assert P.near_layout() == int(np.sum(B["trg_cnt"] * B["src_cnt"]))
    P4 = S.Plan(9, B["trg_dsp"], B["trg_cnt"], B["src_dsp"], B["src_cnt"], nt, ns)
    with pytest.raises(S.SctlAmdError):
        P4.near_layout(None, trg_normal_dot=True)  # Stokes FxUP: 4 outputs per target


def test_near_empty_blocks(gpu):
    """Elements with no near targets and an element with no nodes."""
    S = gpu
    orc = oracle.default()
    cnt = np.array([0, 4, 0, 3], dtype=np.int64)
    dsp = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    nds = np.array([5, 0, 2, 6], dtype=np.int64)
    nds_dsp = np.concatenate([[0], np.cumsum(nds)[:-1]])
    rng = np.random.default_rng(1)
    Xt, Xs = rng.random(7 * 3), rng.random(13 * 3) + 2
    blocks = orc.near_blocks(cnt, dsp, 1)
    F = S.NearField(0, cnt, dsp, nds, nds_dsp, Xt, Xs, None, np.ones(13), blocks=blocks)
    assert F.k_len == 4 * 0 + 3 * 6
    K = F.k_near_subtract(np.zeros(F.k_len))
    kc, kd, n = orc.near_layout(0, False, cnt, nds)
    ref = orc.near_subtract(0, False, blocks, dsp, nds, nds, nds_dsp, kc, kd, Xt, Xt, Xs, Xs, np.ones(13),
                            np.zeros(n))
    assert W.rel_errors(K, ref)[0] < 1e-13
    U = F.plan.near_gemv(K, rng.random(13), np.full(7, np.nan))
    assert np.all(U[:4] == 0)  # element 1 has no nodes: zero rows
