"""GPU parity of Eval and KernelMatrix (through the C-ABI) against the oracle,
the reference's own outputs (golden vectors), and size-independent properties
at full size.

Tolerances (normwise: max|x - ref| / max|ref|):
  FP64 Eval / KernelMatrix   1e-13  (the FP64 bar)
  mixed precision            1e-6   (FP32 pairwise, FP64 Kahan accumulation)
  FP64 with digits d < 15    10^-d  (fewer Newton steps, as the digits knob asks)
"""
import numpy as np
import pytest

import oracle
import pins
from hilbert_ref import hilbert_keys as _hilbert_keys
from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13
MIXED_TOL = 1e-6


def nerr(x, ref):
    return W.rel_errors(x, ref)[0]


def rand_case(rng, k, nt, ns, coincide=0):
    _, K0, K1, ND, _ = oracle.dims(k)
    xs = rng.random(ns * 3)
    xt = rng.random(nt * 3)
    if coincide:
        xt[: 3 * coincide] = xs[: 3 * coincide]
    n = rng.standard_normal(ns * 3)
    n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
    v = rng.uniform(-1, 1, ns * K0)
    return xt, xs, n, v


KERNELS = list(range(12))
IDS = ["LaplaceFxU", "LaplaceDxU", "StokesFxU", "StokesDxU", "LaplaceFxdU", "LaplaceDxdU", "LaplaceFxd2U",
       "StokesFxT", "StokesFSxU", "StokesFxUP", "BiotSavartFxU", "BiotSavartFxdU"]


def golden_for(k):
    """The reference's outputs for kernel k (the first four kernels and the
    other eight are in two fixture files with the same keys)."""
    return pins.golden() if k < 4 else pins.golden_more()


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
def test_eval_fp64_matches_reference_golden(gpu, k):
    """Against the reference's own scalar-build outputs (exact 1/sqrt)."""
    G = golden_for(k)
    _, K0, K1, ND, _ = oracle.dims(k)
    Xt, Xs, Xn, Vs = pins.small_inputs(K0, ND)
    assert nerr(gpu.eval(k, Xt, Xs, Xn, Vs), G["k%d_small_eval_scalar" % k]) < FP64_TOL
    xt, xs, n, v = (G["k%d_rand_%s" % (k, a)] for a in ("xt", "xs", "n", "v"))
    out = gpu.eval(k, xt, xs, n, v)
    assert np.isfinite(out).all()
    assert nerr(out, G["k%d_rand_eval_scalar" % k]) < FP64_TOL
    assert nerr(out, G["k%d_rand_eval_v4" % k]) < FP64_TOL
    # Nt <= VecLen branch with accumulation into a non-zero v_trg
    vt = gpu.eval(k, G["k%d_veclen_xt" % k], xs, n, v, v_trg=G["k%d_veclen_vt0" % k].copy())
    assert nerr(vt, G["k%d_veclen_eval_scalar" % k]) < FP64_TOL


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
@pytest.mark.parametrize("nt,ns,coincide", [(1, 1, 0), (5, 7, 0), (63, 65, 10), (257, 255, 0), (1000, 3000, 40),
                                            (3000, 17, 0), (20000, 700, 0)])
def test_eval_fp64_matches_oracle(gpu, k, nt, ns, coincide):
    rng = np.random.default_rng(nt * 7919 + ns)
    xt, xs, n, v = rand_case(rng, k, nt, ns, coincide)
    ref = oracle.eval(k, xt, xs, n, v)
    out = gpu.eval(k, xt, xs, n, v)
    assert np.isfinite(out).all()
    assert nerr(out, ref) < FP64_TOL


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
def test_eval_fp64_config1_size(gpu, k):
    """The 20,000 x 20,000 uniform-cube case (seed 0), every kernel."""
    _, K0, K1, ND, _ = oracle.dims(k)
    xt, xs, f = W.cube(20000, 20000, seed=0, k0=K0)
    n = W.normal(0, 3, 20000 * 3) if ND else None
    ref = oracle.eval(k, xt, xs, n, f)
    out = gpu.eval(k, xt, xs, n, f)
    e_inf, e_rms = W.rel_errors(out, ref)
    assert e_inf < FP64_TOL and e_rms < FP64_TOL


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
@pytest.mark.parametrize("digits", [3, 7, 11])
def test_eval_fp64_digits(gpu, k, digits):
    rng = np.random.default_rng(digits)
    xt, xs, n, v = rand_case(rng, k, 500, 700)
    ref = oracle.eval(k, xt, xs, n, v)
    out = gpu.eval(k, xt, xs, n, v, digits=digits)
    assert nerr(out, ref) < 10.0 ** (-digits)


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
@pytest.mark.parametrize("newton", [False, True], ids=["rsq", "rsq+newton"])
@pytest.mark.parametrize("nt,ns,coincide", [(5, 7, 0), (300, 257, 30), (3000, 5000, 0), (20000, 20000, 0)])
def test_eval_mixed_matches_oracle(gpu, k, newton, nt, ns, coincide):
    rng = np.random.default_rng(nt + ns)
    xt, xs, n, v = rand_case(rng, k, nt, ns, coincide)
    ref = oracle.eval(k, xt, xs, n, v)
    out = gpu.eval(k, xt, xs, n, v, precision=gpu.MIXED, flags=gpu.MIXED_NEWTON if newton else 0)
    assert np.isfinite(out).all()
    e_inf, e_rms = W.rel_errors(out, ref)
    assert e_inf < MIXED_TOL and e_rms < MIXED_TOL


@pytest.mark.parametrize("case", ["sources_one_point", "targets_one_point", "one_source", "one_target", "far_apart"])
@pytest.mark.parametrize("k", [0, 1, 2, 3], ids=IDS[:4])
def test_eval_degenerate_point_sets(gpu, k, case):
    """Point sets on which the Hilbert ordering of the mixed path degenerates:
    a set collapsed to one point (a zero-extent box, every key equal), a
    single source or target, and targets in a tiny box far from the sources.
    Both precisions against the oracle.  With random-sign densities these
    sums cancel heavily (the far case sums 900 terms of ~1e-4 to ~5e-5), and
    FP32 pairwise terms are accurate relative to their own size, so the
    mixed leg is bounded relative to the sum of the terms' magnitudes,
    |K|^T |f| from the oracle's kernel matrix; FP64 keeps the normwise bar."""
    rng = np.random.default_rng(77)
    _, K0, K1, ND, _ = oracle.dims(k)
    nt = 1 if case == "one_target" else 700
    ns = 1 if case == "one_source" else 900
    xs, xt = rng.random(ns * 3), rng.random(nt * 3)
    if case == "sources_one_point":
        xs = np.tile([0.3, 0.4, 0.5], ns)
    if case == "targets_one_point":
        xt = np.tile([0.7, 0.1, 0.9], nt)
    if case == "far_apart":
        xt = xt * 1e-3 + 50.0
    n = rng.standard_normal(ns * 3)
    n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
    v = rng.uniform(-1, 1, ns * K0)
    ref = oracle.eval(k, xt, xs, n, v)
    out = gpu.eval(k, xt, xs, n, v, precision=gpu.FP64)
    assert np.isfinite(out).all() and nerr(out, ref) < FP64_TOL, (case, nerr(out, ref))
    magnitude = np.abs(oracle.kernel_matrix(k, xt, xs, n)).T @ np.abs(v)
    out = gpu.eval(k, xt, xs, n, v, precision=gpu.MIXED)
    assert np.isfinite(out).all()
    e = np.max(np.abs(out - ref)) / np.max(magnitude)
    assert e < MIXED_TOL, (case, e)


@pytest.mark.parametrize("na,nb", [(300, 5), (70000, 3)])
def test_hilbert_order_zero_extent(gpu, na, nb):
    """A set collapsed to one point has a zero-extent box: every key is 0 and
    the order is the input order (both sorts, with a tiny second set)."""
    a = np.tile([0.25, -3.0, 7.5], na)
    b = np.random.default_rng(na).random(nb * 3)
    pa, pb = gpu.hilbert_order(a, b)
    assert (pa.astype(np.int64) == np.arange(na)).all()
    bits = 5 if na + nb <= 65536 else 10
    assert (pb.astype(np.int64) == np.argsort(_hilbert_keys(b, bits), kind="stable")).all()


@pytest.mark.parametrize("na,nb", [(1, 0), (0, 7), (1, 1), (1000, 3000), (1024, 1024), (5000, 700), (20000, 20000),
                                   (20480, 20480), (20481, 20480), (32768, 32768), (65535, 1), (40000, 26000)])
def test_hilbert_order(gpu, na, nb):
    """The orders a mixed Eval sums in: each set sorted stably by its Hilbert
    key over its own bounding cube (5 bits per axis up to 65 536 points in
    all, by the two-pass radix sort; 10 bits above, by rocPRIM), with
    repeated points and a second set of another extent and offset."""
    rng = np.random.default_rng(na * 7 + nb)
    a = rng.random(na * 3)
    b = rng.random(nb * 3) * 5.0 - 2.0
    if na > 30:
        a[3:30] = np.tile(a[:3], 9)
    if nb > 30:
        b[30:60] = b[0]
    bits = 5 if na + nb <= 65536 else 10
    pa, pb = gpu.hilbert_order(a, b)
    assert (pa.astype(np.int64) == np.argsort(_hilbert_keys(a, bits), kind="stable")).all()
    assert (pb.astype(np.int64) == np.argsort(_hilbert_keys(b, bits), kind="stable")).all()


@pytest.mark.parametrize("k", KERNELS[2:], ids=IDS[2:])
def test_eval_mixed_clusters(gpu, k):
    """Gaussian clusters (sigma 0.02): the case that needs tile centring."""
    _, K0, K1, ND, _ = oracle.dims(k)
    xt, xs, f = W.clusters(32768, 32768, seed=0, k0=K0)
    n = W.normal(0, 30, 32768 * 3) if ND else None
    ref = oracle.eval(k, xt, xs, n, f)
    out = gpu.eval(k, xt, xs, n, f, precision=gpu.MIXED)
    e_inf, e_rms = W.rel_errors(out, ref)
    assert e_inf < MIXED_TOL and e_rms < MIXED_TOL


def test_auto_precision_follows_digits(gpu):
    rng = np.random.default_rng(1)
    xt, xs, n, v = rand_case(rng, 0, 2000, 2000)
    ref = oracle.eval(0, xt, xs, n, v)
    fp64 = gpu.eval(0, xt, xs, n, v, precision=gpu.AUTO, digits=-1)
    mixed = gpu.eval(0, xt, xs, n, v, precision=gpu.AUTO, digits=6)
    assert nerr(fp64, ref) < FP64_TOL
    assert FP64_TOL < nerr(mixed, ref) < MIXED_TOL


@pytest.mark.parametrize("prec", [0, 1], ids=["fp64", "mixed"])
def test_eval_semantics(gpu, prec):
    """Accumulation into v_trg (generic-kernel.txx:129-132), Ns == 0 and
    Nt == 0 leave v_trg untouched, coincident points contribute zero."""
    rng = np.random.default_rng(4)
    xt, xs, n, v = rand_case(rng, 2, 100, 80)
    v0 = rng.standard_normal(300)
    out = gpu.eval(2, xt, xs, n, v, v_trg=v0.copy(), precision=prec)
    ref = oracle.eval(2, xt, xs, n, v, v_trg=v0)
    assert nerr(out, ref) < (FP64_TOL if prec == 0 else MIXED_TOL)
    same = gpu.eval(2, xt, xs[:0], n[:0], v[:0], v_trg=v0.copy(), precision=prec)
    assert np.array_equal(same, v0)
    empty = gpu.eval(2, xt[:0], xs, n, v, v_trg=np.zeros(0), precision=prec)
    assert empty.size == 0
    # every target on a source: only r == 0 pairs are dropped
    out = gpu.eval(0, xs, xs, None, v[:80], precision=prec)
    ref = oracle.eval(0, xs, xs, None, v[:80])
    assert np.isfinite(out).all() and nerr(out, ref) < (FP64_TOL if prec == 0 else MIXED_TOL)


def test_python_generic_kernel_mirror(gpu):
    """GenericKernel.Eval resizes a wrong-sized v_trg and zeroes it, accumulates
    into a right-sized one, and rejects inconsistent shapes."""
    K = gpu.GenericKernel(gpu.LAPLACE3D_FXU)
    assert (K.CoordDim(), K.SrcDim(), K.TrgDim(), K.NormalDim(), K.Name()) == (3, 1, 1, 0, "Laplace3D-FxU")
    sf = 1 / (4 * np.pi)
    v = K.Eval(None, [1, 0, 0, 2, 0, 0, 3, 0, 0, 5, 0, 0], [0, 0, 0], None, [1.0])
    assert np.allclose(v, [sf, sf / 2, sf / 3, sf / 5], rtol=1e-14)
    v2 = K.Eval(v, [1, 0, 0, 2, 0, 0, 3, 0, 0, 5, 0, 0], [0, 0, 0], None, [1.0])
    assert np.allclose(v2, 2 * v, rtol=1e-14)
    with pytest.raises(ValueError):
        K.Eval(None, [1, 0, 0], [0, 0, 0], None, [1.0, 2.0])


@pytest.mark.parametrize("k", KERNELS, ids=IDS)
def test_kernel_matrix_matches_reference(gpu, k):
    G = golden_for(k)
    _, K0, K1, ND, _ = oracle.dims(k)
    Xt, Xs, Xn, Vs = pins.small_inputs(K0, ND)
    M = gpu.kernel_matrix(k, Xt, Xs, Xn)
    assert nerr(M, G["k%d_small_kmat_scalar" % k]) < FP64_TOL
    xt11, xs13, n13 = (G["k%d_kmat_%s" % (k, a)] for a in ("xt", "xs", "n"))
    for tag, a, b, c in (("kmat", xt11, xs13, n13), ("kmat1t", xt11[:3], xs13, n13),
                         ("kmat1s", xt11, xs13[:3], n13[:3])):
        ref = G["k%d_%s_scalar" % (k, tag)]
        M = gpu.kernel_matrix(k, a, b, c)
        assert M.shape == ref.shape and nerr(M, ref) < FP64_TOL
    # Eval == M^T v
    rng = np.random.default_rng(k)
    xt, xs, n, v = rand_case(rng, k, 123, 77, coincide=5)
    M = gpu.kernel_matrix(k, xt, xs, n)
    assert nerr(M, oracle.kernel_matrix(k, xt, xs, n)) < FP64_TOL
    assert nerr(M.T @ v, gpu.eval(k, xt, xs, n, v)) < FP64_TOL


HELMHOLTZ = [12, 13, 14, 15]
H_IDS = ["HelmholtzFxU", "HelmholtzDxU", "HelmholtzFxdU", "HelmholtzDiffFxdU"]


@pytest.mark.parametrize("k", HELMHOLTZ, ids=H_IDS)
def test_helmholtz_matches_reference_golden(gpu, k):
    """Eval and KernelMatrix against the reference's own outputs at four
    wavenumbers (0, 0.7, 3, 25), all branches."""
    G = pins.golden_helmholtz()
    _, K0, K1, ND, _ = oracle.dims(k)
    for j, mu in enumerate(G["mu"]):
        p = "k%d_m%d_" % (k, j)
        Xt, Xs, Xn, Vs = pins.small_inputs(K0, ND)
        assert nerr(gpu.eval(k, Xt, Xs, Xn, Vs, ctx=mu), G[p + "small_eval_scalar"]) < FP64_TOL
        assert nerr(gpu.kernel_matrix(k, Xt, Xs, Xn, ctx=mu), G[p + "small_kmat_scalar"]) < FP64_TOL
        xt, xs, n, v = (G[p + "rand_" + a] for a in ("xt", "xs", "n", "v"))
        out = gpu.eval(k, xt, xs, n, v, ctx=mu)
        assert np.isfinite(out).all()
        assert nerr(out, G[p + "rand_eval_scalar"]) < FP64_TOL, (mu, nerr(out, G[p + "rand_eval_scalar"]))
        vt = gpu.eval(k, G[p + "veclen_xt"], xs, n, v, v_trg=G[p + "veclen_vt0"].copy(), ctx=mu)
        assert nerr(vt, G[p + "veclen_eval_scalar"]) < FP64_TOL
        xt11, xs13, n13 = (G[p + "kmat_" + a] for a in ("xt", "xs", "n"))
        for tag, a, b, c in (("kmat", xt11, xs13, n13), ("kmat1t", xt11[:3], xs13, n13),
                             ("kmat1s", xt11, xs13[:3], n13[:3])):
            ref = G[p + tag + "_scalar"]
            M = gpu.kernel_matrix(k, a, b, c, ctx=mu)
            assert M.shape == ref.shape and nerr(M, ref) < FP64_TOL


@pytest.mark.parametrize("k", HELMHOLTZ, ids=H_IDS)
@pytest.mark.parametrize("mu", [0.5, 4.0, 40.0])
@pytest.mark.parametrize("nt,ns,coincide", [(63, 65, 10), (1000, 3000, 40), (20000, 700, 0)])
def test_helmholtz_fp64_matches_oracle(gpu, k, mu, nt, ns, coincide):
    rng = np.random.default_rng(nt + 3 * ns)
    xt, xs, n, v = rand_case(rng, k, nt, ns, coincide)
    ref = oracle.eval(k, xt, xs, n, v, ctx=mu)
    out = gpu.eval(k, xt, xs, n, v, ctx=mu)
    assert np.isfinite(out).all()
    assert nerr(out, ref) < FP64_TOL
    M = gpu.kernel_matrix(k, xt[:3 * 50], xs[:3 * 40], n[:3 * 40], ctx=mu)
    assert nerr(M, oracle.kernel_matrix(k, xt[:3 * 50], xs[:3 * 40], n[:3 * 40], ctx=mu)) < FP64_TOL


@pytest.mark.parametrize("k", HELMHOLTZ, ids=H_IDS)
@pytest.mark.parametrize("mu,tol", [(1.0, MIXED_TOL), (6.0, 1e-5)])
@pytest.mark.parametrize("nt,ns,coincide", [(300, 257, 30), (20000, 20000, 0)])
def test_helmholtz_mixed_matches_oracle(gpu, k, mu, tol, nt, ns, coincide):
    """Mixed precision: the FP32 phase mu*r carries about 2^-23 relative
    error, so the bound grows with mu times the domain size (unit cube):
    1e-6 at mu = 1, 1e-5 at mu = 6."""
    rng = np.random.default_rng(nt + ns + 1)
    xt, xs, n, v = rand_case(rng, k, nt, ns, coincide)
    ref = oracle.eval(k, xt, xs, n, v, ctx=mu)
    out = gpu.eval(k, xt, xs, n, v, precision=gpu.MIXED, ctx=mu)
    assert np.isfinite(out).all()
    e_inf, e_rms = W.rel_errors(out, ref)
    assert e_inf < tol and e_rms < tol, (e_inf, e_rms)


@pytest.mark.parametrize("k", HELMHOLTZ, ids=H_IDS)
@pytest.mark.parametrize("digits", [3, 11])
def test_helmholtz_digits(gpu, k, digits):
    rng = np.random.default_rng(digits + k)
    xt, xs, n, v = rand_case(rng, k, 500, 700)
    ref = oracle.eval(k, xt, xs, n, v, ctx=1.0)
    out = gpu.eval(k, xt, xs, n, v, digits=digits, ctx=1.0)
    assert nerr(out, ref) < 10.0 ** (-digits)


def test_helmholtz_needs_its_wavenumber(gpu):
    rng = np.random.default_rng(0)
    xt, xs, n, v = rand_case(rng, 12, 10, 10)
    with pytest.raises(gpu.SctlAmdError):
        gpu.eval(12, xt, xs, n, v)
    with pytest.raises(gpu.SctlAmdError):
        gpu.kernel_matrix(12, xt, xs, n)
    K = gpu.GenericKernel(gpu.HELMHOLTZ3D_FXU)
    K.SetCtxPtr(2.0)
    assert nerr(K.Eval(None, xt, xs, None, v), oracle.eval(12, xt, xs, None, v, ctx=2.0)) < FP64_TOL
    # the other kernels ignore a context
    assert nerr(gpu.eval(0, xt, xs, None, v[:10], ctx=5.0), oracle.eval(0, xt, xs, None, v[:10])) < FP64_TOL


@pytest.mark.parametrize("k", [0, 1], ids=["LaplaceFxU", "LaplaceDxU"])
def test_full_size_sphere_shell(gpu, k):
    """Config 2 at full size (2^20 sources on the sphere, 2^20 targets in the
    shell): a subset of targets against the oracle, and linearity of the whole
    output (Eval(f) + Eval(g) == Eval(f + g))."""
    N = 1 << 20
    xt, xs, n, f = W.sphere(N, N, seed=0)
    g = W.normal(1, 5, N)
    d = [gpu.DeviceArray.from_host(a) for a in (xt, xs, n, f, g, f + g)]
    outs = []
    for dens in (d[3], d[4], d[5]):
        o = gpu.DeviceArray.from_host(np.zeros(N))
        gpu.eval(k, d[0], d[1], d[2], dens, o)
        outs.append(o.download())
    sub = 2048
    ref = oracle.eval(k, xt[: 3 * sub], xs, n, f)
    assert nerr(outs[0][:sub], ref) < FP64_TOL
    assert nerr(outs[0] + outs[1], outs[2]) < 1e-12


@pytest.mark.parametrize("k", [0, 3, 7, 12], ids=["LaplaceFxU", "StokesDxU", "StokesFxT", "HelmholtzFxU"])
@pytest.mark.parametrize("offset,scale", [(1e6, 1.0), (-3e3, 1e-4), (0.0, 1e5), (1e9, 1e3)],
                         ids=["far-offset", "small-scale", "large-scale", "offset-and-scale"])
def test_mixed_translation_and_scale(gpu, k, offset, scale):
    """Mixed precision differences positions in FP64 against tile centres, so
    its accuracy does not depend on where the points sit or (within the FP32
    exponent range of r^-5) on their scale: the same configuration translated
    and scaled against the oracle.  Densities and the Helmholtz wavenumber are
    scaled with it so that the kernel values stay comparable."""
    rng = np.random.default_rng(31)
    xt, xs, n, v = rand_case(rng, k, 3000, 2500, coincide=20)
    xt, xs = xt * scale + offset, xs * scale + offset
    mu = 2.0 / scale if k >= 12 else None
    ref = oracle.eval(k, xt, xs, n, v, ctx=mu)
    out = gpu.eval(k, xt, xs, n, v, precision=gpu.MIXED, ctx=mu)
    assert np.isfinite(out).all()
    e_inf, e_rms = W.rel_errors(out, ref)
    assert e_inf < MIXED_TOL and e_rms < MIXED_TOL, (e_inf, e_rms)
    f64 = gpu.eval(k, xt, xs, n, v, ctx=mu)
    assert nerr(f64, ref) < FP64_TOL
