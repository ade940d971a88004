"""Host side of the synthetic workloads: the counter-based generator and the
configuration generators (sizes, ranges, determinism)."""
import numpy as np

from sctl_amd import workloads as W


def test_splitmix64_known_values():
    # SplitMix64 with state 0 advanced by the golden gamma: published first outputs
    # of the reference generator (Steele, Lea, Flood 2014; java.util.SplittableRandom).
    z = np.array([0x9E3779B97F4A7C15], dtype=np.uint64)
    with np.errstate(over="ignore"):
        assert int(W._mix64(z)[0]) == 0xE220A8397B1DCDAF
        z2 = np.array([0x9E3779B97F4A7C15 * 2 % (1 << 64)], dtype=np.uint64)
        assert int(W._mix64(z2)[0]) == 0x6E789E6AA1B965F4


def test_counter_streams_are_deterministic_and_independent():
    a = W.u64(7, 3, 1000)
    assert np.array_equal(a, W.u64(7, 3, 1000))
    assert np.array_equal(a[500:], W.u64(7, 3, 500, offset=500))
    assert not np.array_equal(a, W.u64(7, 4, 1000))
    assert not np.array_equal(a, W.u64(8, 3, 1000))


def test_uniform_and_normal_moments():
    u = W.uniform(0, 0, 200000)
    assert u.min() >= 0 and u.max() < 1 and abs(u.mean() - 0.5) < 0.005
    z = W.normal(0, 1, 200001)
    assert z.size == 200001
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01


def test_config_generators():
    xt, xs, n, f = W.sphere(4096, 2048, seed=0)
    r_s = np.linalg.norm(xs.reshape(-1, 3), axis=1)
    r_t = np.linalg.norm(xt.reshape(-1, 3), axis=1)
    assert np.allclose(r_s, 1) and np.array_equal(n, xs)
    assert r_t.min() > 1.05 and r_t.max() < 2
    xt, xs, f = W.cube(100, 50, seed=0)
    assert xs.size == 300 and xt.size == 150 and f.min() >= -1 and f.max() <= 1
    xt, xs, f = W.clusters(1000, 900, seed=0, k0=3)
    assert xs.size == 3000 and f.size == 3000
    B = W.torus_blocks(8, 4, seed=0)
    assert B["trg_cnt"].size == 32
    assert B["src_cnt"].min() >= 64 and B["src_cnt"].max() <= 4096
    assert B["trg_cnt"].min() >= 16 and B["trg_cnt"].max() <= 2048
    p = np.sqrt(B["src_cnt"])
    assert np.array_equal(p, np.rint(p))  # p x p Gauss-Legendre grids
    assert np.allclose(np.linalg.norm(B["n_src"].reshape(-1, 3), axis=1), 1)
    assert W.interactions(B) == int((B["src_cnt"] * B["trg_cnt"]).sum())


def test_rel_errors():
    ref = np.array([1.0, -2.0, 4.0])
    assert W.rel_errors(ref, ref) == (0.0, 0.0)
    e_inf, e_2 = W.rel_errors(ref + [0, 0, 0.004], ref)
    assert abs(e_inf - 0.001) < 1e-15
