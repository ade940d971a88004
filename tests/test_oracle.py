"""The oracle (oracle/direct_sum.c) pinned to the reference.

Every pin of tests/pins.py must hold for the oracle, and every deliberately
mutated oracle build must fail at least one of them.  The golden vectors come
from the reference's own GenericKernel compiled from its sources
(tests/golden/make_golden.py)."""
import os
import tempfile

import numpy as np
import pytest

import oracle
import pins


@pytest.mark.parametrize("check", pins.CHECKS, ids=[c.__name__ for c in pins.CHECKS])
def test_oracle_pin(check):
    ok, detail = check(oracle.default())
    assert ok, detail


@pytest.fixture(scope="module")
def mutant_dir():
    with tempfile.TemporaryDirectory() as d:
        yield d


@pytest.mark.parametrize("mutant", range(1, 18))
def test_pins_reject_mutant(mutant, mutant_dir):
    orc = oracle.Oracle(oracle.build_mutant(mutant, mutant_dir))
    failed = [c.__name__ for c in pins.CHECKS if not c(orc)[0]]
    assert failed, "mutant %d passes every pin" % mutant


def test_golden_metadata():
    G = pins.golden()
    assert list(G["meta_veclen"]) == [1, 8]  # scalar build, AVX-512 build (vec.txx:27-41)


def test_reference_fp32_path_error_row():
    """The reference's own FP32 Eval (kept in the golden vectors) is a few 1e-7
    to 1e-6 off the FP64 result: the bar the mixed-precision path must beat."""
    G = pins.golden()
    for k in range(4):
        e = np.abs(G["k%d_rand_eval_f32" % k] - G["k%d_rand_eval_v4" % k]).max() / np.abs(G["k%d_rand_eval_v4" % k]).max()
        assert 1e-9 < e < 1e-4


@pytest.mark.skipif(not oracle.Reference.available("v4"), reason="compiled reference not present")
def test_oracle_matches_compiled_reference_medium():
    """Where the compiled reference is present: the oracle against it on a
    fresh 1000 x 1200 case (AVX-512 build, <= 2e-15 normwise)."""
    ref = oracle.Reference("v4")
    rng = np.random.default_rng(99)
    for k in range(4):
        _, K0, K1, ND, _ = oracle.dims(k)
        xs, xt, n = rng.random(1200 * 3), rng.random(1000 * 3), rng.standard_normal(1200 * 3)
        v = rng.standard_normal(1200 * K0)
        a = ref.eval(k, xt, xs, n, v)
        b = oracle.eval(k, xt, xs, n, v)
        assert np.abs(a - b).max() / np.abs(a).max() < 2e-15
        if oracle.have_avx512():
            c = oracle.eval(k, xt, xs, n, v, mode=oracle.AVX512)
            assert np.abs(a - c).max() / np.abs(a).max() < 2e-15
