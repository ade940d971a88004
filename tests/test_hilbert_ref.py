"""The test-side Hilbert keys (tests/hilbert_ref.py) are a Hilbert walk, so
the exact-order GPU tests compare the device order against a Hilbert order
and not against a restatement of the device code."""
import numpy as np
import pytest

from hilbert_ref import hilbert_keys, hilbert_of_cells


@pytest.mark.parametrize("bits", [2, 5])
def test_hilbert_keys_walk(bits):
    """The keys of the whole 2^bits grid are a permutation of 0 .. 8^bits - 1
    whose consecutive cells share a face (a Hilbert walk), starting at the
    origin cell."""
    n = 1 << bits
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    k = hilbert_of_cells(g, bits)
    assert (np.sort(k) == np.arange(n ** 3)).all()
    walk = g[np.argsort(k)]
    assert (np.abs(np.diff(walk, axis=0)).sum(1) == 1).all()
    assert (walk[0] == 0).all()


def test_hilbert_keys_of_points():
    """Points are quantised over their own bounding cube: a copy scaled by a
    power of two (exact in binary floating point) has the same keys; a set of
    one point has key 0."""
    rng = np.random.default_rng(3)
    p = rng.random(3000 * 3)
    assert (hilbert_keys(p, 5) == hilbert_keys(p * 4.0, 5)).all()
    assert (hilbert_keys(np.tile([1.0, 2.0, 3.0], 7), 10) == 0).all()
