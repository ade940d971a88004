"""Test-side reference of the Hilbert-curve keys the mixed-precision path
sorts its sources and targets by (the library computes them on the device)."""
import numpy as np


def hilbert_keys(p, bits):
    """Positions of points p (flat x,y,z) along a 3-D Hilbert curve over the
    set's bounding cube, `bits` per axis: J. Skilling's axes-to-transpose
    transform (Programming the Hilbert curve, AIP Conf. Proc. 707, 2004), then
    the transposed bits interleaved with x's the most significant of each
    triple.  tests/test_hilbert_ref.py pins it as a Hilbert walk."""
    p = p.reshape(-1, 3)
    if len(p) == 0:
        return np.zeros(0, np.int64)
    lo = p.min(0)
    ext = (p.max(0) - lo).max()
    inv = float(1 << bits) / ext if ext > 0 else 0.0
    return hilbert_of_cells(np.minimum(np.maximum((p - lo) * inv, 0.0), float((1 << bits) - 1)).astype(np.int64),
                             bits)


def hilbert_of_cells(q, bits):
    X = q.copy()
    Q = 1 << (bits - 1)
    while Q > 1:
        P = Q - 1
        for i in range(3):
            hit = (X[:, i] & Q) != 0
            X[hit, 0] ^= P
            t = (X[~hit, 0] ^ X[~hit, i]) & P
            X[~hit, 0] ^= t
            X[~hit, i] ^= t
        Q >>= 1
    X[:, 1] ^= X[:, 0]
    X[:, 2] ^= X[:, 1]
    t = np.zeros(len(X), np.int64)
    Q = 1 << (bits - 1)
    while Q > 1:
        t[(X[:, 2] & Q) != 0] ^= Q - 1
        Q >>= 1
    X ^= t[:, None]
    k = np.zeros(len(X), np.int64)
    for b in range(bits - 1, -1, -1):
        for i in range(3):
            k = (k << 1) | ((X[:, i] >> b) & 1)
    return k
