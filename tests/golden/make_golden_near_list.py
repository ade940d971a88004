"""Writes tests/golden/sctl_ref_near_list.npz: inputs and outputs of the
reference library's own BuildNearList (boundary_integral.txx:50-520) on one
process.

Run in the build container, where oracle/_ref/libsctlref_near.so holds the
reference compiled from its sources (``make -C oracle ref``).

Cases:
  torus    48 torus patches of p x p Gauss-Legendre nodes (p in 4..8), 160
           targets near the surface with normals, every radius 0.3
  radii    the same nodes with radii uniform in [0.02, 0.45], no normals
  edge     an element without nodes, an element whose radii are 0, a node at
           the origin with radius 1 and targets at distance exactly 1 (not
           near: the test is strict) and 0.5 (near), an element far away,
           targets near nothing
  cloud    40 clustered elements of 10..50 nodes, mixed radii, 500 targets
           in the unit cube with normals
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from sctl_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sctl_ref_near_list.npz")
KEYS = ("Xtrg_near", "Xn_trg_near", "near_elem_cnt", "near_elem_dsp", "near_scatter_index", "near_trg_cnt",
        "near_trg_dsp")


def cases():
    c = W.near_field_case(n_u=8, n_v=6, n_trg=160, seed=7)
    ns = c["X_far"].size // 3
    yield "torus", c["Xtrg"], c["Xn_trg"], c["X_far"], np.full(ns, 0.3), c["elem_nds_cnt"], c["elem_nds_dsp"]
    rad = 0.02 + 0.43 * W.uniform(7, 50, ns)
    yield "radii", c["Xtrg"], None, c["X_far"], rad, c["elem_nds_cnt"], c["elem_nds_dsp"]
    xs = np.array([[0.2, 0.1, 0.0], [0.3, 0.1, 0.0], [0.2, 0.3, 0.1], [0.25, 0.2, 0.05],  # element 0
                   [0.6, 0.6, 0.6], [0.7, 0.6, 0.6], [0.6, 0.7, 0.6],                       # element 2 (radius 0)
                   [0.0, 0.0, 0.0],                                                          # element 3
                   [50.0, 50.0, 50.0], [51.0, 50.0, 50.0]]).ravel()                          # element 4
    rad = np.array([0.15, 0.15, 0.2, 0.1, 0.0, 0.0, 0.0, 1.0, 0.5, 0.5])
    cnt = np.array([4, 0, 3, 1, 2], dtype=np.int64)
    dsp = np.array([0, 4, 4, 7, 8], dtype=np.int64)
    xt = np.array([[1.0, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, -1.0], [0.22, 0.15, 0.02], [0.6, 0.6, 0.6],
                   [20.0, 20.0, 20.0], [0.3, 0.25, 0.1], [-0.9, -0.2, 0.3], [0.0, 0.0, 0.0]]).ravel()
    yield "edge", xt, None, xs, rad, cnt, dsp
    rng = np.random.default_rng(42)
    ne = 40
    cnt = rng.integers(10, 51, ne).astype(np.int64)
    dsp = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    centres = rng.random((ne, 3))
    xs = np.concatenate([centres[e] + 0.03 * rng.standard_normal((cnt[e], 3)) for e in range(ne)]).ravel()
    rad = np.where(rng.random(cnt.sum()) < 0.5, 0.04, 0.12)
    xt = rng.random((500, 3))
    xn = rng.standard_normal((500, 3))
    xn /= np.linalg.norm(xn, axis=1)[:, None]
    yield "cloud", xt.ravel(), xn.ravel(), xs, rad, cnt, dsp


def main():
    ref = oracle.ReferenceNearList()
    out = {}
    for name, xt, xn, xs, rad, cnt, dsp in cases():
        r = ref.near_list(xt, xs, rad, cnt, dsp, Xn_trg=xn)
        out[name + "_in_Xtrg"], out[name + "_in_Xsrc"], out[name + "_in_rad"] = xt, xs, rad
        out[name + "_in_cnt"], out[name + "_in_dsp"] = cnt, dsp
        if xn is not None:
            out[name + "_in_Xn_trg"] = xn
        for k in KEYS:
            if r[k] is not None:
                out[name + "_" + k] = r[k]
        print(name, "pairs", r["near_scatter_index"].size)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
