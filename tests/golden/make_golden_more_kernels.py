"""Writes tests/golden/sctl_ref_more_kernels.npz: input/output vectors of the
reference library's own GenericKernel code for the real-valued micro-kernels
beyond the four of sctl_ref_vectors.npz (kernel ids 4..11: Laplace3D_FxdU,
Laplace3D_DxdU, Laplace3D_Fxd2U, Stokes3D_FxT, Stokes3D_FSxU, Stokes3D_FxUP,
BiotSavart3D_FxU, BiotSavart3D_FxdU; kernel_functions.hpp:81-123,220-489).

Run in the build container, where oracle/_ref/ holds the reference compiled
from its sources (``make -C oracle ref``).  Cases per kernel, as in
make_golden.py: small (Ns=7, Nt=5), rand (Ns=300, Nt=200, 60 coincident
points), veclen (Nt=3 accumulated into a non-zero v_trg), kmat (KernelMatrix
11 x 13 and its one-target / one-source branches); scalar and AVX-512 builds.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sctl_ref_more_kernels.npz")
KERNELS = range(4, 12)


def main():
    refs = {f: oracle.Reference(f) for f in ("scalar", "v4")}
    rng = np.random.default_rng(20241202)
    out = {}
    for k in KERNELS:
        (_, K0, K1, ND, _), _, name = refs["scalar"].info(k)
        Xs = 0.1 * (np.arange(7 * 3) + 1) - 0.35
        Vs = 0.3 * (np.arange(7 * K0) % 5) - 0.4
        Xt = 1.7 + 0.2 * np.arange(5 * 3)
        Xn = np.tile([0.0, 1.0, 0.0], 7) if ND else np.zeros(0)
        out["k%d_small_in" % k] = np.concatenate([Xt, Xs, Xn, Vs])
        for f, R in refs.items():
            out["k%d_small_eval_%s" % (k, f)] = R.eval(k, Xt, Xs, Xn, Vs, omp=False)
            out["k%d_small_kmat_%s" % (k, f)] = R.kernel_matrix(k, Xt, Xs, Xn)
        xs, xt = rng.random(300 * 3), rng.random(200 * 3)
        xt[:60] = xs[:60]
        n = rng.standard_normal(300 * 3)
        n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
        v = rng.uniform(-1, 1, 300 * K0)
        out["k%d_rand_xt" % k], out["k%d_rand_xs" % k], out["k%d_rand_n" % k], out["k%d_rand_v" % k] = xt, xs, n, v
        for f, R in refs.items():
            out["k%d_rand_eval_%s" % (k, f)] = R.eval(k, xt, xs, n, v)
        xt3, vt0 = rng.random(3 * 3), rng.uniform(-1, 1, 3 * K1)
        out["k%d_veclen_xt" % k], out["k%d_veclen_vt0" % k] = xt3, vt0
        for f, R in refs.items():
            out["k%d_veclen_eval_%s" % (k, f)] = R.eval(k, xt3, xs, n, v, v_trg=vt0)
        xt11, xs13 = rng.random(11 * 3), rng.random(13 * 3)
        n13 = n[:13 * 3]
        out["k%d_kmat_xt" % k], out["k%d_kmat_xs" % k], out["k%d_kmat_n" % k] = xt11, xs13, n13
        for f, R in refs.items():
            out["k%d_kmat_%s" % (k, f)] = R.kernel_matrix(k, xt11, xs13, n13)
            out["k%d_kmat1t_%s" % (k, f)] = R.kernel_matrix(k, xt11[:3], xs13, n13)
            out["k%d_kmat1s_%s" % (k, f)] = R.kernel_matrix(k, xt11, xs13[:3], n13[:3])
        print(name, K0, K1, ND)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
