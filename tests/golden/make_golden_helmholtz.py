"""Writes tests/golden/sctl_ref_helmholtz.npz: input/output vectors of the
reference library's own GenericKernel code for the Helmholtz micro-kernels
(kernel ids 12..15: Helmholtz3D_FxU, Helmholtz3D_DxU, Helmholtz3D_FxdU,
HelmholtzDiff3D_FxdU; kernel_functions.hpp:485-659), with the wavenumber set
through GenericKernel::SetCtxPtr (generic-kernel.txx:64-66).

Run in the build container, where oracle/_ref/ holds the reference compiled
from its sources (``make -C oracle ref``).  Cases per kernel and wavenumber
(MU below), as in make_golden.py: small (Ns=7, Nt=5), rand (Ns=300, Nt=200,
60 coincident points), veclen (Nt=3 accumulated into a non-zero v_trg), kmat
(KernelMatrix 11 x 13 and its one-target / one-source branches); scalar and
AVX-512 builds.  Keys: k<id>_m<j>_<case>_<flavour> for wavenumber MU[j].
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sctl_ref_helmholtz.npz")
KERNELS = range(12, 16)
MU = (0.0, 0.7, 3.0, 25.0)


def main():
    refs = {f: oracle.Reference(f) for f in ("scalar", "v4")}
    rng = np.random.default_rng(20241203)
    out = {"mu": np.array(MU)}
    for k in KERNELS:
        (_, K0, K1, ND, _), _, name = refs["scalar"].info(k)
        for j, mu in enumerate(MU):
            p = "k%d_m%d_" % (k, j)
            Xs = 0.1 * (np.arange(7 * 3) + 1) - 0.35
            Vs = 0.3 * (np.arange(7 * K0) % 5) - 0.4
            Xt = 1.7 + 0.2 * np.arange(5 * 3)
            Xn = np.tile([0.0, 1.0, 0.0], 7) if ND else np.zeros(0)
            for f, R in refs.items():
                out[p + "small_eval_" + f] = R.eval(k, Xt, Xs, Xn, Vs, omp=False, ctx=mu)
                out[p + "small_kmat_" + f] = R.kernel_matrix(k, Xt, Xs, Xn, ctx=mu)
            xs, xt = rng.random(300 * 3), rng.random(200 * 3)
            xt[:60] = xs[:60]
            n = rng.standard_normal(300 * 3)
            n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
            v = rng.uniform(-1, 1, 300 * K0)
            out[p + "rand_xt"], out[p + "rand_xs"], out[p + "rand_n"], out[p + "rand_v"] = xt, xs, n, v
            for f, R in refs.items():
                out[p + "rand_eval_" + f] = R.eval(k, xt, xs, n, v, ctx=mu)
            xt3, vt0 = rng.random(3 * 3), rng.uniform(-1, 1, 3 * K1)
            out[p + "veclen_xt"], out[p + "veclen_vt0"] = xt3, vt0
            for f, R in refs.items():
                out[p + "veclen_eval_" + f] = R.eval(k, xt3, xs, n, v, v_trg=vt0, ctx=mu)
            xt11, xs13 = rng.random(11 * 3), rng.random(13 * 3)
            n13 = n[:13 * 3]
            out[p + "kmat_xt"], out[p + "kmat_xs"], out[p + "kmat_n"] = xt11, xs13, n13
            for f, R in refs.items():
                out[p + "kmat_" + f] = R.kernel_matrix(k, xt11, xs13, n13, ctx=mu)
                out[p + "kmat1t_" + f] = R.kernel_matrix(k, xt11[:3], xs13, n13, ctx=mu)
                out[p + "kmat1s_" + f] = R.kernel_matrix(k, xt11, xs13[:3], n13[:3], ctx=mu)
        print(name, K0, K1, ND)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
