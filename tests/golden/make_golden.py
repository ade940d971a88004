"""Writes tests/golden/sctl_ref_vectors.npz: input/output vectors of the
reference library's own GenericKernel code.

Run in the build container, where oracle/_ref/ holds the reference compiled
from its sources (``make -C oracle ref``).  Two reference builds are used:
  scalar  no -march: DefaultVecLen()==1, exact 1/sqrt (intrin-wrapper.hpp:563-579)
  v4      -march=x86-64-v4: AVX-512 rsqrt14 + deferred Newton (intrin-wrapper.hpp:3383-3408)

Cases, for each of the four kernels:
  small   Ns=7, Nt=5 inputs of test-kernel_functions.cpp:492-499
  rand    Ns=300, Nt=200 uniform points, 20 targets coincident with sources
  veclen  Nt=3 (the Nt <= VecLen branch, generic-kernel.txx:135-156),
          accumulated into a non-zero v_trg
  digits  the rand case through the v4 build with digits 3, 7 and 11
  kmat    KernelMatrix for Nt=11,Ns=13 and the 1-target / 1-source branches
plus approx_rsqrt<digits> (digits 4, 7, 11, -1) of the v4 build and the
reference's own FP32 Eval of the rand case.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sctl_ref_vectors.npz")


def small_inputs(K0, ND):
    Ns, Nt = 7, 5
    Xs = 0.1 * (np.arange(Ns * 3) + 1) - 0.35
    Vs = 0.3 * (np.arange(Ns * K0) % 5) - 0.4
    Xt = 1.7 + 0.2 * np.arange(Nt * 3)
    Xn = np.tile([0.0, 1.0, 0.0], Ns) if ND else np.zeros(0)
    return Xt, Xs, Xn, Vs


def main():
    refs = {f: oracle.Reference(f) for f in ("scalar", "v4")}
    rng = np.random.default_rng(20240601)
    out = {}
    for k in range(4):
        (_, K0, K1, ND, _), _, name = refs["v4"].info(k)
        # small
        Xt, Xs, Xn, Vs = small_inputs(K0, ND)
        out["k%d_small_in" % k] = np.concatenate([Xt, Xs, Xn, Vs])
        for f, R in refs.items():
            out["k%d_small_eval_%s" % (k, f)] = R.eval(k, Xt, Xs, Xn, Vs, omp=False)
            out["k%d_small_kmat_%s" % (k, f)] = R.kernel_matrix(k, Xt, Xs, Xn)
        # rand with coincident points
        Ns, Nt = 300, 200
        xs = rng.random(Ns * 3)
        xt = rng.random(Nt * 3)
        xt[:60] = xs[:60]
        n = rng.standard_normal(Ns * 3)
        n = (n.reshape(-1, 3) / np.linalg.norm(n.reshape(-1, 3), axis=1)[:, None]).ravel()
        v = rng.uniform(-1, 1, Ns * K0)
        out["k%d_rand_xt" % k], out["k%d_rand_xs" % k], out["k%d_rand_n" % k], out["k%d_rand_v" % k] = xt, xs, n, v
        for f, R in refs.items():
            out["k%d_rand_eval_%s" % (k, f)] = R.eval(k, xt, xs, n, v)
        for d in (3, 7, 11):
            out["k%d_rand_eval_v4_d%d" % (k, d)] = refs["v4"].eval(k, xt, xs, n, v, digits=d)
        out["k%d_rand_eval_f32" % k] = refs["v4"].eval_f32(k, xt, xs, n, v)
        # Nt <= VecLen branch, accumulate
        xt3 = rng.random(3 * 3)
        vt0 = rng.uniform(-1, 1, 3 * K1)
        out["k%d_veclen_xt" % k], out["k%d_veclen_vt0" % k] = xt3, vt0
        for f, R in refs.items():
            out["k%d_veclen_eval_%s" % (k, f)] = R.eval(k, xt3, xs, n, v, v_trg=vt0)
        # KernelMatrix branches
        xt11, xs13 = rng.random(11 * 3), rng.random(13 * 3)
        n13 = n[:13 * 3]
        out["k%d_kmat_xt" % k], out["k%d_kmat_xs" % k], out["k%d_kmat_n" % k] = xt11, xs13, n13
        for f, R in refs.items():
            out["k%d_kmat_%s" % (k, f)] = R.kernel_matrix(k, xt11, xs13, n13)
            out["k%d_kmat1t_%s" % (k, f)] = R.kernel_matrix(k, xt11[:3], xs13, n13)
            out["k%d_kmat1s_%s" % (k, f)] = R.kernel_matrix(k, xt11, xs13[:3], n13[:3])
    x = np.concatenate([rng.random(500) * 100, np.logspace(-30, 30, 61), [0.0, 1.0, 4.0]])
    out["rsqrt_x"] = x
    for d in (4, 7, 11, -1):
        out["rsqrt_v4_d%d" % d] = refs["v4"].rsqrt(x, digits=d)
    out["meta_veclen"] = np.array([refs["scalar"].veclen(), refs["v4"].veclen()])
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes,", len(out), "arrays")


if __name__ == "__main__":
    main()
