"""ORACLE -- test infrastructure only.

ctypes bindings for
  * ``liboracle.so`` -- the project's C restatement of SCTL's
    ``GenericKernel::Eval`` / ``KernelMatrix`` (``oracle/direct_sum.c``), and
  * ``_ref/libsctlref_{v4,scalar}.so`` -- the reference library's own
    GenericKernel code compiled from its headers (``oracle/ref_shim.cpp``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only to check or time against.
The product library never depends on it.

Kernel ids follow the product's C-ABI: 0 Laplace3D_FxU, 1 Laplace3D_DxU,
2 Stokes3D_FxU, 3 Stokes3D_DxU.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

EXACT, AVX512, LDOUBLE = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_L = ctypes.c_long


def build(with_ref=None, force=False):
    """Compile liboracle.so (and the _ref libraries when the reference is present)."""
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/include")
    if with_ref:
        targets.append("ref")
    subprocess.check_call(["make", "-s", "-C", HERE] + (["-B"] if force else []) + targets)


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None and a.size else None


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One build of the C restatement (liboracle.so, or a mutant build)."""

    def __init__(self, path=None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build(with_ref=False)
        L = ctypes.CDLL(path)
        L.oracle_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _L, _dp, _dp, _dp, _L, _dp,
                                  ctypes.c_int]
        L.oracle_kernel_matrix.argtypes = [ctypes.c_int, _dp, _L, _dp, _dp, _L, _dp, ctypes.c_int]
        L.oracle_rsqrt.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _L, _dp]
        L.oracle_dims.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_scale.argtypes = [ctypes.c_int]
        L.oracle_scale.restype = ctypes.c_double
        L.oracle_newton_iters_avx512.argtypes = [ctypes.c_int]
        self.lib = L
        self.path = path

    def dims(self, kernel):
        d = (ctypes.c_int * 5)()
        if self.lib.oracle_dims(kernel, d) != 0:
            raise ValueError("unknown kernel %r" % (kernel,))
        return tuple(d)  # (coord, K0, K1, N_DIM, FLOPS)

    def scale(self, kernel):
        return self.lib.oracle_scale(kernel)

    def have_avx512(self):
        return bool(self.lib.oracle_have_avx512())

    def eval(self, kernel, r_trg, r_src, n_src, v_src, v_trg=None, mode=EXACT, digits=-1, nthreads=None):
        """v_trg += Eval(...); returns v_trg (a fresh zero array when v_trg is None)."""
        _, K0, K1, ND, _ = self.dims(kernel)
        r_trg, r_src, v_src = _c(r_trg), _c(r_src), _c(v_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        assert r_trg.size == 3 * nt and r_src.size == 3 * ns and v_src.size == K0 * ns
        assert not ND or n_src.size == 3 * ns
        out = np.zeros(nt * K1) if v_trg is None else np.ascontiguousarray(v_trg, dtype=np.float64).copy()
        nth = nthreads or os.cpu_count() or 1
        rc = self.lib.oracle_eval(kernel, mode, digits, _p(r_trg), nt, _p(r_src), _p(n_src), _p(v_src), ns,
                                  _p(out), nth)
        if rc != 0:
            raise RuntimeError("oracle_eval failed rc=%d" % rc)
        return out

    def kernel_matrix(self, kernel, r_trg, r_src, n_src, nthreads=None):
        _, K0, K1, ND, _ = self.dims(kernel)
        r_trg, r_src = _c(r_trg), _c(r_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        M = np.zeros((ns * K0, nt * K1))
        rc = self.lib.oracle_kernel_matrix(kernel, _p(r_trg), nt, _p(r_src), _p(n_src), ns, _p(M),
                                           nthreads or os.cpu_count() or 1)
        if rc != 0:
            raise RuntimeError("oracle_kernel_matrix failed rc=%d" % rc)
        return M

    def rsqrt(self, x, mode=EXACT, digits=-1):
        x = _c(x)
        y = np.zeros_like(x)
        rc = self.lib.oracle_rsqrt(mode, digits, _p(x), x.size, _p(y))
        if rc != 0:
            raise RuntimeError("oracle_rsqrt failed rc=%d" % rc)
        return y


_default = None


def default():
    global _default
    if _default is None:
        _default = Oracle()
    return _default


def dims(kernel):
    return default().dims(kernel)


def scale(kernel):
    return default().scale(kernel)


def have_avx512():
    return default().have_avx512()


def eval(*args, **kw):
    return default().eval(*args, **kw)


def kernel_matrix(*args, **kw):
    return default().kernel_matrix(*args, **kw)


def rsqrt(*args, **kw):
    return default().rsqrt(*args, **kw)


def build_mutant(index, out_dir):
    """Compiles the oracle with ORACLE_MUTANT=index into out_dir; returns the path."""
    out = os.path.join(out_dir, "liboracle_mutant%d.so" % index)
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-std=c11",
                           "-DORACLE_MUTANT=%d" % index, "-o", out, os.path.join(HERE, "direct_sum.c")])
    return out


# ------------------------------------------------------------------ reference
class Reference:
    """The reference library's own Eval/KernelMatrix (compiled reference)."""

    def __init__(self, flavour="v4"):
        path = os.path.join(REF_DIR, "libsctlref_%s.so" % flavour)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.sctlref_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _L, _dp, _dp, _dp, _L, _dp]
        L.sctlref_eval_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _L, _fp, _fp, _fp, _L, _fp]
        L.sctlref_kernel_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _L, _dp, _dp, _L, _dp]
        L.sctlref_rsqrt.argtypes = [ctypes.c_int, _dp, _L, _dp]
        L.sctlref_info.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp, ctypes.c_char_p, ctypes.c_int]
        self.lib = L
        self.flavour = flavour

    @staticmethod
    def available(flavour="v4"):
        return os.path.exists(os.path.join(REF_DIR, "libsctlref_%s.so" % flavour))

    def veclen(self):
        return self.lib.sctlref_veclen()

    def info(self, kernel):
        d = (ctypes.c_int * 5)()
        s = ctypes.c_double()
        name = ctypes.create_string_buffer(64)
        self.lib.sctlref_info(kernel, d, ctypes.byref(s), name, 64)
        return tuple(d), s.value, name.value.decode()

    def eval(self, kernel, r_trg, r_src, n_src, v_src, v_trg=None, digits=-1, omp=True):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        r_trg, r_src, v_src = _c(r_trg), _c(r_src), _c(v_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        out = np.zeros(nt * K1) if v_trg is None else np.ascontiguousarray(v_trg, dtype=np.float64).copy()
        self.lib.sctlref_eval(kernel, digits, int(omp), _p(r_trg), nt, _p(r_src), _p(n_src), _p(v_src), ns,
                              _p(out))
        return out

    def eval_f32(self, kernel, r_trg, r_src, n_src, v_src, digits=-1, omp=True):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        r_trg, r_src, v_src = f(r_trg), f(r_src), f(v_src)
        n_src = f(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        out = np.zeros(nt * K1, dtype=np.float32)
        q = lambda a: a.ctypes.data_as(_fp) if a is not None and a.size else None
        self.lib.sctlref_eval_f32(kernel, digits, int(omp), q(r_trg), nt, q(r_src), q(n_src), q(v_src), ns, q(out))
        return out

    def kernel_matrix(self, kernel, r_trg, r_src, n_src, digits=-1):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        r_trg, r_src = _c(r_trg), _c(r_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        M = np.zeros((ns * K0, nt * K1))
        self.lib.sctlref_kernel_matrix(kernel, digits, _p(r_trg), nt, _p(r_src), _p(n_src), ns, _p(M))
        return M

    def rsqrt(self, x, digits=-1):
        x = _c(x)
        y = np.zeros_like(x)
        if self.lib.sctlref_rsqrt(digits, _p(x), x.size, _p(y)) != 0:
            raise ValueError("digits %r not instantiated" % digits)
        return y
