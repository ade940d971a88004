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
2 Stokes3D_FxU, 3 Stokes3D_DxU, 4 Laplace3D_FxdU, 5 Laplace3D_DxdU,
6 Laplace3D_Fxd2U, 7 Stokes3D_FxT, 8 Stokes3D_FSxU, 9 Stokes3D_FxUP,
10 BiotSavart3D_FxU, 11 BiotSavart3D_FxdU, 12 Helmholtz3D_FxU,
13 Helmholtz3D_DxU, 14 Helmholtz3D_FxdU, 15 HelmholtzDiff3D_FxdU.  The
Helmholtz kernels take the wavenumber as ``ctx``.
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


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _ctx(ctx):
    """Pointer to the kernel context (one double: the Helmholtz wavenumber), or NULL."""
    if ctx is None:
        return None
    return ctypes.pointer(ctypes.c_double(float(ctx)))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long)) if a is not None and a.size else None


def _near_list(fn, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg=None):
    """Calls a BuildNearList-shaped C function twice (size, then fill)."""
    Xt, Xs, rad = _c(Xtrg), _c(Xsrc), _c(src_radius)
    Xn = None if Xn_trg is None else _c(Xn_trg)
    ec, ed = _i(elem_nds_cnt), _i(elem_nds_dsp)
    nt, ns, ne = Xt.size // 3, Xs.size // 3, ec.size
    n = fn(_p(Xt), nt, _p(Xn), _p(Xs), ns, _p(rad), ne, _ip(ec), _ip(ed), 0, None, None, None, None, None, None,
           None)
    if n < 0:
        raise RuntimeError("near list failed rc=%d" % n)
    out = dict(Xtrg_near=np.zeros(3 * n), Xn_trg_near=np.zeros(3 * n) if Xn is not None else None,
               near_elem_cnt=np.zeros(ne, dtype=np.int64), near_elem_dsp=np.zeros(ne, dtype=np.int64),
               near_scatter_index=np.zeros(n, dtype=np.int64), near_trg_cnt=np.zeros(nt, dtype=np.int64),
               near_trg_dsp=np.zeros(nt, dtype=np.int64))
    m = fn(_p(Xt), nt, _p(Xn), _p(Xs), ns, _p(rad), ne, _ip(ec), _ip(ed), n, _p(out["Xtrg_near"]),
           _p(out["Xn_trg_near"]), _ip(out["near_elem_cnt"]), _ip(out["near_elem_dsp"]),
           _ip(out["near_scatter_index"]), _ip(out["near_trg_cnt"]), _ip(out["near_trg_dsp"]))
    if m != n:
        raise RuntimeError("near list size changed between calls (%d, %d)" % (n, m))
    return out


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
        L.oracle_eval_ctx.argtypes = L.oracle_eval.argtypes + [_dp]
        L.oracle_kernel_matrix_ctx.argtypes = L.oracle_kernel_matrix.argtypes + [_dp]
        L.oracle_rsqrt.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _L, _dp]
        L.oracle_dims.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_scale.argtypes = [ctypes.c_int]
        L.oracle_scale.restype = ctypes.c_double
        L.oracle_newton_iters_avx512.argtypes = [ctypes.c_int]
        _lp = ctypes.POINTER(ctypes.c_long)
        _ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_near_blocks.argtypes = [_L, _lp, _lp, ctypes.c_int, _lp, _lp, _lp]
        L.oracle_near_blocks.restype = _L
        L.oracle_near_layout.argtypes = [ctypes.c_int, ctypes.c_int, _L, _lp, _lp, _ip, _lp, _lp]
        L.oracle_near_layout.restype = _L
        L.oracle_near_subtract.argtypes = [ctypes.c_int, ctypes.c_int, _L] + [_lp] * 9 + [_ip] + [_dp] * 7
        L.oracle_near_gemv.argtypes = [ctypes.c_int, ctypes.c_int, _L] + [_lp] * 8 + [_dp] * 3
        L.oracle_near_accumulate.argtypes = [_L, ctypes.c_int, _lp, _lp, _L, _lp, _dp, _dp]
        L.oracle_near_list.argtypes = [_dp, _L, _dp, _dp, _L, _dp, _L, _lp, _lp, _L, _dp, _dp, _lp, _lp, _lp, _lp, _lp]
        L.oracle_near_list.restype = _L
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

    def eval(self, kernel, r_trg, r_src, n_src, v_src, v_trg=None, mode=EXACT, digits=-1, nthreads=None, ctx=None):
        """v_trg += Eval(...); returns v_trg (a fresh zero array when v_trg is None).
        ctx: the wavenumber of the Helmholtz kernels."""
        _, K0, K1, ND, _ = self.dims(kernel)
        r_trg, r_src, v_src = _c(r_trg), _c(r_src), _c(v_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        assert r_trg.size == 3 * nt and r_src.size == 3 * ns and v_src.size == K0 * ns
        assert not ND or n_src.size == 3 * ns
        out = np.zeros(nt * K1) if v_trg is None else np.ascontiguousarray(v_trg, dtype=np.float64).copy()
        nth = nthreads or os.cpu_count() or 1
        rc = self.lib.oracle_eval_ctx(kernel, mode, digits, _p(r_trg), nt, _p(r_src), _p(n_src), _p(v_src), ns,
                                      _p(out), nth, _ctx(ctx))
        if rc != 0:
            raise RuntimeError("oracle_eval failed rc=%d" % rc)
        return out

    def kernel_matrix(self, kernel, r_trg, r_src, n_src, nthreads=None, ctx=None):
        _, K0, K1, ND, _ = self.dims(kernel)
        r_trg, r_src = _c(r_trg), _c(r_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        M = np.zeros((ns * K0, nt * K1))
        rc = self.lib.oracle_kernel_matrix_ctx(kernel, _p(r_trg), nt, _p(r_src), _p(n_src), ns, _p(M),
                                               nthreads or os.cpu_count() or 1, _ctx(ctx))
        if rc != 0:
            raise RuntimeError("oracle_kernel_matrix failed rc=%d" % rc)
        return M

    # ------------------------------------------------------- near field
    # Arguments are the reference's own arrays (boundary_integral.txx).
    def near_blocks(self, near_elem_cnt, near_elem_dsp, nthreads):
        """(near_blk_elem, near_blk_t0, near_blk_cnt), boundary_integral.txx:1009-1032."""
        cnt, dsp = _i(near_elem_cnt), _i(near_elem_dsp)
        n = self.lib.oracle_near_blocks(cnt.size, _ip(cnt), _ip(dsp), nthreads, None, None, None)
        e, t0, c = (np.zeros(n, dtype=np.int64) for _ in range(3))
        self.lib.oracle_near_blocks(cnt.size, _ip(cnt), _ip(dsp), nthreads, _ip(e), _ip(t0), _ip(c))
        return e, t0, c

    def near_layout(self, kernel, normal_dot, near_elem_cnt, elem_nds_cnt, matrix_free=None):
        """(K_near_cnt, K_near_dsp, len(K_near)), boundary_integral.txx:1039-1057."""
        cnt, nds = _i(near_elem_cnt), _i(elem_nds_cnt)
        mf = None if matrix_free is None else np.ascontiguousarray(matrix_free, dtype=np.int32)
        kc, kd = np.zeros(cnt.size, dtype=np.int64), np.zeros(cnt.size, dtype=np.int64)
        n = self.lib.oracle_near_layout(kernel, int(normal_dot), cnt.size, _ip(cnt), _ip(nds),
                                        None if mf is None else mf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        _ip(kc), _ip(kd))
        if n < 0:
            raise ValueError("bad near layout")
        return kc, kd, n

    def near_subtract(self, kernel, normal_dot, blocks, near_elem_dsp, elem_nds_cnt, elem_nds_cnt_far,
                      elem_nds_dsp_far, K_near_cnt, K_near_dsp, Xtrg_near, Xn_trg_near, X_far, Xn_far, wts_far,
                      K_near, matrix_free=None, ctx=None):
        """KNearSubtract (boundary_integral.txx:1165-1240) on a copy of K_near."""
        be, bt, bc = (_i(a) for a in blocks)
        arrs = [_i(a) for a in (near_elem_dsp, elem_nds_cnt, elem_nds_cnt_far, elem_nds_dsp_far, K_near_cnt,
                                K_near_dsp)]
        mf = None if matrix_free is None else np.ascontiguousarray(matrix_free, dtype=np.int32)
        K = np.ascontiguousarray(K_near, dtype=np.float64).copy()
        f = [_c(a) for a in (Xtrg_near, Xn_trg_near, X_far, Xn_far, wts_far)]
        rc = self.lib.oracle_near_subtract(kernel, int(normal_dot), be.size, _ip(be), _ip(bt), _ip(bc),
                                           *[_ip(a) for a in arrs],
                                           None if mf is None else mf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           *[_p(a) for a in f], _p(K), _ctx(ctx))
        if rc != 0:
            raise RuntimeError("oracle_near_subtract failed rc=%d" % rc)
        return K

    def near_gemv(self, kernel, normal_dot, blocks, near_elem_dsp, elem_nds_cnt, elem_nds_dsp, K_near_cnt,
                  K_near_dsp, K_near, F, n_near):
        """U_near of ComputeNearInterac's matrix route (boundary_integral.txx:1324-1340)."""
        _, K0, K1, _, _ = self.dims(kernel)
        dof = K1 // 3 if normal_dot else K1
        be, bt, bc = (_i(a) for a in blocks)
        arrs = [_i(a) for a in (near_elem_dsp, elem_nds_cnt, elem_nds_dsp, K_near_cnt, K_near_dsp)]
        U = np.zeros(n_near * dof)
        rc = self.lib.oracle_near_gemv(kernel, int(normal_dot), be.size, _ip(be), _ip(bt), _ip(bc),
                                       *[_ip(a) for a in arrs], _p(_c(K_near)), _p(_c(F)), _p(U))
        if rc != 0:
            raise RuntimeError("oracle_near_gemv failed rc=%d" % rc)
        return U

    def near_accumulate(self, U, dof, near_trg_cnt, near_trg_dsp, scatter_index, U_near):
        """U += per-target sums after ScatterForward (boundary_integral.txx:1368-1379)."""
        cnt, dsp, sc = _i(near_trg_cnt), _i(near_trg_dsp), _i(scatter_index)
        out = np.ascontiguousarray(U, dtype=np.float64).copy()
        rc = self.lib.oracle_near_accumulate(cnt.size, dof, _ip(cnt), _ip(dsp), sc.size, _ip(sc),
                                             _p(_c(U_near)), _p(out))
        if rc != 0:
            raise RuntimeError("oracle_near_accumulate failed rc=%d" % rc)
        return out

    def near_list(self, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg=None):
        """BuildNearList (boundary_integral.txx:50-520), one process, brute force:
        dict of Xtrg_near, Xn_trg_near, near_elem_cnt, near_elem_dsp,
        near_scatter_index, near_trg_cnt, near_trg_dsp."""
        return _near_list(self.lib.oracle_near_list, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg)

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
                           "-DORACLE_MUTANT=%d" % index, "-o", out, os.path.join(HERE, "direct_sum.c"),
                           os.path.join(HERE, "near_field.c")])
    return out


# ------------------------------------------------------------------ reference
class Reference:
    """The reference library's own Eval/KernelMatrix (compiled reference)."""

    def __init__(self, flavour="v4"):
        path = os.path.join(REF_DIR, "libsctlref_%s.so" % flavour)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.sctlref_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _L, _dp, _dp, _dp, _L, _dp, _dp]
        L.sctlref_eval_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _L, _fp, _fp, _fp, _L, _fp,
                                       _dp]
        L.sctlref_kernel_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _L, _dp, _dp, _L, _dp, _dp]
        L.sctlref_rsqrt.argtypes = [ctypes.c_int, _dp, _L, _dp]
        _lp = ctypes.POINTER(ctypes.c_long)
        L.sctlref_blocks.argtypes = [ctypes.c_int, ctypes.c_int, _L, _lp, _lp, _lp, _lp, _dp, _dp, _dp, _dp, _dp, _dp]
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

    def eval(self, kernel, r_trg, r_src, n_src, v_src, v_trg=None, digits=-1, omp=True, ctx=None):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        r_trg, r_src, v_src = _c(r_trg), _c(r_src), _c(v_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        out = np.zeros(nt * K1) if v_trg is None else np.ascontiguousarray(v_trg, dtype=np.float64).copy()
        rc = self.lib.sctlref_eval(kernel, digits, int(omp), _p(r_trg), nt, _p(r_src), _p(n_src), _p(v_src), ns,
                                   _p(out), _ctx(ctx))
        if rc != 0:
            raise RuntimeError("sctlref_eval failed rc=%d" % rc)
        return out

    def eval_f32(self, kernel, r_trg, r_src, n_src, v_src, digits=-1, omp=True, ctx=None):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float32)
        r_trg, r_src, v_src = f(r_trg), f(r_src), f(v_src)
        n_src = f(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        out = np.zeros(nt * K1, dtype=np.float32)
        q = lambda a: a.ctypes.data_as(_fp) if a is not None and a.size else None
        rc = self.lib.sctlref_eval_f32(kernel, digits, int(omp), q(r_trg), nt, q(r_src), q(n_src), q(v_src), ns,
                                       q(out), _ctx(ctx))
        if rc != 0:
            raise RuntimeError("sctlref_eval_f32 failed rc=%d" % rc)
        return out

    def kernel_matrix(self, kernel, r_trg, r_src, n_src, digits=-1, ctx=None):
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        r_trg, r_src = _c(r_trg), _c(r_src)
        n_src = _c(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        M = np.zeros((ns * K0, nt * K1))
        rc = self.lib.sctlref_kernel_matrix(kernel, digits, _p(r_trg), nt, _p(r_src), _p(n_src), ns, _p(M),
                                            _ctx(ctx))
        if rc != 0:
            raise RuntimeError("sctlref_kernel_matrix failed rc=%d" % rc)
        return M

    def rsqrt(self, x, digits=-1):
        x = _c(x)
        y = np.zeros_like(x)
        if self.lib.sctlref_rsqrt(digits, _p(x), x.size, _p(y)) != 0:
            raise ValueError("digits %r not instantiated" % digits)
        return y

    def blocks(self, kernel, mode, trg_dsp, trg_cnt, src_dsp, src_cnt, r_trg, r_src, n_src, v_src=None,
               v_trg=None):
        """BoundaryIntegralOp's per-block OpenMP loop (boundary_integral.txx:1165-1211): Eval<double,false>
        per block into v_trg (mode 0, returned), or KernelMatrix<double,false> per block (mode 1; returns a
        checksum of the matrices, which are not kept)."""
        (_, K0, K1, ND, _), _, _ = self.info(kernel)
        a = [np.ascontiguousarray(x, dtype=np.int64) for x in (trg_dsp, trg_cnt, src_dsp, src_cnt)]
        lp = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_long))
        r_trg, r_src = _c(r_trg), _c(r_src)
        n_src = _c(n_src) if ND else None
        v_src = _c(v_src) if v_src is not None else None
        if mode == 0 and v_trg is None:
            v_trg = np.zeros(r_trg.size // 3 * K1)
        chk = ctypes.c_double(0)
        rc = self.lib.sctlref_blocks(kernel, mode, a[0].size, lp(a[0]), lp(a[1]), lp(a[2]), lp(a[3]), _p(r_trg),
                                     _p(r_src), _p(n_src), _p(v_src), _p(v_trg), ctypes.pointer(chk))
        if rc != 0:
            raise RuntimeError("sctlref_blocks failed rc=%d" % rc)
        return v_trg if mode == 0 else chk.value


class ReferenceNearList:
    """The reference library's own BuildNearList (compiled reference,
    oracle/ref_near_shim.cpp)."""

    def __init__(self):
        path = os.path.join(REF_DIR, "libsctlref_near.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        _lp = ctypes.POINTER(ctypes.c_long)
        L.sctlref_build_near_list.argtypes = [_dp, _L, _dp, _dp, _L, _dp, _L, _lp, _lp, _L, _lp, _dp, _dp, _lp, _lp,
                                              _lp, _lp, _lp]
        self.lib = L

    @staticmethod
    def available():
        return os.path.exists(os.path.join(REF_DIR, "libsctlref_near.so"))

    def build_seconds(self, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp):
        """(seconds, pairs) of one BuildNearList call (the sizing call builds the whole list)."""
        import time
        Xt, Xs, rad = _c(Xtrg), _c(Xsrc), _c(src_radius)
        ec, ed = _i(elem_nds_cnt), _i(elem_nds_dsp)
        n = ctypes.c_long()
        t = time.perf_counter()
        self.lib.sctlref_build_near_list(_p(Xt), Xt.size // 3, None, _p(Xs), Xs.size // 3, _p(rad), ec.size, _ip(ec),
                                         _ip(ed), 0, ctypes.byref(n), None, None, None, None, None, None, None)
        return time.perf_counter() - t, n.value

    def near_list(self, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg=None):
        def fn(xt, nt, xn, xs, ns, rad, ne, ec, ed, max_near, *outs):
            n = ctypes.c_long()
            rc = self.lib.sctlref_build_near_list(xt, nt, xn, xs, ns, rad, ne, ec, ed, max_near, ctypes.byref(n),
                                                  *outs)
            return n.value if rc in (0, 1) else -1
        return _near_list(fn, Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg)

