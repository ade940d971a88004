"""sctl_amd -- MI355X (gfx950) direct kernel summation, Python binding.

A thin ctypes layer over the C-ABI of ``libsctl_amd.so`` (include/sctl_amd.h),
used by the tests and the benchmark.  The drop-in for C++ callers of SCTL is
``include/sctl_amd/generic_kernel.hpp``; this module mirrors the same
operations for Python hosts:

  * ``GenericKernel(kernel).Eval(v_trg, r_trg, r_src, n_src, v_src)``
    -- ``GenericKernel<uKernel>::Eval`` (generic-kernel.txx:102-211): resizes
    and zeroes ``v_trg`` when its length is wrong, otherwise accumulates.
  * ``GenericKernel(kernel).KernelMatrix(r_trg, r_src, n_src)``
    -- ``GenericKernel<uKernel>::KernelMatrix`` (generic-kernel.txx:213-329).
  * ``Plan`` -- batched near-field blocks in one persistent launch.

The library is required: importing works without it, but every call raises
``LibraryNotBuilt`` (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsctl_amd.so")

LAPLACE3D_FXU, LAPLACE3D_DXU, STOKES3D_FXU, STOKES3D_DXU = 0, 1, 2, 3
LAPLACE3D_FXDU, LAPLACE3D_DXDU, LAPLACE3D_FXD2U = 4, 5, 6
STOKES3D_FXT, STOKES3D_FSXU, STOKES3D_FXUP = 7, 8, 9
BIOTSAVART3D_FXU, BIOTSAVART3D_FXDU = 10, 11
HELMHOLTZ3D_FXU, HELMHOLTZ3D_DXU, HELMHOLTZ3D_FXDU, HELMHOLTZDIFF3D_FXDU = 12, 13, 14, 15
KERNEL_NAMES = {0: "Laplace3D_FxU", 1: "Laplace3D_DxU", 2: "Stokes3D_FxU", 3: "Stokes3D_DxU",
                4: "Laplace3D_FxdU", 5: "Laplace3D_DxdU", 6: "Laplace3D_Fxd2U", 7: "Stokes3D_FxT",
                8: "Stokes3D_FSxU", 9: "Stokes3D_FxUP", 10: "BiotSavart3D_FxU", 11: "BiotSavart3D_FxdU",
                12: "Helmholtz3D_FxU", 13: "Helmholtz3D_DxU", 14: "Helmholtz3D_FxdU", 15: "HelmholtzDiff3D_FxdU"}
# Kernels that take a context value (the wavenumber), as GenericKernel::SetCtxPtr.
CTX_KERNELS = frozenset(range(12, 16))
NUM_KERNELS = len(KERNEL_NAMES)
FP64, MIXED, AUTO = 0, 1, 2
HOST_PTRS, MIXED_NEWTON = 0x1, 0x2
NEAR_SUBTRACT, NEAR_STORE = 0, 1

STATUS = {0: "SUCCESS", 1: "INVALID_VALUE", 2: "UNKNOWN_KERNEL", 3: "UNSUPPORTED", 4: "NO_DEVICE",
          5: "OUT_OF_MEMORY", 6: "HIP", 7: "QUEUE"}

# Exported symbols declared in include/sctl_amd.h (checked by the tests).
SYMBOLS = [
    "sctl_amd_init", "sctl_amd_last_error", "sctl_amd_version", "sctl_amd_kernel_info", "sctl_amd_eval",
    "sctl_amd_kernel_matrix", "sctl_amd_plan_create", "sctl_amd_plan_set_geometry", "sctl_amd_plan_apply",
    "sctl_amd_plan_interactions", "sctl_amd_plan_destroy", "sctl_amd_malloc", "sctl_amd_free", "sctl_amd_memcpy",
    "sctl_amd_stream_create", "sctl_amd_stream_destroy", "sctl_amd_stream_synchronize", "sctl_amd_profile_enable",
    "sctl_amd_profile_read", "sctl_amd_profile_reset", "sctl_amd_plan_graph_create", "sctl_amd_graph_launch",
    "sctl_amd_graph_interactions", "sctl_amd_graph_destroy", "sctl_amd_gen_normal", "sctl_amd_plan_get_shape",
    "sctl_amd_plan_set_shape", "sctl_amd_tune_db_path", "sctl_amd_tune_eval", "sctl_amd_plan_tune",
    "sctl_amd_near_blocks", "sctl_amd_plan_near_layout", "sctl_amd_plan_near_matrix", "sctl_amd_plan_near_gemv",
    "sctl_amd_near_accumulate", "sctl_amd_dense_shape_get", "sctl_amd_dense_shape_set", "sctl_amd_near_list_build",
    "sctl_amd_near_list_size", "sctl_amd_near_list_get", "sctl_amd_near_list_destroy", "sctl_amd_eval_ctx",
    "sctl_amd_kernel_matrix_ctx", "sctl_amd_plan_set_ctx", "sctl_amd_dense_plan_create", "sctl_amd_dense_plan_apply",
    "sctl_amd_dense_plan_shape", "sctl_amd_dense_plan_destroy", "sctl_amd_dense_plan_graph_create",
    "sctl_amd_eval_noalloc", "sctl_amd_eval_noalloc_scratch_bytes", "sctl_amd_hilbert_order",
]


class LibraryNotBuilt(RuntimeError):
    pass


class SctlAmdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(status, "?"), status, msg))
        self.status = status


_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_lib = None


def lib():
    """Load libsctl_amd.so (built in tree by sctl_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryNotBuilt("%s is missing: run `python -m sctl_amd.build` (or __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    i32, i64, u32 = ctypes.c_int, ctypes.c_int64, ctypes.c_uint32
    L.sctl_amd_last_error.restype = ctypes.c_char_p
    L.sctl_amd_version.restype = ctypes.c_char_p
    L.sctl_amd_init.argtypes = [i32]
    L.sctl_amd_kernel_info.argtypes = [i32] + [ctypes.POINTER(i32)] * 5 + [_dp, ctypes.POINTER(ctypes.c_char_p)]
    L.sctl_amd_eval.argtypes = [i32, i32, i32, _vp, i64, _vp, _vp, _vp, i64, _vp, u32, _vp]
    L.sctl_amd_eval_ctx.argtypes = [i32, _dp, i32, i32, _vp, i64, _vp, _vp, _vp, i64, _vp, u32, _vp]
    L.sctl_amd_eval_noalloc.argtypes = [i32, _dp, i32, i32, _vp, i64, _vp, _vp, _vp, i64, _vp, u32, _vp]
    L.sctl_amd_eval_noalloc_scratch_bytes.restype = ctypes.c_size_t
    L.sctl_amd_kernel_matrix.argtypes = [i32, i32, _vp, i64, _vp, _vp, i64, _vp, u32, _vp]
    L.sctl_amd_kernel_matrix_ctx.argtypes = [i32, _dp, i32, _vp, i64, _vp, _vp, i64, _vp, u32, _vp]
    L.sctl_amd_plan_set_ctx.argtypes = [_vp, _dp]
    L.sctl_amd_dense_plan_create.argtypes = [ctypes.POINTER(_vp), i32, i32, i32, u32, _vp, i64, _vp, _vp, i64, _dp, _vp]
    L.sctl_amd_dense_plan_apply.argtypes = [_vp, _vp, _vp, u32, _vp]
    L.sctl_amd_dense_plan_shape.argtypes = [_vp] + [ctypes.POINTER(i32)] * 3
    L.sctl_amd_dense_plan_destroy.argtypes = [_vp]
    L.sctl_amd_dense_plan_graph_create.argtypes = [ctypes.POINTER(_vp), _vp, i32, ctypes.c_uint64, ctypes.c_uint64,
                                                   _vp, _vp]
    L.sctl_amd_plan_create.argtypes = [ctypes.POINTER(_vp), i32, i32, i32, i64, _i64p, _i64p, _i64p, _i64p, i64, i64]
    L.sctl_amd_plan_set_geometry.argtypes = [_vp, _vp, _vp, _vp, u32, _vp]
    L.sctl_amd_plan_apply.argtypes = [_vp, _vp, _vp, u32, _vp]
    L.sctl_amd_plan_interactions.argtypes = [_vp, _i64p]
    L.sctl_amd_plan_destroy.argtypes = [_vp]
    L.sctl_amd_malloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
    L.sctl_amd_free.argtypes = [_vp]
    L.sctl_amd_memcpy.argtypes = [_vp, _vp, ctypes.c_size_t, i32, _vp]
    L.sctl_amd_stream_create.argtypes = [ctypes.POINTER(_vp)]
    L.sctl_amd_stream_destroy.argtypes = [_vp]
    L.sctl_amd_stream_synchronize.argtypes = [_vp]
    L.sctl_amd_plan_graph_create.argtypes = [ctypes.POINTER(_vp), _vp, i32, ctypes.c_uint64, ctypes.c_uint64, _vp, _vp]
    L.sctl_amd_graph_launch.argtypes = [_vp, _vp]
    L.sctl_amd_graph_interactions.argtypes = [_vp, _i64p]
    L.sctl_amd_graph_destroy.argtypes = [_vp]
    L.sctl_amd_gen_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, i64, _vp, _vp]
    L.sctl_amd_hilbert_order.argtypes = [_vp, i64, _vp, _vp, i64, _vp, _vp]
    ip = ctypes.POINTER(i32)
    L.sctl_amd_plan_get_shape.argtypes = [_vp, ip, ip, ip]
    L.sctl_amd_plan_set_shape.argtypes = [_vp, i32, i32, i32]
    L.sctl_amd_tune_db_path.argtypes = [ctypes.c_char_p]
    L.sctl_amd_tune_eval.argtypes = [i32, i32, i32, _vp, i64, _vp, _vp, _vp, i64, _vp, u32, _vp, ip, ip, _dp]
    L.sctl_amd_dense_shape_get.argtypes = [i32, i32, i32, u32, i64, i64, ip, ip]
    L.sctl_amd_dense_shape_set.argtypes = [i32, i32, i32, u32, i64, i64, i32, i32]
    L.sctl_amd_plan_tune.argtypes = [_vp, _vp, _vp, _vp, ip, ip, ip, _dp]
    L.sctl_amd_near_blocks.argtypes = [i64, _i64p, _i64p, i32, _i64p, _i64p, _i64p, _i64p]
    L.sctl_amd_plan_near_layout.argtypes = [_vp, _i64p, i32, _i64p]
    L.sctl_amd_plan_near_matrix.argtypes = [_vp, i32, _vp, _vp, _vp, u32, _vp]
    L.sctl_amd_plan_near_gemv.argtypes = [_vp, _vp, _vp, _vp, u32, _vp]
    L.sctl_amd_near_accumulate.argtypes = [i64, i32, _vp, _vp, i64, _vp, i64, _vp, _vp, u32, _vp]
    L.sctl_amd_near_list_build.argtypes = [ctypes.POINTER(_vp), _vp, i64, _vp, _vp, i64, _vp, i64, _vp, _vp, u32, _vp]
    L.sctl_amd_near_list_size.argtypes = [_vp, _i64p]
    L.sctl_amd_near_list_get.argtypes = [_vp] + [_vp] * 7 + [u32, _vp]
    L.sctl_amd_near_list_destroy.argtypes = [_vp]
    L.sctl_amd_profile_enable.argtypes = [i32]
    L.sctl_amd_profile_read.argtypes = [_i64p, _i64p, _i64p, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def check(status):
    if status != 0:
        raise SctlAmdError(status, lib().sctl_amd_last_error().decode())


def kernel_info(kernel):
    """(coord_dim, src_dim, trg_dim, normal_dim, flops, scale, name)."""
    c = [ctypes.c_int() for _ in range(5)]
    s = ctypes.c_double()
    n = ctypes.c_char_p()
    check(lib().sctl_amd_kernel_info(kernel, *[ctypes.byref(x) for x in c], ctypes.byref(s), ctypes.byref(n)))
    return tuple(x.value for x in c) + (s.value, n.value.decode())


# ------------------------------------------------------------ device plumbing
class DeviceArray:
    """A device buffer owned by libsctl_amd (sctl_amd_malloc/free)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = (shape,) if np.isscalar(shape) else tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        check(lib().sctl_amd_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        d.upload(a)
        return d

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        if self.nbytes:
            check(lib().sctl_amd_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0, stream))

    def download(self, stream=None):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(lib().sctl_amd_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1, stream))
        return out

    def free(self):
        if self.ptr:
            lib().sctl_amd_free(self.ptr)
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def profile_enable(on=True):
    check(lib().sctl_amd_profile_enable(1 if on else 0))


def profile_reset():
    check(lib().sctl_amd_profile_reset())


def profile_read():
    """dict(interactions, flops, timed_launches, kernel_ms); waits for pending timing events."""
    a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    d = ctypes.c_double()
    check(lib().sctl_amd_profile_read(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
    return dict(interactions=a.value, flops=b.value, timed_launches=c.value, kernel_ms=d.value)


def synchronize(stream=None):
    check(lib().sctl_amd_stream_synchronize(stream))


def _host(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _ctx(ctx):
    """Pointer to one double (the wavenumber of the Helmholtz kernels), or NULL."""
    return None if ctx is None else ctypes.pointer(ctypes.c_double(float(ctx)))


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data if a.size else None


# ----------------------------------------------------------------- Eval API
def eval(kernel, r_trg, r_src, n_src, v_src, v_trg=None, precision=FP64, digits=-1, flags=0, stream=None, ctx=None):
    """v_trg += Eval(...).

    With numpy arrays the call stages through the device and returns the
    (new, or updated in place) host v_trg.  With DeviceArray arguments the call
    is asynchronous on `stream` and returns v_trg.  ctx: the wavenumber of the
    Helmholtz kernels."""
    _, K0, K1, ND, _, _, _ = kernel_info(kernel)
    dev = isinstance(r_trg, DeviceArray)
    if not dev:
        r_trg, r_src, v_src = _host(r_trg), _host(r_src), _host(v_src)
        n_src = _host(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        if v_trg is None:
            v_trg = np.zeros(nt * K1)
        flags |= HOST_PTRS
    else:
        nt, ns = r_trg.nbytes // 24, r_src.nbytes // 24
        if not ND:
            n_src = None
    check(lib().sctl_amd_eval_ctx(kernel, _ctx(ctx), precision, digits, _ptr(r_trg), nt, _ptr(r_src), _ptr(n_src),
                                  _ptr(v_src), ns, _ptr(v_trg), flags, stream))
    return v_trg


def kernel_matrix(kernel, r_trg, r_src, n_src, digits=-1, M=None, flags=0, stream=None, ctx=None):
    _, K0, K1, ND, _, _, _ = kernel_info(kernel)
    dev = isinstance(r_trg, DeviceArray)
    if not dev:
        r_trg, r_src = _host(r_trg), _host(r_src)
        n_src = _host(n_src) if ND else None
        nt, ns = r_trg.size // 3, r_src.size // 3
        if M is None:
            M = np.zeros((ns * K0, nt * K1))
        flags |= HOST_PTRS
    else:
        nt, ns = r_trg.nbytes // 24, r_src.nbytes // 24
        if not ND:
            n_src = None
    check(lib().sctl_amd_kernel_matrix_ctx(kernel, _ctx(ctx), digits, _ptr(r_trg), nt, _ptr(r_src), _ptr(n_src), ns,
                                           _ptr(M), flags, stream))
    return M


class GenericKernel:
    """Python mirror of sctl::GenericKernel<uKernel> (SetCtxPtr takes the
    wavenumber itself)."""

    def __init__(self, kernel):
        (self._coord, self._k0, self._k1, self._nd, self._flops, self._scale,
         self._name) = kernel_info(kernel)
        self.kernel = kernel
        self._ctx = None

    def SetCtxPtr(self, ctx):
        self._ctx = ctx

    def GetCtxPtr(self):
        return self._ctx

    def CoordDim(self):
        return self._coord

    def SrcDim(self):
        return self._k0

    def TrgDim(self):
        return self._k1

    def NormalDim(self):
        return self._nd

    def Name(self):
        return self._name

    def FLOPS(self):
        return self._flops

    def ScaleFactor(self):
        return self._scale

    def Eval(self, v_trg, r_trg, r_src, n_src, v_src, digits=-1, precision=FP64):
        """Returns v_trg: accumulated into when its length is Nt*TrgDim, otherwise a fresh zeroed vector
        (generic-kernel.txx:129-132).  Shape mismatches raise ValueError (SCTL_ASSERT, :125-128)."""
        r_trg = _host(r_trg).ravel()
        r_src = _host(r_src).ravel()
        v_src = _host(v_src).ravel()
        nt, ns = r_trg.size // 3, r_src.size // 3
        if r_trg.size != nt * 3 or r_src.size != ns * 3 or v_src.size != ns * self._k0:
            raise ValueError("shape mismatch")
        if self._nd:
            n_src = _host(n_src).ravel()
            if n_src.size != ns * self._nd:
                raise ValueError("normal shape mismatch")
        v = None if v_trg is None else _host(v_trg).ravel()
        if v is None or v.size != nt * self._k1:
            v = np.zeros(nt * self._k1)
        else:
            v = v.copy()
        return eval(self.kernel, r_trg, r_src, n_src, v_src, v, precision=precision, digits=digits, ctx=self._ctx)

    def KernelMatrix(self, r_trg, r_src, n_src, digits=-1):
        return kernel_matrix(self.kernel, r_trg, r_src, n_src, digits=digits, ctx=self._ctx)


# ------------------------------------------------------------------ batched
class Plan:
    """Batched (target-set, source-set) blocks evaluated in one persistent launch."""

    def __init__(self, kernel, trg_dsp, trg_cnt, src_dsp, src_cnt, nt_total, ns_total, precision=FP64, digits=-1,
                 ctx=None):
        a = [np.ascontiguousarray(x, dtype=np.int64) for x in (trg_dsp, trg_cnt, src_dsp, src_cnt)]
        self._keep = a
        self.kernel = kernel
        self.nblocks = a[0].size
        h = _vp()
        P = lambda x: x.ctypes.data_as(_i64p)
        check(lib().sctl_amd_plan_create(ctypes.byref(h), kernel, precision, digits, self.nblocks, P(a[0]), P(a[1]),
                                         P(a[2]), P(a[3]), int(nt_total), int(ns_total)))
        self.handle = h
        if ctx is not None:
            self.set_ctx(ctx)

    def set_ctx(self, ctx):
        """The wavenumber of the Helmholtz kernels (None clears it)."""
        check(lib().sctl_amd_plan_set_ctx(self.handle, _ctx(ctx)))

    def set_geometry(self, r_trg, r_src, n_src=None, stream=None):
        flags = 0
        if not isinstance(r_trg, DeviceArray):
            r_trg, r_src, n_src = _host(r_trg), _host(r_src), _host(n_src)
            flags = HOST_PTRS
        check(lib().sctl_amd_plan_set_geometry(self.handle, _ptr(r_trg), _ptr(r_src), _ptr(n_src), flags, stream))

    def apply(self, v_src, v_trg, stream=None):
        flags = 0
        if not isinstance(v_src, DeviceArray):
            v_src = _host(v_src)
            flags = HOST_PTRS
        check(lib().sctl_amd_plan_apply(self.handle, _ptr(v_src), _ptr(v_trg), flags, stream))
        return v_trg

    def interactions(self):
        n = ctypes.c_int64()
        check(lib().sctl_amd_plan_interactions(self.handle, ctypes.byref(n)))
        return n.value

    def shape(self):
        """(sources_per_lane, targets_per_lane, wg_per_cu_cap)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().sctl_amd_plan_get_shape(self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_shape(self, sources_per_lane, targets_per_lane, wg_per_cu_cap=0):
        check(lib().sctl_amd_plan_set_shape(self.handle, sources_per_lane, targets_per_lane, wg_per_cu_cap))

    def tune(self, v_src, v_scratch, stream=None):
        """Autotunes the launch shape on these device densities; returns (spl, tpl, cap, ms)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().sctl_amd_plan_tune(self.handle, v_src.ptr, v_scratch.ptr, stream, ctypes.byref(a), ctypes.byref(b),
                                       ctypes.byref(c), ctypes.byref(ms)))
        return a.value, b.value, c.value, ms.value

    def graph(self, nsteps, seed, stream_id, v_src, v_trg):
        """Captures nsteps applies with densities regenerated on the device (see Graph)."""
        return Graph(self, nsteps, seed, stream_id, v_src, v_trg)

    # ------------------------------------------------- near-field operator
    def near_layout(self, offsets=None, trg_normal_dot=False):
        """Places block b's target-major matrix at offsets[b] of K_near (-1: none;
        None: packed in block order).  Returns the length K_near must have."""
        n = ctypes.c_int64()
        off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int64)
        self._near_off = off
        check(lib().sctl_amd_plan_near_layout(self.handle, None if off is None else off.ctypes.data_as(_i64p),
                                              int(bool(trg_normal_dot)), ctypes.byref(n)))
        self.near_len = n.value
        return n.value

    def near_matrix(self, K, wts=None, n_trg=None, mode=NEAR_SUBTRACT, stream=None):
        """K -= w * M^T per block (mode NEAR_SUBTRACT) or K = w * M^T (NEAR_STORE).
        Host K is updated in place; a DeviceArray K asynchronously on `stream`."""
        flags = 0
        if not isinstance(K, DeviceArray):
            assert isinstance(K, np.ndarray) and K.dtype == np.float64 and K.flags.c_contiguous
            wts, n_trg = _host(wts), _host(n_trg)
            flags = HOST_PTRS
        check(lib().sctl_amd_plan_near_matrix(self.handle, mode, _ptr(wts), _ptr(n_trg), _ptr(K), flags, stream))
        return K

    def near_gemv(self, K, F, U_near, stream=None):
        """U_near[block rows] = K_block * F[block sources] (overwrites those rows)."""
        flags = 0
        if not isinstance(K, DeviceArray):
            assert isinstance(U_near, np.ndarray) and U_near.dtype == np.float64 and U_near.flags.c_contiguous
            K, F = _host(K), _host(F)
            flags = HOST_PTRS
        check(lib().sctl_amd_plan_near_gemv(self.handle, _ptr(K), _ptr(F), _ptr(U_near), flags, stream))
        return U_near

    def destroy(self):
        if self.handle:
            lib().sctl_amd_plan_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class DensePlan:
    """One (targets, sources) Eval with resident geometry: create once, apply
    with new densities (the far field of a GMRES loop)."""

    def __init__(self, kernel, r_trg, r_src, n_src=None, precision=FP64, digits=-1, flags=0, ctx=None, stream=None):
        _, K0, K1, ND, _, _, _ = kernel_info(kernel)
        self.kernel, self.K0, self.K1 = kernel, K0, K1
        if isinstance(r_trg, DeviceArray):
            nt, ns = r_trg.nbytes // 24, r_src.nbytes // 24
        else:
            r_trg, r_src = _host(r_trg), _host(r_src)
            n_src = _host(n_src) if ND else None
            nt, ns = r_trg.size // 3, r_src.size // 3
            flags |= HOST_PTRS
        if not ND:
            n_src = None
        self.nt, self.ns = nt, ns
        h = _vp()
        check(lib().sctl_amd_dense_plan_create(ctypes.byref(h), kernel, precision, digits, flags, _ptr(r_trg), nt,
                                               _ptr(r_src), _ptr(n_src), ns, _ctx(ctx), stream))
        self.handle = h

    def apply(self, v_src, v_trg=None, stream=None):
        """v_trg += Eval(v_src); host arrays are staged (and v_trg returned),
        DeviceArrays run asynchronously on `stream`."""
        flags = 0
        if not isinstance(v_src, DeviceArray):
            v_src = _host(v_src)
            if v_trg is None:
                v_trg = np.zeros(self.nt * self.K1)
            flags = HOST_PTRS
        check(lib().sctl_amd_dense_plan_apply(self.handle, _ptr(v_src), _ptr(v_trg), flags, stream))
        return v_trg

    def shape(self):
        """(precision, targets_per_lane, wg_per_cu_cap)."""
        v = [ctypes.c_int() for _ in range(3)]
        check(lib().sctl_amd_dense_plan_shape(self.handle, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def graph(self, nsteps, seed, stream_id, v_src, v_trg):
        """nsteps applies with densities regenerated on the device, as one hipGraph."""
        return Graph(self, nsteps, seed, stream_id, v_src, v_trg)

    def destroy(self):
        if self.handle:
            lib().sctl_amd_dense_plan_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Graph:
    """nsteps applications of a plan captured as one hipGraph: for step s the
    device buffer v_src is refilled with N(0,1) draws of (seed + s, stream_id),
    then v_trg += apply(v_src)."""

    def __init__(self, plan, nsteps, seed, stream_id, v_src, v_trg):
        h = _vp()
        create = lib().sctl_amd_dense_plan_graph_create if isinstance(plan, DensePlan) else \
            lib().sctl_amd_plan_graph_create
        check(create(ctypes.byref(h), plan.handle, nsteps, seed, stream_id, v_src.ptr, v_trg.ptr))
        self.handle = h
        self._keep = (plan, v_src, v_trg)

    def launch(self, stream=None):
        check(lib().sctl_amd_graph_launch(self.handle, stream))

    def interactions(self):
        n = ctypes.c_int64()
        check(lib().sctl_amd_graph_interactions(self.handle, ctypes.byref(n)))
        return n.value

    def destroy(self):
        if self.handle:
            lib().sctl_amd_graph_destroy(self.handle)
            self.handle = _vp()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def tune_db_path(path):
    check(lib().sctl_amd_tune_db_path(None if path is None else str(path).encode()))


def tune_eval(kernel, r_trg, r_src, n_src, v_src, v_scratch, precision=FP64, digits=-1, flags=0, stream=None):
    """Autotunes the dense Eval shape for these device arrays; returns
    (targets_per_lane, wg_per_cu_cap, ms)."""
    nt, ns = r_trg.nbytes // 24, r_src.nbytes // 24
    a, b = ctypes.c_int(), ctypes.c_int()
    ms = ctypes.c_double()
    check(lib().sctl_amd_tune_eval(kernel, precision, digits, r_trg.ptr, nt, r_src.ptr, _ptr(n_src), v_src.ptr, ns,
                                   v_scratch.ptr, flags, stream, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ms)))
    return a.value, b.value, ms.value


def dense_shape(kernel, nt, ns, precision=FP64, digits=-1, flags=0):
    """(targets_per_lane, wg_per_cu_cap) that Eval uses for this size class."""
    a, b = ctypes.c_int(), ctypes.c_int()
    check(lib().sctl_amd_dense_shape_get(kernel, precision, digits, flags, int(nt), int(ns), ctypes.byref(a),
                                         ctypes.byref(b)))
    return a.value, b.value


def set_dense_shape(kernel, nt, ns, targets_per_lane, wg_per_cu_cap=0, precision=FP64, digits=-1, flags=0):
    """Records the Eval launch shape for this size class in the tuning database."""
    check(lib().sctl_amd_dense_shape_set(kernel, precision, digits, flags, int(nt), int(ns), targets_per_lane,
                                         wg_per_cu_cap))


def gen_normal(seed, stream_id, out, stream=None):
    """Device generator: fills DeviceArray `out` with N(0,1) draws of (seed, stream_id)."""
    check(lib().sctl_amd_gen_normal(seed, stream_id, out.nbytes // 8, out.ptr, stream))
    return out


def hilbert_order(r_a, r_b):
    """Hilbert-curve orders of two host point sets (flat x,y,z) as the mixed path
    sums them: (perm_a, perm_b) as uint32 arrays."""
    r_a = np.ascontiguousarray(r_a, dtype=np.float64).ravel()
    r_b = np.ascontiguousarray(r_b, dtype=np.float64).ravel()
    na, nb = r_a.size // 3, r_b.size // 3
    da, db = DeviceArray.from_host(r_a), DeviceArray.from_host(r_b)
    pa, pb = DeviceArray(na, np.uint32), DeviceArray(nb, np.uint32)
    check(lib().sctl_amd_hilbert_order(da.ptr, na, pa.ptr, db.ptr, nb, pb.ptr, None))
    return pa.download(), pb.download()


# -------------------------------------------------------------- near field
def near_blocks(near_elem_cnt, near_elem_dsp, nthreads):
    """BoundaryIntegralOp's near-block partition (boundary_integral.txx:1009-1032):
    returns (near_blk_elem, near_blk_t0, near_blk_cnt)."""
    cnt = np.ascontiguousarray(near_elem_cnt, dtype=np.int64)
    dsp = np.ascontiguousarray(near_elem_dsp, dtype=np.int64)
    n = ctypes.c_int64()
    P = lambda x: x.ctypes.data_as(_i64p)
    check(lib().sctl_amd_near_blocks(cnt.size, P(cnt), P(dsp), int(nthreads), ctypes.byref(n), None, None, None))
    e, t0, c = (np.zeros(n.value, dtype=np.int64) for _ in range(3))
    check(lib().sctl_amd_near_blocks(cnt.size, P(cnt), P(dsp), int(nthreads), ctypes.byref(n), P(e), P(t0), P(c)))
    return e, t0, c


def near_accumulate(U, dof, near_trg_cnt, near_trg_dsp, scatter_index, U_near, n_near=None, stream=None):
    """U[i] += sum of U_near over target i's near elements after ScatterForward
    (boundary_integral.txx:1368-1379).  Host arrays: U is updated in place and
    returned.  DeviceArrays: asynchronous on `stream` (give n_near)."""
    if isinstance(U, DeviceArray):
        ntrg = U.shape[0] // dof
        n_sc = scatter_index.shape[0]
        n_near = U_near.shape[0] // dof if n_near is None else n_near
        check(lib().sctl_amd_near_accumulate(ntrg, dof, near_trg_cnt.ptr, near_trg_dsp.ptr, n_sc, scatter_index.ptr,
                                             n_near, U_near.ptr, U.ptr, 0, stream))
        return U
    assert isinstance(U, np.ndarray) and U.dtype == np.float64 and U.flags.c_contiguous
    cnt = np.ascontiguousarray(near_trg_cnt, dtype=np.int64)
    dsp = np.ascontiguousarray(near_trg_dsp, dtype=np.int64)
    sc = np.ascontiguousarray(scatter_index, dtype=np.int64)
    un = _host(U_near)
    check(lib().sctl_amd_near_accumulate(U.size // dof, dof, _ptr(cnt), _ptr(dsp), sc.size, _ptr(sc), un.size // dof,
                                         _ptr(un), _ptr(U), HOST_PTRS, stream))
    return U


def near_offsets(K0, dof, blocks, near_elem_cnt, elem_nds_cnt, matrix_free=None):
    """Offsets of each block's rows in the reference's K_near and its length.

    K_near_cnt[e] = elem_nds_cnt[e]*near_elem_cnt[e] (0 for a matrix-free
    element), K_near_dsp its exclusive scan (boundary_integral.txx:1039-1053);
    block (e, t0, cnt) starts at K_near_dsp[e]*KDIM0*KDIM1_ + t0*KDIM1_*src_dof
    (:1187-1189), -1 for a matrix-free element.  dof is KDIM1_."""
    be, bt, _ = (np.asarray(a, dtype=np.int64) for a in blocks)
    cnt = np.asarray(near_elem_cnt, dtype=np.int64)
    nds = np.asarray(elem_nds_cnt, dtype=np.int64)
    mf = np.zeros(cnt.size, dtype=bool) if matrix_free is None else np.asarray(matrix_free, dtype=bool)
    if cnt.size == 0:
        return np.zeros(be.size, dtype=np.int64), 0
    k_cnt = np.where(mf, 0, nds * cnt)
    k_dsp = np.concatenate(([0], np.cumsum(k_cnt)[:-1])).astype(np.int64)
    off = k_dsp[be] * K0 * dof + bt * dof * (nds[be] * K0)
    off = np.where(mf[be], -1, off).astype(np.int64)
    return off, int((k_dsp[-1] + k_cnt[-1]) * K0 * dof)


class NearField:
    """The device near-field operator of one BoundaryIntegralOp, built from the
    reference's own arrays (boundary_integral.txx):

      near_elem_cnt/dsp     near targets of each element (Xtrg_near rows)
      elem_nds_cnt/dsp_far  each element's far nodes (X_far rows); with the
                            identity FarFieldDensityOperatorTranspose they are
                            the element's nodes, so F is indexed the same way
      matrix_free           per element: no K_near matrix (MatrixFree())
      trg_normal_dot        contract the output with the target normal
      ctx                   the wavenumber of the Helmholtz kernels

    Methods:
      k_near_subtract(K_near)           KNearSubtract (:1165-1240)
      near_interac(U, F, K_near, ...)   ComputeNearInterac's matrix route
                                        (:1311-1381): GEMV, ScatterForward,
                                        per-target sum
    K_near uses the reference's layout (K_near_dsp[e]*KDIM0*KDIM1_ +
    t0*KDIM1_*src_dof per block, :1187-1189)."""

    def __init__(self, kernel, near_elem_cnt, near_elem_dsp, elem_nds_cnt_far, elem_nds_dsp_far, Xtrg_near,
                 X_far, Xn_far=None, wts_far=None, Xn_trg_near=None, matrix_free=None, trg_normal_dot=False,
                 nthreads=None, blocks=None, ctx=None):
        _, K0, K1, ND, _, _, _ = kernel_info(kernel)
        self.kernel, self.K0, self.K1 = kernel, K0, K1
        self.dof = K1 // 3 if trg_normal_dot else K1
        cnt = np.ascontiguousarray(near_elem_cnt, dtype=np.int64)
        dsp = np.ascontiguousarray(near_elem_dsp, dtype=np.int64)
        ncf = np.ascontiguousarray(elem_nds_cnt_far, dtype=np.int64)
        ndf = np.ascontiguousarray(elem_nds_dsp_far, dtype=np.int64)
        nelem = cnt.size
        mf = np.zeros(nelem, dtype=bool) if matrix_free is None else np.asarray(matrix_free, dtype=bool)
        if blocks is None:
            blocks = near_blocks(cnt, dsp, nthreads or os.cpu_count() or 1)
        be, bt, bc = (np.ascontiguousarray(a, dtype=np.int64) for a in blocks)
        self.blocks = (be, bt, bc)
        off, self.k_len = near_offsets(K0, self.dof, (be, bt, bc), cnt, ncf, mf)
        nt_total = int(dsp[-1] + cnt[-1]) if nelem else 0
        ns_total = int(np.max(ndf + ncf)) if nelem else 0
        self.nt_total, self.ns_total = nt_total, ns_total
        self.plan = Plan(kernel, dsp[be] + bt, bc, ndf[be], ncf[be], nt_total, ns_total, ctx=ctx)
        self.plan.set_geometry(Xtrg_near, X_far, Xn_far if ND else None)
        need = self.plan.near_layout(off, trg_normal_dot)
        assert need <= self.k_len
        self.wts_far = None if wts_far is None else _host(wts_far)
        self.Xn_trg_near = None if Xn_trg_near is None else _host(Xn_trg_near)

    def k_near_subtract(self, K_near, stream=None):
        """K_near -= direct part of every block (in place; numpy or DeviceArray)."""
        wts, ntr = self.wts_far, self.Xn_trg_near
        if isinstance(K_near, DeviceArray):
            wts = None if wts is None else DeviceArray.from_host(wts)
            ntr = None if ntr is None else DeviceArray.from_host(ntr)
            self._keep = (wts, ntr)
        return self.plan.near_matrix(K_near, wts, ntr, NEAR_SUBTRACT, stream)

    def near_interac(self, U, F, K_near, near_trg_cnt, near_trg_dsp, scatter_index):
        """U += near interactions (host arrays): the GEMV route of
        ComputeNearInterac followed by ScatterForward and the per-target sum.
        U is resized and zeroed if its length is not Ntrg*KDIM1_ (:1316-1319)."""
        ntrg = np.asarray(near_trg_cnt).size
        if U is None or U.size != ntrg * self.dof:
            U = np.zeros(ntrg * self.dof)
        U_near = np.zeros(self.nt_total * self.dof)
        self.plan.near_gemv(K_near, F, U_near)
        return near_accumulate(U, self.dof, near_trg_cnt, near_trg_dsp, scatter_index, U_near)


def near_list(Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg=None, stream=None):
    """BuildNearList (boundary_integral.txx:50-520) on the device, one process.

    Host arrays in, host arrays out: dict of Xtrg_near, Xn_trg_near (None
    without Xn_trg), near_elem_cnt, near_elem_dsp, near_scatter_index,
    near_trg_cnt, near_trg_dsp.  With DeviceArray inputs (all of them) the
    build runs on device memory and the results are still returned on the host."""
    dev = isinstance(Xtrg, DeviceArray)
    if dev:
        xt, xs, rad, ec, ed, xn = Xtrg, Xsrc, src_radius, elem_nds_cnt, elem_nds_dsp, Xn_trg
        nt, ns, ne = xt.shape[0] // 3, xs.shape[0] // 3, ec.shape[0]
        flags = 0
    else:
        xt, xs, rad, xn = _host(Xtrg), _host(Xsrc), _host(src_radius), _host(Xn_trg)
        ec = np.ascontiguousarray(elem_nds_cnt, dtype=np.int64)
        ed = np.ascontiguousarray(elem_nds_dsp, dtype=np.int64)
        nt, ns, ne = xt.size // 3, xs.size // 3, ec.size
        flags = HOST_PTRS
    h = _vp()
    check(lib().sctl_amd_near_list_build(ctypes.byref(h), _ptr(xt), nt, _ptr(xn), _ptr(xs), ns, _ptr(rad), ne,
                                         _ptr(ec), _ptr(ed), flags, stream))
    try:
        n = ctypes.c_int64()
        check(lib().sctl_amd_near_list_size(h, ctypes.byref(n)))
        n = n.value
        out = dict(Xtrg_near=np.zeros(3 * n), Xn_trg_near=np.zeros(3 * n) if Xn_trg is not None else None,
                   near_elem_cnt=np.zeros(ne, dtype=np.int64), near_elem_dsp=np.zeros(ne, dtype=np.int64),
                   near_scatter_index=np.zeros(n, dtype=np.int64), near_trg_cnt=np.zeros(nt, dtype=np.int64),
                   near_trg_dsp=np.zeros(nt, dtype=np.int64))
        keys = ("Xtrg_near", "Xn_trg_near", "near_elem_cnt", "near_elem_dsp", "near_scatter_index", "near_trg_cnt",
                "near_trg_dsp")
        check(lib().sctl_amd_near_list_get(h, *[_ptr(out[k]) for k in keys], HOST_PTRS, stream))
    finally:
        lib().sctl_amd_near_list_destroy(h)
    return out

