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
KERNEL_NAMES = {0: "Laplace3D_FxU", 1: "Laplace3D_DxU", 2: "Stokes3D_FxU", 3: "Stokes3D_DxU"}
FP64, MIXED, AUTO = 0, 1, 2
HOST_PTRS, MIXED_NEWTON = 0x1, 0x2

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
    L.sctl_amd_kernel_matrix.argtypes = [i32, i32, _vp, i64, _vp, _vp, i64, _vp, u32, _vp]
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
    ip = ctypes.POINTER(i32)
    L.sctl_amd_plan_get_shape.argtypes = [_vp, ip, ip, ip]
    L.sctl_amd_plan_set_shape.argtypes = [_vp, i32, i32, i32]
    L.sctl_amd_tune_db_path.argtypes = [ctypes.c_char_p]
    L.sctl_amd_tune_eval.argtypes = [i32, i32, i32, _vp, i64, _vp, _vp, _vp, i64, _vp, u32, _vp, ip, ip, _dp]
    L.sctl_amd_plan_tune.argtypes = [_vp, _vp, _vp, _vp, ip, ip, ip, _dp]
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


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    return a.ctypes.data if a.size else None


# ----------------------------------------------------------------- Eval API
def eval(kernel, r_trg, r_src, n_src, v_src, v_trg=None, precision=FP64, digits=-1, flags=0, stream=None):
    """v_trg += Eval(...).

    With numpy arrays the call stages through the device and returns the
    (new, or updated in place) host v_trg.  With DeviceArray arguments the call
    is asynchronous on `stream` and returns v_trg."""
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
    check(lib().sctl_amd_eval(kernel, precision, digits, _ptr(r_trg), nt, _ptr(r_src), _ptr(n_src), _ptr(v_src), ns,
                              _ptr(v_trg), flags, stream))
    return v_trg


def kernel_matrix(kernel, r_trg, r_src, n_src, digits=-1, M=None, flags=0, stream=None):
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
    check(lib().sctl_amd_kernel_matrix(kernel, digits, _ptr(r_trg), nt, _ptr(r_src), _ptr(n_src), ns, _ptr(M),
                                       flags, stream))
    return M


class GenericKernel:
    """Python mirror of sctl::GenericKernel<uKernel> for the four kernels."""

    def __init__(self, kernel):
        (self._coord, self._k0, self._k1, self._nd, self._flops, self._scale,
         self._name) = kernel_info(kernel)
        self.kernel = kernel

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
        return eval(self.kernel, r_trg, r_src, n_src, v_src, v, precision=precision, digits=digits)

    def KernelMatrix(self, r_trg, r_src, n_src, digits=-1):
        return kernel_matrix(self.kernel, r_trg, r_src, n_src, digits=digits)


# ------------------------------------------------------------------ batched
class Plan:
    """Batched (target-set, source-set) blocks evaluated in one persistent launch."""

    def __init__(self, kernel, trg_dsp, trg_cnt, src_dsp, src_cnt, nt_total, ns_total, precision=FP64, digits=-1):
        a = [np.ascontiguousarray(x, dtype=np.int64) for x in (trg_dsp, trg_cnt, src_dsp, src_cnt)]
        self._keep = a
        self.kernel = kernel
        self.nblocks = a[0].size
        h = _vp()
        P = lambda x: x.ctypes.data_as(_i64p)
        check(lib().sctl_amd_plan_create(ctypes.byref(h), kernel, precision, digits, self.nblocks, P(a[0]), P(a[1]),
                                         P(a[2]), P(a[3]), int(nt_total), int(ns_total)))
        self.handle = h

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

    def destroy(self):
        if self.handle:
            lib().sctl_amd_plan_destroy(self.handle)
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
        check(lib().sctl_amd_plan_graph_create(ctypes.byref(h), plan.handle, nsteps, seed, stream_id, v_src.ptr,
                                               v_trg.ptr))
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
    """Autotunes the dense Eval shape for these device arrays; returns (tpl, cap, ms)."""
    nt, ns = r_trg.nbytes // 24, r_src.nbytes // 24
    a, b = ctypes.c_int(), ctypes.c_int()
    ms = ctypes.c_double()
    check(lib().sctl_amd_tune_eval(kernel, precision, digits, r_trg.ptr, nt, r_src.ptr, _ptr(n_src), v_src.ptr, ns,
                                   v_scratch.ptr, flags, stream, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ms)))
    return a.value, b.value, ms.value


def gen_normal(seed, stream_id, out, stream=None):
    """Device generator: fills DeviceArray `out` with N(0,1) draws of (seed, stream_id)."""
    check(lib().sctl_amd_gen_normal(seed, stream_id, out.nbytes // 8, out.ptr, stream))
    return out
