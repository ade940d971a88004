"""sctl_amd_eval_noalloc: the Eval of sctl_amd_eval_ctx with its scratch in
the library's fixed device array, for callers that cannot allocate.  Its
output is checked against sctl_amd_eval_ctx's on the same device inputs, on
calls in a row (each later call smaller than one before it, so stale scratch
contents would show), including sizes whose split partial sums exceed the
array and run in groups of launches, and one split per launch."""
import ctypes

import numpy as np
import pytest

from sctl_amd import workloads as W

pytestmark = pytest.mark.gpu

_HIP = None


def _noalloc(L):
    f = L.sctl_amd_eval_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                  ctypes.c_uint32, ctypes.c_void_p]
    return f


def _pool_used(attr):
    """Default memory pool of device 0: used bytes (attr 7: current, 8: high-water mark)."""
    global _HIP
    if _HIP is None:
        _HIP = ctypes.CDLL("libamdhip64.so")
    pool = ctypes.c_void_p()
    assert _HIP.hipDeviceGetDefaultMemPool(ctypes.byref(pool), 0) == 0
    v = ctypes.c_uint64()
    assert _HIP.hipMemPoolGetAttribute(pool, attr, ctypes.byref(v)) == 0
    return pool, v.value


def _reset_high(pool):
    zero = ctypes.c_uint64(0)
    assert _HIP.hipMemPoolSetAttribute(pool, 8, ctypes.byref(zero)) == 0


def _run(gpu, kernel, precision, xt, xs, n, f, ctx=None, expect_groups=None):
    """Both entry points on the same device inputs; returns (noalloc, eval)."""
    _, K0, K1, ND, _, _, _ = gpu.kernel_info(kernel)
    nt, ns = xt.size // 3, xs.size // 3
    print("noalloc case: kernel %d precision %d nt %d ns %d" % (kernel, precision, nt, ns), flush=True)
    d = [gpu.DeviceArray.from_host(a) for a in (xt, xs, f)]
    dn = gpu.DeviceArray.from_host(n) if ND else None
    init = np.linspace(-1, 1, nt * K1)  # both accumulate into the same start values
    out_a = gpu.DeviceArray.from_host(init)
    out_b = gpu.DeviceArray.from_host(init)
    c = None if ctx is None else ctypes.pointer(ctypes.c_double(ctx))
    gpu.synchronize()
    pool, used = _pool_used(7)
    _reset_high(pool)
    rc = _noalloc(gpu.lib())(kernel, c, precision, -1, d[0].ptr, nt, d[1].ptr, dn.ptr if dn else None, d[2].ptr, ns,
                             out_a.ptr, 0, None)
    gpu.check(rc)
    gpu.synchronize()
    _, high = _pool_used(8)
    assert high <= used, "the no-alloc entry point took %d bytes from the memory pool" % (high - used)
    gpu.check(gpu.lib().sctl_amd_eval_ctx(kernel, c, precision, -1, d[0].ptr, nt, d[1].ptr, dn.ptr if dn else None,
                                          d[2].ptr, ns, out_b.ptr, 0, None))
    gpu.synchronize()
    return out_a.download() - init, out_b.download() - init


def test_noalloc_matches_eval_on_calls_in_a_row(gpu):
    # 1. mixed Laplace FxU, 2^20 sphere sources x 2^20 shell targets: 16 splits
    #    of 8 MiB partials do not fit beside the orders and tiles, so they run
    #    in groups of launches.
    xt, xs, n, f = W.sphere(1 << 20, 1 << 20, seed=0)
    a, b = _run(gpu, 0, gpu.MIXED, xt, xs, n, f)
    assert W.rel_errors(a, b)[0] < 1e-13
    ref64 = gpu.eval(0, gpu.DeviceArray.from_host(xt), gpu.DeviceArray.from_host(xs), None,
                     gpu.DeviceArray.from_host(f), gpu.DeviceArray.from_host(np.zeros(xt.size // 3)),
                     precision=gpu.FP64).download()
    assert W.rel_errors(a, ref64)[0] < 1e-6
    # 2. a smaller mixed Stokes FxU call (one launch): bit-identical
    xt, xs, f = W.cube(5000, 3000, seed=1, k0=3)
    a, b = _run(gpu, 2, gpu.MIXED, xt, xs, None, f)
    assert np.array_equal(a, b)
    # 3. FP64 Stokes FxU, 2^20 targets x 2^19 sources: 8 splits of 24 MiB
    #    partials, in groups of 5 within the array
    xt, xs, f = W.cube(1 << 19, 1 << 20, seed=2, k0=3)
    a, b = _run(gpu, 2, gpu.FP64, xt, xs, None, f)
    assert W.rel_errors(a, b)[0] < 1e-13
    # 4. FP64 Stokes FxT (9 outputs), 2^21 targets x 2^17 sources: one split's
    #    partials (144 MiB) exceed the array, so each split accumulates
    #    straight into v_trg
    xt, xs, f = W.cube(1 << 17, 1 << 21, seed=3, k0=3)
    a, b = _run(gpu, 7, gpu.FP64, xt, xs, None, f)
    assert W.rel_errors(a, b)[0] < 1e-13
    # 5. smaller again: mixed Laplace DxU and FP64 Helmholtz, one launch each
    xt, xs, f = W.cube(2000, 7000, seed=4)
    n = W.normal(4, 9, 3 * 2000)
    a, b = _run(gpu, 1, gpu.MIXED, xt, xs, n, f)
    assert np.array_equal(a, b)
    xt, xs, f = W.cube(900, 1100, seed=5, k0=2)
    a, b = _run(gpu, 12, gpu.FP64, xt, xs, None, f, ctx=2.0)
    assert np.array_equal(a, b)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    L = gpu.lib()
    L.sctl_amd_eval_noalloc_scratch_bytes.restype = ctypes.c_size_t
    cap = L.sctl_amd_eval_noalloc_scratch_bytes()
    assert cap == 128 << 20
    nt = ns = 1 << 22  # orders, tiles and sort workspace alone exceed 128 MiB
    xt, xs, f = gpu.DeviceArray(3 * nt), gpu.DeviceArray(3 * ns), gpu.DeviceArray(ns)
    out = gpu.DeviceArray(nt)
    rc = _noalloc(L)(0, None, gpu.MIXED, -1, xt.ptr, nt, xs.ptr, None, f.ptr, ns, out.ptr, 0, None)
    assert rc == 1 and b"more than its" in L.sctl_amd_last_error()  # SCTL_AMD_ERROR_INVALID_VALUE
    # host pointers need staging buffers: refused
    h = np.zeros(30)
    rc = _noalloc(L)(0, None, gpu.FP64, -1, h.ctypes.data, 10, h.ctypes.data, None, h.ctypes.data, 10,
                     h.ctypes.data, gpu.HOST_PTRS, None)
    assert rc == 1
    # empty sizes: nothing to do
    rc = _noalloc(L)(0, None, gpu.FP64, -1, None, 0, xs.ptr, None, f.ptr, 5, None, 0, None)
    assert rc == 0
