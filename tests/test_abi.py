"""The C-ABI library: it loads, exports every entry point include/sctl_amd.h
declares, and rejects bad arguments with status codes (no GPU needed: none of
these calls reaches a kernel)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sctl_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sctl_amd.h")).read()
    return sorted(set(re.findall(r"\b(sctl_amd_\w+)\s*\(", text)))


def test_library_built_in_tree():
    assert os.path.exists(S.LIB_PATH), "libsctl_amd.so not built (run __graft_entry__.build())"


def test_exports_every_declared_symbol():
    L = S.lib()
    decl = declared_symbols()
    assert len(decl) >= 17
    missing = [s for s in decl if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(S.SYMBOLS) == decl


def test_library_does_not_link_the_oracle():
    """The product never depends on the oracle or the compiled reference."""
    data = open(S.LIB_PATH, "rb").read()
    for needle in (b"liboracle", b"oracle_eval", b"sctlref_"):
        assert needle not in data


def test_kernel_info_matches_reference_metadata():
    # generic-kernel.txx:48-62 dims; kernel_functions.hpp FLOPS() and Name()
    want = {0: (3, 1, 1, 0, 6, "Laplace3D-FxU"), 1: (3, 1, 1, 3, 14, "Laplace3D-DxU"),
            2: (3, 3, 3, 0, 23, "Stokes3D-FxU"), 3: (3, 3, 3, 3, 26, "Stokes3D-DxU"),
            4: (3, 1, 3, 0, 11, "Laplace3D-FxdU"), 5: (3, 1, 3, 3, 22, "Laplace3D-DxdU"),
            6: (3, 1, 9, 0, 25, "Laplace3D-Fxd2U"), 7: (3, 3, 9, 0, 39, "Stokes3D-FxT"),
            8: (3, 4, 3, 0, 26, "Stokes3D-FSxU"), 9: (3, 3, 4, 0, 26, "Stokes3D-FxUP"),
            10: (3, 3, 3, 0, 15, "BiotSavart3D-FxU"), 11: (3, 3, 9, 0, 51, "BiotSavart3D-FxdU"),
            12: (3, 2, 2, 0, 20, "Helmholtz3D-FxU"), 13: (3, 2, 2, 3, 29, "Helmholtz3D-DxU"),
            14: (3, 2, 6, 0, 35, "Helmholtz3D-FxdU"), 15: (3, 2, 6, 0, 41, "HelmholtzDiff3D-FxdU")}
    import oracle
    assert len(want) == S.NUM_KERNELS
    for k, w in want.items():
        c, k0, k1, nd, fl, sc, name = S.kernel_info(k)
        assert (c, k0, k1, nd, fl, name) == w
        assert sc == oracle.scale(k)  # same rounding as uKerScaleFactor<double>()


def _eval_rc(*a):
    return S.lib().sctl_amd_eval(*a)


def test_error_codes_before_any_device_work():
    L = S.lib()
    x = np.zeros(3)
    v = np.zeros(3)
    p = x.ctypes.data
    assert _eval_rc(16, 0, -1, p, 1, p, None, p, 1, p, 1, None) == 2  # unknown kernel
    assert b"unknown kernel" in L.sctl_amd_last_error()
    assert _eval_rc(0, 7, -1, p, 1, p, None, p, 1, p, 1, None) == 1  # unknown precision
    assert _eval_rc(0, 0, -2, p, 1, p, None, p, 1, p, 1, None) == 1  # digits < -1
    assert _eval_rc(0, 0, -1, None, 1, p, None, p, 1, p, 1, None) == 1  # NULL r_trg
    assert _eval_rc(1, 0, -1, p, 1, p, None, p, 1, p, 1, None) == 1  # DxU without normals
    assert b"n_src" in L.sctl_amd_last_error()
    assert _eval_rc(0, 0, -1, p, -1, p, None, p, 1, p, 1, None) == 1  # negative size
    assert _eval_rc(0, 0, -1, p, 1, p, None, p, 1, p, 0x80, None) == 1  # unknown flag
    assert L.sctl_amd_kernel_matrix(16, -1, p, 1, p, None, 1, p, 1, None) == 2
    assert _eval_rc(12, 0, -1, p, 1, p, None, p, 1, p, 1, None) == 1  # Helmholtz without its wavenumber
    assert b"wavenumber" in L.sctl_amd_last_error()
    inf = ctypes.pointer(ctypes.c_double(float("inf")))
    assert L.sctl_amd_eval_ctx(12, inf, 0, -1, p, 1, p, None, p, 1, p, 1, None) == 1  # non-finite wavenumber
    assert L.sctl_amd_memcpy(p, p, 8, 3, None) == 1
    with pytest.raises(S.SctlAmdError):
        S.kernel_info(16)
    with pytest.raises(S.SctlAmdError):
        S.kernel_info(-1)


def test_plan_rejects_bad_blocks():
    L = S.lib()
    h = ctypes.c_void_p()
    a = lambda *v: np.array(v, dtype=np.int64)
    P = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    td, tc, sd, sc = a(0, 5), a(10, 10), a(0, 0), a(4, 4)  # targets [0,10) and [5,15) overlap
    assert L.sctl_amd_plan_create(ctypes.byref(h), 0, 0, -1, 2, P(td), P(tc), P(sd), P(sc), 20, 4) == 1
    assert b"overlap" in L.sctl_amd_last_error()
    td, tc = a(0, 10), a(10, 20)  # second block runs past nt_total
    assert L.sctl_amd_plan_create(ctypes.byref(h), 0, 0, -1, 2, P(td), P(tc), P(sd), P(sc), 20, 4) == 1
    assert L.sctl_amd_plan_create(ctypes.byref(h), 16, 0, -1, 2, P(td), P(tc), P(sd), P(sc), 40, 4) == 2
    assert L.sctl_amd_plan_apply(None, None, None, 0, None) == 1


def test_no_device_is_reported(gpu_present):
    if gpu_present:
        pytest.skip("a GPU is present")
    x = np.zeros(3)
    p = x.ctypes.data
    assert _eval_rc(0, 0, -1, p, 1, p, None, p, 1, p, 1, None) == 4
    assert b"no HIP device" in S.lib().sctl_amd_last_error()
