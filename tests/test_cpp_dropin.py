"""The C++ drop-in header (include/sctl_amd/generic_kernel.hpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_INC = "/root/reference/include"


def _syntax(src, extra=()):
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I" + os.path.join(ROOT, "include")] + list(extra) + [src]
    return subprocess.run(cmd, capture_output=True, text=True)


def test_dropin_header_compiles():
    r = _syntax(os.path.join(ROOT, "tests", "cpp", "test_generic_kernel.cpp"))
    assert r.returncode == 0, r.stderr


def test_near_field_header_compiles():
    r = _syntax(os.path.join(ROOT, "tests", "cpp", "test_near_field_adapter.cpp"))
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
@pytest.mark.parametrize("memdebug", [False, True], ids=["default", "SCTL_MEMDEBUG"])
def test_dropin_accepts_sctl_types(memdebug):
    """sctl::Vector / sctl::Matrix arguments, and ParticleFMM's own assignment
    `data.ker_s2t_eval = KerS2T::template Eval<Real,false>;` (fmm-wrapper.txx:450-451)
    in the default build and under SCTL_MEMDEBUG, where ConstIterator<char> is a class."""
    extra = ["-I" + REF_INC, "-DSCTL_GLOBAL_MEM_BUFF=0", "-Wno-unknown-pragmas"]
    if memdebug:
        extra.append("-DSCTL_MEMDEBUG")
    r = _syntax(os.path.join(ROOT, "tests", "cpp", "dropin_sctl_types.cpp"), extra)
    assert r.returncode == 0, r.stderr


def test_dropin_rejects_other_kernels():
    src = os.path.join(ROOT, "_cpp_reject_probe.cpp")
    code = ('#include "sctl_amd/generic_kernel.hpp"\nstruct Other { static constexpr int KERNEL_ID = 0; '
            'static constexpr long DIM = 3, KDIM0 = 1, KDIM1 = 1, N_DIM = 0; };\n'
            'sctl_amd::GenericKernel<Other> k;\n')
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "reject.cpp")
        open(src, "w").write(code)
        r = _syntax(src)
    assert r.returncode != 0 and "only the sctl_amd::kernel_impl micro-kernels run on the device" in r.stderr


def _run_program(name):
    exe = os.path.join(ROOT, "tests", "cpp", name)
    assert os.path.exists(exe), "tests/cpp/%s not built (run __graft_entry__.build())" % name
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASS (0 check(s) failed)" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_dropin_on_gpu(gpu):
    _run_program("test_generic_kernel")


@pytest.mark.gpu
def test_cpp_dropin_concurrent_openmp_calls(gpu):
    """KernelMatrix<double,false> / Eval<double,false> from 8 OpenMP threads at once, as
    BoundaryIntegralOp calls them (boundary_integral.txx:1165-1211): bit-identical to one
    thread, and the oracle at 1e-13."""
    _run_program("test_concurrent_calls")


@pytest.mark.gpu
def test_cpp_near_field_from_element_lists(gpu):
    """sctl_amd::NearFieldOp set up from two element lists (one matrix-free) through the
    ElementListBase interface: near list, KNearSubtract + NearInterac and the direct part
    against long-double sums over the lists' own nodes."""
    _run_program("test_near_field_adapter")


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "tests", "cpp", "test_dropin_sctl")),
                    reason="built only where the reference headers are present")
def test_cpp_dropin_sctl_memdebug_and_profile(gpu):
    """ParticleFMM's binding with a checked ConstIterator<char> self (SCTL_MEMDEBUG), and the
    reference profiler's FLOP counter moving by Ns*Nt*FLOPS() per call."""
    _run_program("test_dropin_sctl")
