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


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_dropin_accepts_sctl_types():
    """sctl::Vector / sctl::Matrix arguments and the ParticleFMM function-pointer type."""
    r = _syntax(os.path.join(ROOT, "tests", "cpp", "dropin_sctl_types.cpp"), ["-I" + REF_INC, "-DSCTL_GLOBAL_MEM_BUFF=0"])
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


@pytest.mark.gpu
def test_cpp_dropin_on_gpu(gpu):
    exe = os.path.join(ROOT, "tests", "cpp", "test_generic_kernel")
    assert os.path.exists(exe), "tests/cpp/test_generic_kernel not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASS (0 check(s) failed)" in r.stdout, r.stdout + r.stderr
