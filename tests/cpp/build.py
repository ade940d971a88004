"""Build recipe of the C++ test programs under tests/cpp/ (host code, g++).

  test_generic_kernel    the drop-in on the library's own containers
  test_concurrent_calls  the drop-in called from 8 OpenMP threads at once,
                         checked against the oracle (links oracle/liboracle.so)
  test_near_field_adapter  the near-field operator set up from element lists
  test_dropin_sctl       the drop-in with SCTL's own types under SCTL_MEMDEBUG
                         and SCTL_PROFILE; built only where the reference
                         headers are present (nothing from them is needed at
                         run time)

Each program is linked to sctl_amd/libsctl_amd.so with an rpath relative to
its own directory, so the tree can be moved as a whole.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "sctl_amd")
ORACLE = os.path.join(ROOT, "oracle")
REF_INC = "/root/reference/include"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C++ test build failed: %s\n%s" % (" ".join(cmd), r.stderr))


def _base(name, extra_src=()):
    src = os.path.join(HERE, name + ".cpp")
    out = os.path.join(HERE, name)
    return ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), src] + list(extra_src), out


def build(verbose=True):
    built = []
    cmd, out = _base("test_generic_kernel")
    _run(cmd + ["-L" + PKG, "-lsctl_amd", "-Wl,-rpath,$ORIGIN/../../sctl_amd", "-o", out])
    built.append(out)

    cmd, out = _base("test_concurrent_calls")
    _run(cmd + ["-fopenmp", "-L" + PKG, "-lsctl_amd", "-L" + ORACLE, "-loracle",
                "-Wl,-rpath,$ORIGIN/../../sctl_amd", "-Wl,-rpath,$ORIGIN/../../oracle", "-o", out])
    built.append(out)

    cmd, out = _base("test_near_field_adapter")
    _run(cmd + ["-L" + PKG, "-lsctl_amd", "-Wl,-rpath,$ORIGIN/../../sctl_amd", "-o", out])
    built.append(out)

    if os.path.isdir(REF_INC):
        cmd, out = _base("test_dropin_sctl")
        _run(cmd + ["-I" + REF_INC, "-fopenmp", "-Wno-unknown-pragmas", "-DSCTL_MEMDEBUG", "-DSCTL_PROFILE=5",
                    "-DSCTL_GLOBAL_MEM_BUFF=0", "-L" + PKG, "-lsctl_amd", "-Wl,-rpath,$ORIGIN/../../sctl_amd",
                    "-o", out])
        built.append(out)
    if verbose:
        for b in built:
            print("[sctl_amd] built", b, flush=True)
    return built


if __name__ == "__main__":
    build()
    sys.exit(0)
