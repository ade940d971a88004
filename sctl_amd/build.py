"""Build recipe for libsctl_amd.so (gfx950).

Each source under sctl_amd/csrc/ is compiled to an object with hipcc and the
objects are linked into sctl_amd/libsctl_amd.so, in tree.  Objects are rebuilt
when their source or a header is newer; ``force=True`` rebuilds everything.

Every HIP compile carries ``-Xclang -target-feature -Xclang -packed-fp32-ops``:
packed FP32 VALU instructions are kept out of the generated code.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "libsctl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["dense.hip", "mixed.hip", "sort.hip", "batched.hip", "near.hip", "nearlist.hip", "graph.hip", "noalloc.hip",
           "capi.cpp", "tune.cpp", "profile.cpp", "dense_plan.cpp"]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function",
]


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    hs.append(os.path.join(ROOT, "include", "sctl_amd.h"))
    return hs


def _needs(obj, src, hdr_mtime, force):
    if force or not os.path.exists(obj):
        return True
    return os.path.getmtime(obj) < max(os.path.getmtime(src), hdr_mtime)


def _compile(src, obj):
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(force=False, jobs=None, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    hdr = max(os.path.getmtime(h) for h in _headers())
    work = []
    objs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, os.path.splitext(s)[0] + ".o")
        objs.append(obj)
        if _needs(obj, src, hdr, force):
            work.append((src, obj))
    if work:
        if verbose:
            print("[sctl_amd] compiling", ", ".join(os.path.basename(s) for s, _ in work), flush=True)
        with ThreadPoolExecutor(max_workers=jobs or min(len(work), 8)) as ex:
            list(ex.map(lambda a: _compile(*a), work))
    if work or not os.path.exists(LIB) or force:
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv)
    print(LIB)
