"""Build libpyqed_amd.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

    python -m pyqed_amd.build

The library is written next to this file; tests, smoke() and bench.py load it
from there.  Packed-FP32 VALU instructions are disabled for every device
compilation (-target-feature -packed-fp32-ops).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libpyqed_amd.so")
SOURCES = ["capi.cpp", "heom.hip", "liouville.hip", "lindblad.hip", "legacy.hip", "spectral.hip"]
ARCH = os.environ.get("PYQED_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build pyqed_amd)")


def build(verbose=True, out=None, defines=()):
    cc = hipcc()
    flags = [*(f"-D{d}" for d in defines), "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INCLUDE, "-I", CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(bdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc, *flags, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    target = out or LIB
    tmp = target + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, target)
    return target


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="output path (default: pyqed_amd/libpyqed_amd.so)")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="preprocessor define NAME=VALUE")
    a = ap.parse_args()
    build(out=a.out, defines=a.defines)
    sys.exit(0)
