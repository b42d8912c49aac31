"""Build libharmony_amd.so (host C++ + HIP kernels for gfx950) in-tree with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libharmony_amd.so")
ARCH = os.environ.get("HV_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hv_build.hip", "hv_scan.hip", "hv_gt.hip", "hv_index.cpp"]

# -packed-fp32-ops off: hipcc -O3 would otherwise pair FP32 ops into v_pk_* forms.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"]
        _run(["hipcc"] + lang + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
