"""Build libharmony_amd.so (host C++ + HIP kernels for gfx950) in-tree with hipcc.

python -m harmony_amd.build [--incremental]
  --incremental  recompile only sources newer than their object (headers count for all)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libharmony_amd.so")
ARCH = os.environ.get("HV_OFFLOAD_ARCH", "gfx950")

SOURCES = ["hv_build.hip", "hv_scan.hip", "hv_select.hip", "hv_gt.hip", "hv_bf16.hip", "hv_learn.hip",
           "hv_coarse.hip", "hv_ip.hip", "hv_index.cpp"]
HEADERS = ["hv_device.h", "hv_kernels.h", os.path.join("..", "..", "include", "harmony_ivf.h")]

# -packed-fp32-ops off: hipcc -O3 would otherwise pair FP32 ops into v_pk_* forms.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def _stale(src, obj):
    if not os.path.isfile(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, h) for h in HEADERS]
    return any(os.path.getmtime(p) > t for p in deps if os.path.isfile(p))


def build(incremental=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if incremental and not _stale(path, obj):
            continue
        jobs.append(["hipcc", "-x", "hip"] + HIPFLAGS + ["-c", path, "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, min(len(jobs), 4))) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
