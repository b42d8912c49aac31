"""Build recipe for the CPU oracle (test infrastructure only).

Outputs:
  oracle/build/liboracle.so          exact restatement (explicit fmaf chain, no fast-math):
                                     the parity checker.
  oracle/build/liboracle_fast_v4.so  same source with the reference's distance loop, built
  oracle/build/liboracle_fast_v3.so  with the reference Release flags (-Ofast -O3, AVX-512 /
                                     AVX2 targets): the timed CPU baseline. v4 is picked when
                                     the host has avx512f, as the reference's -mavx512f needs.
  oracle/_ref/libref_heap.so         the reference's own src/heap.hpp compiled from where it
                                     lies, behind oracle/ref_heap_driver.cpp. Only built when
                                     the reference checkout is present.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
REF_OUT = os.path.join(HERE, "_ref")
REFERENCE_SRC = "/root/reference/src"

SRC = os.path.join(HERE, "harmony_oracle.cpp")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_oracle():
    os.makedirs(BUILD, exist_ok=True)
    common = ["g++", "-std=c++17", "-shared", "-fPIC", "-fopenmp", SRC]
    # exact checker: portable AVX2+FMA target, contraction off so only the explicit fmaf
    # chain fuses
    _run(common + ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-o",
                   os.path.join(BUILD, "liboracle.so")])
    # timed baseline: the reference Release flags (CMakeLists.txt:37 + CMake Release)
    for tag, arch in (("v4", "x86-64-v4"), ("v3", "x86-64-v3")):
        _run(common + ["-DORACLE_REFERENCE_FLAGS", "-Ofast", "-O3", "-DNDEBUG", "-march=" + arch,
                       "-mtune=generic", "-o", os.path.join(BUILD, "liboracle_fast_%s.so" % tag)])


def build_ref():
    """Compile the reference heap from its own source. Returns False when absent."""
    if not os.path.isfile(os.path.join(REFERENCE_SRC, "heap.hpp")):
        return False
    os.makedirs(REF_OUT, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-shared", "-fPIC", "-fopenmp", "-I", REFERENCE_SRC,
          os.path.join(HERE, "ref_heap_driver.cpp"), "-o", os.path.join(REF_OUT, "libref_heap.so")])
    return True


def build_all():
    build_oracle()
    build_ref()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
