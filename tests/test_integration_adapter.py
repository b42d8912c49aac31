"""The reference-side C++ adapter that INTEGRATION.md shows (harmony::GpuIndex over the C-ABI)
compiles against include/harmony_ivf.h, links against libharmony_amd.so and runs: constructing
it with no GPU visible raises the library's error through std::runtime_error, as the reference's
own Index raises its errors (Index.cpp:1642-1645)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MAIN = r"""
#include <cstdio>
#include <vector>
int main() {
    try {
        harmony::GpuIndex idx(16, 4, 2);
        std::vector<float> x(64 * 16);
        for (size_t i = 0; i < x.size(); ++i) x[i] = (float)(i % 97);
        idx.train(64, x.data());
        idx.add(64, x.data());
        idx.preSearch(10, 0, true, 4);
        std::vector<float> D(2 * 3);
        std::vector<int64_t> I(2 * 3);
        idx.search(2, x.data(), 3, D.data(), I.data());
        std::printf("searched %lld\n", (long long)I[0]);
    } catch (const std::runtime_error& e) {
        std::printf("error: %s\n", e.what());
    }
    return 0;
}
"""


def adapter_source():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", txt, flags=re.S)
    assert len(blocks) == 1, "INTEGRATION.md should hold exactly one C++ adapter block"
    return blocks[0]


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_adapter_compiles_links_and_runs(tmp_path):
    from harmony_amd import _lib
    _lib.lib()  # built
    src = tmp_path / "adapter_main.cpp"
    src.write_text(adapter_source() + MAIN)
    exe = tmp_path / "adapter_main"
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
           "-lharmony_amd", "-Wl,-rpath," + libdir, "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout
    if _lib.lib().hv_device_count() > 0:
        assert out.startswith("searched "), out
    else:
        assert "no GPU device visible" in out, out
