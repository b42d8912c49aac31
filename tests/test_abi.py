"""The C-ABI library loads and exports every entry point include/harmony_ivf.h declares; host
error paths that need no GPU (here: creating an index with no device visible)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "harmony_ivf.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hv_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("hv_create", "hv_train", "hv_add", "hv_search", "hv_search_device", "hv_set_search_params",
              "hv_last_error", "hv_ground_truth_f64"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from harmony_amd import _lib
    L = _lib.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers the whole header too
    assert not [s for s in declared_symbols() if s not in _lib.SIGNATURES]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from harmony_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.HipExtensionMissing):
        _lib.lib()


def test_create_without_device_reports_error():
    from harmony_amd import _lib
    L = _lib.lib()
    if L.hv_device_count() > 0:
        pytest.skip("a GPU is visible")
    h = L.hv_create(128, 100, _lib.HV_METRIC_L2, 0)
    assert not h
    assert b"no GPU" in L.hv_last_error(None)


def test_create_rejects_bad_arguments():
    from harmony_amd import _lib
    L = _lib.lib()
    assert not L.hv_create(0, 100, _lib.HV_METRIC_L2, 0)
    assert not L.hv_create(128, 100, 5, 0)
    assert b"metric" in L.hv_last_error(None)
    assert not L.hv_create(128, 100, _lib.HV_METRIC_L2, 9)


def test_null_handle_calls_return_invalid_arg():
    from harmony_amd import _lib
    L = _lib.lib()
    assert L.hv_add(None, 1, None) == _lib.HV_ERR_INVALID_ARG
    assert L.hv_search(None, 1, None, 1, None, None, None) == _lib.HV_ERR_INVALID_ARG
    assert L.hv_ntotal(None) == -1


def test_last_getters_null_and_no_search():
    from harmony_amd import _lib
    L = _lib.lib()
    assert L.hv_get_last_probes(None, None, 0) == _lib.HV_ERR_INVALID_ARG
    assert L.hv_get_last_bounds(None, None, 0) == _lib.HV_ERR_INVALID_ARG
    assert L.hv_get_last_nprobe(None, None, 0) == _lib.HV_ERR_INVALID_ARG
    assert L.hv_last_n(None) == 0
