"""Error paths of the C-ABI that need a built index (hence a GPU): the last-search getters
refuse buffers shorter than the last search, and bound modes that a form cannot honour are
refused rather than ignored."""
import ctypes

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, datagen
from harmony_amd._lib import HV_BOUND_EXTERNAL, HV_BOUND_WARMUP, HV_ERR_INVALID_ARG, HV_OK, HvError, lib
from harmony_amd.index import Index

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w1():
    xb, xq = datagen.make(1)
    g = Index(128, 100, nprobe=10)
    g.train(xb, niter=10, seed=0)
    g.add(xb)
    return g, xq


def test_last_getters_check_capacity(w1):
    g, xq = w1
    g.preSearch(pruning=True, dim_blocks=4, precision=0, learned=False)
    g.search(xq[:37], 10)
    assert g.last_n() == 37
    stride = int(lib().hv_last_probe_stride(g._h))
    assert stride == 10
    L = lib()
    short = np.empty(37 * stride - 1, np.int64)
    assert L.hv_get_last_probes(g._h, short.ctypes.data_as(ctypes.c_void_p), short.size) == HV_ERR_INVALID_ARG
    full = np.empty(37 * stride, np.int64)
    assert L.hv_get_last_probes(g._h, full.ctypes.data_as(ctypes.c_void_p), full.size) == HV_OK
    b = np.empty(36, np.float32)
    assert L.hv_get_last_bounds(g._h, b.ctypes.data_as(ctypes.c_void_p), b.size) == HV_ERR_INVALID_ARG
    c = np.empty(36, np.int32)
    assert L.hv_get_last_nprobe(g._h, c.ctypes.data_as(ctypes.c_void_p), c.size) == HV_ERR_INVALID_ARG
    # the Python wrappers size from the library's count
    assert g.last_probes().shape == (37, 10)
    assert g.last_nprobe().shape == (37,)
    with pytest.raises(HvError):
        g.last_probes(36)


def test_external_bound_refused_with_fp8_and_learned(w1):
    g, xq = w1
    g.set_external_bounds(np.full(8, 1e30, np.float32))
    g.preSearch(bound_mode=HV_BOUND_EXTERNAL, precision=HV_PRECISION_FP8, learned=False)
    with pytest.raises(HvError, match="HV_BOUND_EXTERNAL"):
        g.search(xq[:8], 10)
    g.preSearch(bound_mode=HV_BOUND_EXTERNAL, precision=HV_PRECISION_BF16, learned=False)
    D, I = g.search(xq[:8], 10)  # the BF16 form honours it
    assert (I >= 0).all()
    g.preSearch(bound_mode=HV_BOUND_WARMUP, precision=0, learned=False)
