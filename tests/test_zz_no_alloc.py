"""hv_search_device_noalloc: the search with every scratch buffer carved from the library's fixed
per-GPU block. Results must equal hv_search_device's for every form, on two calls in a row (the
second with fewer queries); a batch too large for the block is refused before any work; and the
call allocates no device memory (the free memory the runtime reports does not move, while an
ordinary first search on the same kind of handle does allocate)."""
import ctypes

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen
from harmony_amd._lib import HV_ERR_CAPACITY, HvError
from harmony_amd.index import DeviceArray, Index

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def build(xb):
    g = Index(128, 100, nprobe=10)
    g.train(xb, niter=10, seed=0)
    g.add(xb)
    return g


@pytest.fixture(scope="module")
def w1():
    say("no-alloc: workload 1, 10000 x 128, nlist 100")
    xb, xq = datagen.make(1)
    return xb, xq, build(xb)


def both(g, q, k):
    n = q.shape[0]
    qd = DeviceArray.from_host(q)
    D1, I1 = DeviceArray(n * k * 4), DeviceArray(n * k * 8)
    D2, I2 = DeviceArray(n * k * 4), DeviceArray(n * k * 8)
    g.search_device(qd, n, k, D1, I1)
    g.search_device_noalloc(qd, n, k, D2, I2)
    return ((D1.to_host(np.float32, (n, k)), I1.to_host(np.int64, (n, k))),
            (D2.to_host(np.float32, (n, k)), I2.to_host(np.int64, (n, k))))


FORMS = [("fp32", dict(precision=HV_PRECISION_FP32)), ("bf16", dict(precision=HV_PRECISION_BF16)),
         ("fp8", dict(precision=HV_PRECISION_FP8, rerank_r=2)),
         ("fp32-nopruning", dict(precision=HV_PRECISION_FP32, pruning=False))]


@pytest.mark.parametrize("name,kw", FORMS)
def test_noalloc_matches_search_device_twice(w1, name, kw):
    xb, xq, g = w1
    g.nprobe = 10
    g.preSearch(dim_blocks=4, **kw)
    for n in (100, 37):  # the second call is smaller
        say("no-alloc %s n=%d" % (name, n))
        (D1, I1), (D2, I2) = both(g, xq[:n], 10)
        np.testing.assert_array_equal(I2, I1)
        np.testing.assert_array_equal(D2, D1)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32, rerank_r=4)


def test_noalloc_uneven_and_learned(w1):
    xb, xq, g = w1
    g.nprobe = 10
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    for n in (100, 60):
        g.set_query_nprobe(np.random.default_rng(n).integers(1, 11, size=n).astype(np.int32))
        say("no-alloc uneven n=%d" % n)
        (D1, I1), (D2, I2) = both(g, xq[:n], 10)
        np.testing.assert_array_equal(I2, I1)
        np.testing.assert_array_equal(D2, D1)
    g.set_query_nprobe(None)
    # a predictor with untrained weights still exercises the learned buffers and kernels
    g.mlp_set_norm(np.zeros(65), np.ones(65), np.array([6.0, 0.5]), np.ones(2))
    g.mlp_init(3)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16, learned=True, learned_bound_scale=4.0)
    for n in (100, 37):
        say("no-alloc learned n=%d" % n)
        (D1, I1), (D2, I2) = both(g, xq[:n], 10)
        np.testing.assert_array_equal(I2, I1)
        np.testing.assert_array_equal(D2, D1)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32, learned=False)


def test_noalloc_refuses_oversized_batch(w1):
    xb, xq, g = w1
    g.nprobe = 100
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16, survivor_cap=8192)
    n = 5000
    q = np.ascontiguousarray(np.tile(xq, (n // xq.shape[0], 1)))
    qd = DeviceArray.from_host(q)
    D, I = DeviceArray(n * 100 * 4), DeviceArray(n * 100 * 8)
    sentinel = np.full((n, 100), -7, np.int64)
    from harmony_amd._lib import lib
    assert lib().hv_memcpy_htod(I.ptr, sentinel.ctypes.data_as(ctypes.c_void_p), sentinel.nbytes) == 0
    with pytest.raises(HvError) as e:
        g.search_device_noalloc(qd, n, 100, D, I)
    assert e.value.code == HV_ERR_CAPACITY and "MiB of scratch" in str(e.value)
    np.testing.assert_array_equal(I.to_host(np.int64, (n, 100)), sentinel)  # nothing was queued
    g.nprobe = 10
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32)


def test_noalloc_allocates_no_device_memory(w1):
    xb, xq, _ = w1
    g = build(xb)  # fresh handle: its search scratch is still unallocated
    g.nprobe = 10
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    n, k = 1000, 10
    q = np.ascontiguousarray(np.tile(xq, (10, 1)))
    qd = DeviceArray.from_host(q)
    D, I = DeviceArray(n * k * 4), DeviceArray(n * k * 8)
    before = free_device_bytes()
    g.search_device_noalloc(qd, n, k, D, I)
    g.search_device_noalloc(qd, n, k, D, I)
    after = free_device_bytes()
    say("no-alloc: free device memory moved by %d bytes" % (before - after))
    assert before - after <= (4 << 20)
    g.search_device(qd, n, k, D, I)  # the ordinary path sizes its own buffers
    assert before - free_device_bytes() >= (8 << 20)
