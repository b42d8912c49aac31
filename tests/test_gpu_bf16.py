"""GPU parity of the BF16 precision form: MFMA filter on BF16 residuals with a certified error
bound, then the FP32 distance of every survivor. Its results must be the FP32 path's results,
bit for bit (ids and distances), so every check here is against the CPU oracle or the FP32 path
on the same inputs."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP32, datagen
from harmony_amd.index import DeviceArray, Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def say(*a):
    print("[gpu-test bf16]", *a, flush=True)
    sys.stdout.flush()


def gpu_index(xb, nlist, nprobe, niter=10, dim_order=None):
    idx = Index(xb.shape[1], nlist, nprobe=nprobe, dim_order=dim_order)
    idx.train(xb, niter=niter, seed=0, max_points_per_centroid=256)
    idx.add(xb)
    return idx


@pytest.fixture(scope="module")
def w1():
    say("workload 1: 10000 x 128, nlist 100")
    xb, xq = datagen.make(1)
    g = gpu_index(xb, 100, 10)
    o = O.build_index(xb, 100, niter=10, seed=0)
    return xb, xq, g, o


@pytest.mark.parametrize("B,refine,wl", [(4, True, 10), (4, False, 10), (1, True, 10), (3, True, 10),
                                         (8, False, 10), (4, True, 0), (4, True, 3)])
def test_w1_bf16_matches_oracle(w1, B, refine, wl):
    xb, xq, g, o = w1
    say("w1 bf16 B=%d refine=%s warmup_list=%d" % (B, refine, wl))
    g.nprobe = 10
    g.preSearch(warmup_list=wl, dim_blocks=B, refine_bound=refine, precision=HV_PRECISION_BF16)
    D, I, st = g.search(xq, 10, return_stats=True)
    g.preSearch(precision=HV_PRECISION_FP32)
    oD, oI, oprobes, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=B, warmup_list=wl)
    np.testing.assert_array_equal(g.last_probes(xq.shape[0]), oprobes)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    assert st["overflow_queries"] == 0 and st["candidates"] > 0


def test_w1_bf16_bound_is_upper_bound(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, 10)
    b = g.last_bounds(xq.shape[0])
    g.preSearch(precision=HV_PRECISION_FP32)
    assert np.all(b >= D[:, -1])


@pytest.mark.parametrize("k", [1, 100])
def test_w1_bf16_full_probe_and_k(w1, k):
    xb, xq, g, o = w1
    say("w1 bf16 nprobe=100 (all lists) k=%d" % k)
    g.nprobe = 100
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, k)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32)
    fD, fI = g.search(xq, k)
    np.testing.assert_array_equal(I, fI)
    np.testing.assert_array_equal(D, fD)
    if k == 1:
        GD, GI = O.bruteforce_f64(xb, xq, 1)
        np.testing.assert_array_equal(I, GI)
    g.nprobe = 10


def test_w1_bf16_overflow_falls_back_to_exact_path(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    # no warm-up, so the buffer really holds 10 pairs (a warm-up raises it to its candidate count)
    g.preSearch(warmup_list=0, dim_blocks=4, survivor_cap=10, precision=HV_PRECISION_BF16)
    D, I, st = g.search(xq, 10, return_stats=True)
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP32)
    assert st["overflow_queries"] > 0
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_w1_bf16_wide_buffers(w1):
    """k = 600 over 16 lists (one scan phase after the nearest list) with no usable bound (each
    nearest list holds fewer than k vectors): over a thousand kept keys per query, so the final
    select reads the buffer in steps and hands the widest ones to the workgroup select."""
    xb, xq, g, o = w1
    g.nprobe = 16
    g.preSearch(warmup_list=0, dim_blocks=4, survivor_cap=8192, precision=HV_PRECISION_BF16)
    D, I, st = g.search(xq, 600, return_stats=True)
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP32)
    g.nprobe = 10
    assert st["overflow_queries"] == 0 and st["survivors"] > 1024 * xq.shape[0]
    oD, oI, _, _ = O.search(o, xq, 600, nprobe=16, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_w1_bf16_external_bound(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    oD, oI, _, obound = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    g.preSearch(dim_blocks=4, bound_mode=1, precision=HV_PRECISION_BF16)
    for b in (obound, np.full(xq.shape[0], FMAX, np.float32), oD[:, -1].copy()):
        g.set_external_bounds(b)
        D, I = g.search(xq, 10)
        np.testing.assert_array_equal(I, oI)
        np.testing.assert_array_equal(D, oD)
    g.preSearch(dim_blocks=4, bound_mode=0, precision=HV_PRECISION_FP32)


def test_w1_bf16_device_resident(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, 10)
    qd = DeviceArray.from_host(xq)
    Dd = DeviceArray(xq.shape[0] * 10 * 4)
    Id = DeviceArray(xq.shape[0] * 10 * 8)
    g.search_device(qd, xq.shape[0], 10, Dd, Id)
    g.preSearch(precision=HV_PRECISION_FP32)
    np.testing.assert_array_equal(Dd.to_host(np.float32, D.shape), D)
    np.testing.assert_array_equal(Id.to_host(np.int64, I.shape), I)


@pytest.mark.parametrize("d", [8, 33, 100])
def test_bf16_small_and_ragged_dims(d):
    """d not a multiple of 32 (zero-padded MFMA steps), k larger than the candidate count."""
    rng = np.random.default_rng(d)
    xb = rng.integers(0, 50, size=(300, d)).astype(np.float32)
    xq = rng.integers(0, 50, size=(7, d)).astype(np.float32)
    say("tiny index 300 x %d, k=64" % d)
    g = gpu_index(xb, 5, 2, niter=3)
    o = O.IndexData.from_gpu_index(g)
    g.preSearch(dim_blocks=2, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, 64)
    oD, oI, _, _ = O.search(o, xq, 64, nprobe=2, dim_blocks=2)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    g.nprobe = 5
    D, I = g.search(xq, 400)
    oD, oI, _, _ = O.search(o, xq, 400, nprobe=5, dim_blocks=2)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    assert np.all(I[:, 300:] == -1) and np.all(D[:, 300:] == FMAX)


@pytest.mark.parametrize("B", [8, 1])
def test_bf16_float_data_bit_exact(B):
    """Float data with a wide dynamic range (workload 3 sample, d = 420), variance dim order."""
    say("float workload 3 sample: 4000 x 420, nlist 40, B=%d" % B)
    xb, xq = datagen.make(3, nb=4000, nq=40)
    perm = datagen.variance_order(xb)
    g = gpu_index(xb, 40, 12, niter=4, dim_order=perm)
    o = O.IndexData.from_gpu_index(g)
    o = O.IndexData(o.centroids[:, perm], o.list_off, o.ids, o.codes[:, perm])
    g.preSearch(dim_blocks=B, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, 20)
    oD, oI, _, _ = O.search(o, xq[:, perm], 20, nprobe=12, dim_blocks=B)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    # the caller's bound at the exact 20th distance: the 20th pair survives only if its lower
    # bound is certified
    g.preSearch(dim_blocks=B, bound_mode=1, precision=HV_PRECISION_BF16)
    g.set_external_bounds(oD[:, -1].copy())
    D, I = g.search(xq, 20)
    g.preSearch(dim_blocks=B, bound_mode=0, precision=HV_PRECISION_BF16)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


@pytest.mark.parametrize("wl", [4])
def test_bf16_float_data_gaussian(wl):
    """Continuous data close to ties: FP32 gaussian rows, many near-equal distances."""
    rng = np.random.default_rng(11)
    xb = rng.standard_normal((20000, 96)).astype(np.float32)
    xq = rng.standard_normal((64, 96)).astype(np.float32)
    say("gaussian 20000 x 96, nlist 64")
    g = gpu_index(xb, 64, 16, niter=5)
    o = O.IndexData.from_gpu_index(g)
    g.preSearch(warmup_list=wl, dim_blocks=4, precision=HV_PRECISION_BF16)
    D, I = g.search(xq, 50)
    oD, oI, _, _ = O.search(o, xq, 50, nprobe=16, dim_blocks=4, warmup_list=wl)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


@pytest.fixture(scope="module")
def w2():
    say("workload 2: generating 1,000,000 x 128")
    xb, xq = datagen.make(2)
    say("workload 2: train + add on the GPU (nlist 1024)")
    g = gpu_index(xb, 1024, 16)
    return xb, xq, g


@pytest.mark.parametrize("nprobe", [16, 64, 256])
def test_w2_bf16_equals_fp32_full_batch(w2, nprobe):
    xb, xq, g = w2
    g.nprobe = nprobe
    say("w2 full batch nprobe=%d fp32" % nprobe)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32)
    D0, I0, s0 = g.search(xq, 100, return_stats=True)
    say("w2 full batch nprobe=%d bf16" % nprobe)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16)
    D1, I1, s1 = g.search(xq, 100, return_stats=True)
    g.preSearch(precision=HV_PRECISION_FP32)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    say("w2 nprobe=%d scan fp32 %.2f ms, bf16 %.2f ms; survivors/query fp32 %.0f bf16 %.0f; overflow %d"
        % (nprobe, s0["ms_scan"], s1["ms_scan"], s0["survivors"] / xq.shape[0], s1["survivors"] / xq.shape[0],
           s1["overflow_queries"]))


def test_w2_bf16_tight_external_bound(w2):
    """1M rows, 10000 queries: with the caller's bound at each query's exact 100th distance,
    every pair at the boundary must keep a lower bound at or below its exact distance."""
    xb, xq, g = w2
    g.nprobe = 16
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32)
    D0, I0 = g.search(xq, 100)
    g.preSearch(dim_blocks=4, bound_mode=1, precision=HV_PRECISION_BF16)
    g.set_external_bounds(D0[:, -1].copy())
    D1, I1, st = g.search(xq, 100, return_stats=True)
    g.preSearch(dim_blocks=4, bound_mode=0, precision=HV_PRECISION_FP32)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    say("w2 tight external bound: survivors/query %.1f" % (st["survivors"] / xq.shape[0]))
