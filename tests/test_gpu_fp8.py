"""GPU checks of the FP8 precision form: e4m3 MFMA estimates with power-of-two scales per list
and per query slot, a shortlist of the k * rerank_r smallest estimates per query, then the FP32
distance of every shortlisted pair. The form is approximate, so the checks are what it does
guarantee: every returned distance is the plain path's exact distance of the returned id, a
shortlist wide enough to hold every candidate gives the plain path's results bit for bit, the
overflow path is exact, runs repeat bit for bit, and the recall against the plain path's top-k
at the same nprobe is measured on the 1M workload."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_FP8, HV_PRECISION_FP32, HvError, datagen
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test fp8]", *a, flush=True)
    sys.stdout.flush()


def gpu_index(xb, nlist, nprobe, niter=10):
    idx = Index(xb.shape[1], nlist, nprobe=nprobe)
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


def restore(g):
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP32, rerank_r=4)
    g.nprobe = 10


def test_fp8_wide_shortlist_equals_plain(w1):
    # every nearest list holds fewer than k * r vectors, so nothing is filtered and every
    # candidate is re-ranked: the plain path's results exactly
    xb, xq, g, o = w1
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP8, rerank_r=100)
    D, I = g.search(xq, 10)
    restore(g)
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


@pytest.mark.parametrize("r", [1, 2, 4])
def test_fp8_distances_are_exact_and_recall(w1, r):
    xb, xq, g, o = w1
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP8, rerank_r=r)
    D, I = g.search(xq, 10)
    D2, I2 = g.search(xq, 10)
    restore(g)
    np.testing.assert_array_equal(I, I2)
    np.testing.assert_array_equal(D, D2)
    # all candidates of the probed lists, exactly (k above every query's candidate count)
    aD, aI, _, _ = O.search(o, xq, 1024, nprobe=10, dim_blocks=4)
    hit = 0
    for q in range(xq.shape[0]):
        exact = dict(zip(aI[q][aI[q] >= 0].tolist(), aD[q][aI[q] >= 0].tolist()))
        assert np.all(np.diff(D[q]) >= 0)
        for dist, i in zip(D[q], I[q]):
            assert i >= 0 and int(i) in exact
            assert np.float32(exact[int(i)]) == dist
        hit += len(set(I[q].tolist()) & set(aI[q][:10].tolist()))
    rec = hit / (10.0 * xq.shape[0])
    say("r", r, "recall@10 vs the plain path", rec)
    assert rec >= {1: 0.7, 2: 0.85, 4: 0.95}[r]


def test_fp8_overflow_is_exact(w1):
    xb, xq, g, o = w1
    g.nprobe = 40
    g.preSearch(warmup_list=10, dim_blocks=4, precision=HV_PRECISION_FP8, rerank_r=20, survivor_cap=200)
    D, I, st = g.search(xq, 10, return_stats=True)
    restore(g)
    # no filter (each nearest list is shorter than k * r): every buffer overflows, and the exact
    # path redoes every query
    assert st["overflow_queries"] == xq.shape[0]
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=40, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_fp8_parameter_checks(w1):
    xb, xq, g, o = w1
    with pytest.raises(HvError):
        g.preSearch(precision=HV_PRECISION_FP8, rerank_r=0)
    g.preSearch(precision=HV_PRECISION_FP8, rerank_r=50)
    with pytest.raises(HvError):
        g.search(xq, 100)  # k * r = 5000 > 4096
    restore(g)


def test_w2_fp8_recall():
    say("workload 2: 1M x 128, nlist 1024")
    xb, xq = datagen.make(2)
    g = gpu_index(xb, 1024, 16)
    del xb
    q = xq[:2000]
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32)
    pD, pI = g.search(q, 100)
    for r in (2, 4):
        g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP8, rerank_r=r)
        D, I = g.search(q, 100)
        rec = np.mean([len(set(I[i].tolist()) & set(pI[i].tolist())) / 100.0 for i in range(q.shape[0])])
        say("1M, nprobe 16, r", r, "recall@100 vs the plain path", rec)
        assert rec >= (0.9 if r == 2 else 0.95)
        # returned distances equal the plain path's wherever the id is in the plain top-100
        for i in range(0, q.shape[0], 97):
            m = dict(zip(pI[i].tolist(), pD[i].tolist()))
            for dist, j in zip(D[i], I[i]):
                if int(j) in m:
                    assert np.float32(m[int(j)]) == dist
    g.close()
