"""The uneven batch (Zipf-skewed unit-vector mixture, d = 96): per-query nprobe drawn log-uniform
in [8, 512] and k in {10, 100}, every query served by ONE search call (graph-captured, replayed).

At an oracle-covered size (200,000 x 96, nlist 256, 2,000 queries) the plain and BF16 forms are
compared with the oracle query by query at each query's own (nprobe, k): ids and distances bit for
bit, padding after the query's k. The FP8 form's returned distances are the plain path's exact
distances of those ids, and its recall against the plain path is measured. At the full size
(10,000,000 x 96, nlist 4096, 20,000 queries) the BF16 form must equal the plain form exactly and
the FP8 form must keep recall@k >= 0.95 against it."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, datagen, metrics
from harmony_amd.index import DeviceArray, Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def say(*a):
    print("[gpu-test uneven]", *a, flush=True)
    sys.stdout.flush()


def one_call(g, xq, npq, kq, kmax, graph=True, reps=3):
    """Per-query nprobe and k set once, the whole batch in one call; with graph mode the first
    call sizes the buffers, the second captures, the later ones replay (all must agree)."""
    n = xq.shape[0]
    g.set_query_nprobe(npq)
    g.set_query_k(kq)
    g._params.graph = 1 if graph else 0
    g._push()
    qd, Dd, Id = DeviceArray.from_host(xq), DeviceArray(n * kmax * 4), DeviceArray(n * kmax * 8)
    out = None
    for _ in range(reps):
        g.search_device(qd, n, kmax, Dd, Id)
        cur = (Dd.to_host(np.float32, (n, kmax)), Id.to_host(np.int64, (n, kmax)))
        if out is not None:
            np.testing.assert_array_equal(cur[1], out[1])
            np.testing.assert_array_equal(cur[0], out[0])
        out = cur
    g.set_query_nprobe(None)
    g.set_query_k(None)
    g._params.graph = 0
    g._push()
    return out


@pytest.fixture(scope="module")
def w4s():
    say("uneven sample: 200,000 x 96 Zipf unit mixture, nlist 256, 2,000 queries")
    xb, xq = datagen.make(4, nb=200000, nq=2000)
    npq, kq = datagen.uneven_plan(4, 2000)
    g = Index(96, 256, nprobe=512)
    g.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    g.add(xb)
    o = O.IndexData.from_gpu_index(g)
    return xq, npq, kq, g, o


def oracle_rows(o, xq, npq, kq, kmax, nlist):
    """The oracle at each query's own (min(nprobe, nlist), k), padded to kmax."""
    D = np.full((xq.shape[0], kmax), FMAX, np.float32)
    I = np.full((xq.shape[0], kmax), -1, np.int64)
    eff = np.minimum(npq, nlist)
    for p, kk in sorted(set(zip(eff.tolist(), kq.tolist()))):
        rows = np.nonzero((eff == p) & (kq == kk))[0]
        oD, oI, _, _ = O.search(o, xq[rows], int(kk), nprobe=int(p), dim_blocks=4)
        D[rows, :kk] = oD
        I[rows, :kk] = oI
    return D, I


@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
def test_uneven_one_call_matches_oracle(w4s, precision):
    xq, npq, kq, g, o = w4s
    g.nprobe = 512
    g.preSearch(pruning=True, dim_blocks=4, precision=precision)
    say("one call, precision %d: nprobe mean %.1f, k in {10, 100}" % (precision, npq.mean()))
    D, I = one_call(g, xq, npq, kq, 100)
    oD, oI = oracle_rows(o, xq, npq, kq, 100, 256)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    assert np.all(I[kq == 10][:, 10:] == -1) and np.all(D[kq == 10][:, 10:] == FMAX)
    g.preSearch(pruning=True, dim_blocks=4, precision=HV_PRECISION_FP32)


def test_uneven_fp8_exact_distances_and_recall(w4s):
    xq, npq, kq, g, o = w4s
    g.nprobe = 512
    g.preSearch(pruning=True, dim_blocks=4, precision=HV_PRECISION_FP32)
    P = one_call(g, xq, npq, kq, 100, graph=False, reps=1)
    g.preSearch(pruning=True, dim_blocks=4, precision=HV_PRECISION_FP8, rerank_r=2)
    D, I = one_call(g, xq, npq, kq, 100)
    rec = []
    for q in range(xq.shape[0]):
        kk = int(kq[q])
        m = dict(zip(P[1][q, :kk].tolist(), P[0][q, :kk].tolist()))
        for dist, j in zip(D[q, :kk], I[q, :kk]):
            if int(j) in m:  # a returned id that the plain path also returns: the same FP32 distance
                assert dist == m[int(j)]
        rec.append(len(set(I[q, :kk].tolist()) & set(m)) / kk)
        assert np.all(I[q, kk:] == -1)
        assert np.all(np.diff(D[q, :kk]) >= 0)
    say("uneven sample fp8 r=2: recall@k vs plain %.5f" % np.mean(rec))
    assert np.mean(rec) >= 0.95
    g.preSearch(pruning=True, dim_blocks=4, precision=HV_PRECISION_FP32, rerank_r=4)


def test_uneven_full_size_properties():
    say("uneven full size: generating 10,000,000 x 96")
    xb, xq = datagen.make(4)
    npq, kq = datagen.uneven_plan(4, xq.shape[0])
    say("uneven full size: train + add (nlist 4096)")
    g = Index(96, 4096, nprobe=512)
    g.train(xb, niter=10, seed=0, max_points_per_centroid=256)
    g.add(xb)
    del xb
    sizes = np.diff(g.list_offsets())
    say("list sizes min %d median %d max %d" % (sizes.min(), int(np.median(sizes)), sizes.max()))
    res = {}
    for name, prec, r in (("plain", HV_PRECISION_FP32, 4), ("bf16", HV_PRECISION_BF16, 4), ("fp8", HV_PRECISION_FP8, 2)):
        g.preSearch(pruning=True, dim_blocks=4, precision=prec, rerank_r=r)
        say("full size %s: one graph call for 20,000 queries" % name)
        res[name] = one_call(g, xq, npq, kq, 100)
    D, I = res["plain"]
    np.testing.assert_array_equal(res["bf16"][1], I)
    np.testing.assert_array_equal(res["bf16"][0], D)
    for q in np.nonzero(kq == 10)[0]:
        assert np.all(I[q, 10:] == -1)
    assert np.all(I[kq == 100] >= 0)
    for name in ("plain", "fp8"):
        Dn = res[name][0]
        for kk in (10, 100):
            rows = kq == kk
            assert np.all(np.diff(Dn[rows][:, :kk], axis=1) >= 0)
    rec = np.mean([metrics.recall(res["fp8"][1][kq == kk][:, :kk], I[kq == kk][:, :kk]) for kk in (10, 100)])
    say("full size fp8 r=2 recall vs plain %.5f" % rec)
    assert rec >= 0.95
