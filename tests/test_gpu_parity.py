"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the same inputs.

Bar: bit-exact ids and FP32 distances (the GPU uses the same FMA chain and block sums as the
oracle), on integer-valued and on float data. At the 1M-vector workload, oracle parity is
checked on query subsets and the full batch is checked through size-independent properties
(pruning on == off, results sorted, bound >= k-th distance)."""
import sys

import numpy as np
import pytest

from harmony_amd import datagen
from harmony_amd.index import DeviceArray, Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def say(*a):
    print("[gpu-test]", *a, flush=True)
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


def test_w1_build_matches_oracle(w1):
    xb, xq, g, o = w1
    np.testing.assert_array_equal(g.centroids(), o.centroids)
    np.testing.assert_array_equal(g.list_offsets(), o.list_off)
    np.testing.assert_array_equal(g.list_ids(), o.ids)
    np.testing.assert_array_equal(g.list_codes(), o.codes)


@pytest.mark.parametrize("B,pruning,refine", [(4, True, True), (4, True, False), (4, False, True), (1, True, True),
                                              (3, True, True), (8, True, False)])
def test_w1_search_matches_oracle(w1, B, pruning, refine):
    xb, xq, g, o = w1
    say("w1 search B=%d pruning=%s refine=%s" % (B, pruning, refine))
    g.nprobe = 10
    g.preSearch(warmup_list=10, warmup_list_size=0, pruning=pruning, dim_blocks=B, refine_bound=refine)
    D, I, st = g.search(xq, 10, return_stats=True)
    oD, oI, oprobes, obound = O.search(o, xq, 10, nprobe=10, dim_blocks=B, pruning=pruning)
    np.testing.assert_array_equal(g.last_probes(xq.shape[0]), oprobes)
    np.testing.assert_array_equal(g.last_bounds(xq.shape[0]), obound)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    assert st["overflow_queries"] == 0 and st["candidates"] > 0


def test_w1_exact_vs_fp64_at_full_probe(w1):
    xb, xq, g, o = w1
    g.nprobe = 100
    g.preSearch(pruning=True, dim_blocks=4)
    D, I = g.search(xq, 10)
    GD, GI = O.bruteforce_f64(xb, xq, 10)
    np.testing.assert_array_equal(I, GI)
    np.testing.assert_array_equal(D.astype(np.float64), GD)
    gD, gI = g.ground_truth_f64(xq, 10)  # the library's FP64 brute force
    np.testing.assert_array_equal(gI, GI)
    np.testing.assert_array_equal(gD, GD)


def test_overflow_falls_back_to_exact_path(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(pruning=True, dim_blocks=4, survivor_cap=10)
    D, I, st = g.search(xq, 10, return_stats=True)
    assert st["overflow_queries"] > 0
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


@pytest.mark.parametrize("precision", [0, 1])
def test_overflow_redo_in_several_batches(w1, precision):
    """Every query overflows a 1024-pair buffer at k = 1024 over all 100 lists (no warm-up bound):
    the exact redo runs in batches (its scratch holds about 80 such queries), the later ones
    added by the host."""
    xb, xq, g, o = w1
    _, q = datagen.make(1, nb=1, nq=200)
    g.nprobe = 100
    g.preSearch(warmup_list=0, dim_blocks=4, survivor_cap=1024, precision=precision)
    D, I, st = g.search(q, 1024, return_stats=True)
    g.preSearch(warmup_list=10, dim_blocks=4, precision=0)
    g.nprobe = 10
    assert st["overflow_queries"] == q.shape[0]
    oD, oI, _, _ = O.search(o, q, 1024, nprobe=100, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nmax", [10, 60])
def test_uneven_nprobe_matches_oracle_per_query(w1, precision, nmax):
    """Per-query probe counts: each query's results are the oracle's at its own nprobe. Up to 60
    probes the BF16 form scans the ranks in groups ([1, 4), [4, 16), [16, 60)) with a
    compaction after each, which skips the queries that gained nothing."""
    xb, xq, g, o = w1
    npq = np.random.default_rng(3).integers(1, nmax + 1, size=xq.shape[0]).astype(np.int32)
    g.nprobe = nmax
    g.preSearch(pruning=True, dim_blocks=4, precision=precision)
    g.set_query_nprobe(npq)
    D, I = g.search(xq, 10)
    got = g.last_nprobe(xq.shape[0])
    g.set_query_nprobe(None)
    g.preSearch(pruning=True, dim_blocks=4, precision=0)
    np.testing.assert_array_equal(got, npq)
    for p in np.unique(npq):
        rows = np.where(npq == p)[0]
        oD, oI, _, _ = O.search(o, xq[rows], 10, nprobe=int(p), dim_blocks=4)
        np.testing.assert_array_equal(I[rows], oI)
        np.testing.assert_array_equal(D[rows], oD)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_graph_replays_match_eager(w1, precision):
    """Graph mode: the first call sizes the buffers, the second captures, later ones replay.
    Every call (with eager searches of another shape in between) returns the eager results."""
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(pruning=True, dim_blocks=4, precision=precision, graph=False)
    eD, eI = g.search(xq, 10)
    for it in range(5):
        g.preSearch(pruning=True, dim_blocks=4, precision=precision, graph=True)
        D, I = g.search(xq, 10)
        np.testing.assert_array_equal(I, eI)
        np.testing.assert_array_equal(D, eD)
        g.preSearch(pruning=True, dim_blocks=4, precision=0, graph=False)
        sD, sI = g.search(xq[:7], 5)
        oD, oI, _, _ = O.search(o, xq[:7], 5, nprobe=10, dim_blocks=4)
        np.testing.assert_array_equal(sI, oI)
    g.preSearch(pruning=True, dim_blocks=4, precision=0, graph=False)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_graph_replay_follows_per_query_nprobe(w1, precision):
    """Per-query probe counts change between replays of one captured graph (they are copied from
    pinned memory at replay). FP8: a shortlist wider than any query's candidates, so exact."""
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(pruning=True, dim_blocks=4, precision=precision, graph=True, rerank_r=300)
    rng = np.random.default_rng(5)
    for it in range(4):
        npq = rng.integers(1, 11, size=xq.shape[0]).astype(np.int32)
        g.set_query_nprobe(npq)
        D, I = g.search(xq, 10)
        for p in np.unique(npq):
            rows = np.where(npq == p)[0]
            oD, oI, _, _ = O.search(o, xq[rows], 10, nprobe=int(p), dim_blocks=4)
            np.testing.assert_array_equal(I[rows], oI)
            np.testing.assert_array_equal(D[rows], oD)
    g.set_query_nprobe(None)
    g.preSearch(pruning=True, dim_blocks=4, precision=0, graph=False, rerank_r=4)


@pytest.mark.parametrize("precision", [0, 1])
def test_autotune_keeps_results(w1, precision):
    from harmony_amd.autotune import autotune
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(pruning=True, dim_blocks=4, precision=precision)
    best, trials = autotune(g, xq, 10, reps=1, log=say)
    assert len(trials) >= 3 and best["ms"] == min(trials.values())
    D, I = g.search(xq, 10)
    g.preSearch(pruning=True, dim_blocks=4, precision=0, graph=False, tile_shape=0, tile_shape_first=0,
                scan_vsplit=0)
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_external_bound(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    oD, oI, _, obound = O.search(o, xq, 10, nprobe=10, dim_blocks=4)
    g.preSearch(pruning=True, dim_blocks=4, bound_mode=1)
    for b in (obound, np.full(xq.shape[0], FMAX, np.float32), oD[:, -1].copy()):
        g.set_external_bounds(b)
        D, I = g.search(xq, 10)
        np.testing.assert_array_equal(I, oI)
        np.testing.assert_array_equal(D, oD)
    g.preSearch(pruning=True, dim_blocks=4)


def test_device_resident_search_matches_host(w1):
    xb, xq, g, o = w1
    g.nprobe = 10
    g.preSearch(pruning=True, dim_blocks=4)
    D, I = g.search(xq, 10)
    qd = DeviceArray.from_host(xq)
    Dd = DeviceArray(xq.shape[0] * 10 * 4)
    Id = DeviceArray(xq.shape[0] * 10 * 8)
    g.search_device(qd, xq.shape[0], 10, Dd, Id)
    np.testing.assert_array_equal(Dd.to_host(np.float32, D.shape), D)
    np.testing.assert_array_equal(Id.to_host(np.int64, I.shape), I)


def test_padding_and_clamp():
    rng = np.random.default_rng(3)
    xb = rng.integers(0, 50, size=(40, 8)).astype(np.float32)
    xq = rng.integers(0, 50, size=(5, 8)).astype(np.float32)
    say("tiny index 40 x 8, k=64 > candidates")
    g = gpu_index(xb, 4, 99, niter=3)
    o = O.IndexData.from_gpu_index(g)
    g.preSearch(pruning=True, dim_blocks=2)
    D, I = g.search(xq, 64)
    oD, oI, _, _ = O.search(o, xq, 64, nprobe=99, dim_blocks=2)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    assert np.all(I[:, 40:] == -1) and np.all(D[:, 40:] == FMAX)


@pytest.fixture(scope="module")
def w3s():
    say("float workload 3 sample: 4000 x 420, nlist 40")
    xb, xq = datagen.make(3, nb=4000, nq=40)
    perm = datagen.variance_order(xb)
    g = gpu_index(xb, 40, 12, niter=4, dim_order=perm)
    o_c = O.kmeans(xb[:, perm], 40, niter=4)
    np.testing.assert_array_equal(g.centroids()[:, perm], o_c)
    o = O.IndexData.from_gpu_index(g)
    o = O.IndexData(o.centroids[:, perm], o.list_off, o.ids, o.codes[:, perm])
    return xq, perm, g, o


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("B,pruning", [(8, True), (1, True), (8, False), (4, False)])
def test_float_data_bit_exact(w3s, B, pruning, precision):
    """Float data (d = 420, wide per-dimension scales): pruning on sums B per-block chains
    (node.cpp:639-651), pruning off one full chain (node.cpp:658-667, the ORIGINAL scan of
    Index.cpp:1420-1554); the two differ in the last bits here, and each matches the oracle's."""
    xq, perm, g, o = w3s
    say("w3 sample B=%d pruning=%s precision=%d" % (B, pruning, precision))
    g.nprobe = 12
    g.preSearch(pruning=pruning, dim_blocks=B, precision=precision)
    D, I = g.search(xq, 20)
    g.preSearch(pruning=True, dim_blocks=4, precision=0)
    oD, oI, _, _ = O.search(o, xq[:, perm], 20, nprobe=12, dim_blocks=B, pruning=pruning)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    if not pruning:
        # the full chain is not the blocked sum on this data, so the test tells them apart
        bD, _, _, _ = O.search(o, xq[:, perm], 20, nprobe=12, dim_blocks=8, pruning=True)
        assert np.any(bD != oD)


@pytest.fixture(scope="module")
def w2():
    say("workload 2: generating 1,000,000 x 128")
    xb, xq = datagen.make(2)
    say("workload 2: train + add on the GPU (nlist 1024)")
    g = gpu_index(xb, 1024, 16)
    say("workload 2: export index for the oracle")
    o = O.IndexData.from_gpu_index(g)
    return xb, xq, g, o


@pytest.mark.parametrize("nprobe", [16, 64, 256])
def test_w2_pruning_lossless_full_batch(w2, nprobe):
    xb, xq, g, o = w2
    g.nprobe = nprobe
    say("w2 full batch nprobe=%d pruning on" % nprobe)
    g.preSearch(pruning=True, dim_blocks=4)
    D1, I1, s1 = g.search(xq, 100, return_stats=True)
    say("w2 full batch nprobe=%d pruning off" % nprobe)
    g.preSearch(pruning=False, dim_blocks=4)
    D0, I0, s0 = g.search(xq, 100, return_stats=True)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    assert np.all(np.diff(D1, axis=1) >= 0)
    assert np.all(I1 >= 0)
    b = g.last_bounds(xq.shape[0])
    assert np.all(b >= D0[:, -1])
    say("w2 nprobe=%d executed block fraction %.3f, overflow %d, scan %.2f ms"
        % (nprobe, s1["block_evals"] / max(1, s1["block_evals_max"]), s1["overflow_queries"], s1["ms_scan"]))
    g.preSearch(pruning=True, dim_blocks=4)


@pytest.mark.parametrize("nprobe", [16, 64, 256])
def test_w2_subset_matches_oracle(w2, nprobe):
    xb, xq, g, o = w2
    sub = xq[::50]  # 200 queries
    g.nprobe = nprobe
    g.preSearch(pruning=True, dim_blocks=4)
    say("w2 subset nprobe=%d: GPU" % nprobe)
    D, I = g.search(sub, 100)
    say("w2 subset nprobe=%d: oracle" % nprobe)
    oD, oI, oprobes, obound = O.search(o, sub, 100, nprobe=nprobe, dim_blocks=4)
    np.testing.assert_array_equal(g.last_probes(sub.shape[0]), oprobes)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_w2_ground_truth_and_recall(w2):
    xb, xq, g, o = w2
    sub = xq[:40]
    g.nprobe = 64
    g.preSearch(pruning=True, dim_blocks=4)
    say("w2 fp64 ground truth (GPU) on 40 queries")
    gD, gI = g.ground_truth_f64(sub, 100)
    say("w2 fp64 brute force (CPU oracle) on 40 queries")
    GD, GI = O.bruteforce_f64(xb, sub, 100)
    np.testing.assert_array_equal(gI, GI)
    np.testing.assert_array_equal(gD, GD)
    D, I = g.search(sub, 100)
    rec = O.recall(I, GI)
    say("w2 recall@100 at nprobe 64 on 40 queries: %.4f" % rec)
    assert rec > 0.9


def test_w1_matches_golden_fixture(w1):
    import os
    xb, xq, g, o = w1
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "w1_golden.npz"),
                   allow_pickle=False)
    g.nprobe = 10
    g.preSearch(warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=4)
    D, I = g.search(xq, 10)
    np.testing.assert_array_equal(I, gold["I"])
    np.testing.assert_array_equal(D, gold["D"])
    np.testing.assert_array_equal(g.list_offsets(), gold["list_off"])
