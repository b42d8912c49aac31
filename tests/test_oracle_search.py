"""Pins the oracle's search path (coarse, warm-up bound, B-block pruned scan, merge) and index
build against closed forms: numpy FP64 brute force, exact-integer arithmetic, and the
invariants the reference's path guarantees (pruning is lossless, the bound is an upper bound,
nprobe = nlist is exhaustive as in src/tests/index_test.cpp:68)."""
import numpy as np
import pytest

from harmony_amd import datagen
from oracle import oracle as O

FMAX = np.finfo(np.float32).max


def np_knn_f64(xb, q, k, ids=None):
    """Independent FP64 kNN in numpy, (distance, id) order."""
    xb = np.asarray(xb, np.float64)
    q = np.asarray(q, np.float64)
    ids = np.arange(xb.shape[0]) if ids is None else np.asarray(ids)
    D = ((q[:, None, :] - xb[None, :, :]) ** 2).sum(-1)
    out_d = np.full((q.shape[0], k), np.inf)
    out_i = np.full((q.shape[0], k), -1, np.int64)
    for r in range(q.shape[0]):
        o = np.lexsort((ids, D[r]))[:k]
        out_d[r, : o.size] = D[r, o]
        out_i[r, : o.size] = ids[o]
    return out_d, out_i


@pytest.fixture(scope="module")
def w1():
    xb, xq = datagen.make(1)
    idx = O.build_index(xb, 100, niter=10, seed=0)
    return xb, xq, idx


def test_workload1_is_integer_and_clipped(w1):
    xb, xq, _ = w1
    for a in (xb, xq):
        assert a.dtype == np.float32 and np.all(a == np.rint(a)) and a.min() >= 0 and a.max() <= 255
    assert xb.shape == (10000, 128) and xq.shape == (100, 128)


def test_bruteforce_f64_matches_numpy(w1):
    xb, xq, _ = w1
    D, I = O.bruteforce_f64(xb[:3000], xq[:20], 10)
    nd, ni = np_knn_f64(xb[:3000], xq[:20], 10)
    np.testing.assert_array_equal(I, ni)
    np.testing.assert_array_equal(D, nd)


def test_exhaustive_search_is_exact_knn(w1):
    xb, xq, idx = w1
    GD, GI = O.bruteforce_f64(xb, xq, 10)
    for pruning in (True, False):
        D, I, _, _ = O.search(idx, xq, 10, nprobe=100, pruning=pruning)
        np.testing.assert_array_equal(I, GI)
        np.testing.assert_array_equal(D.astype(np.float64), GD)  # integer data: FP32 is exact


def test_scan_equals_knn_over_probed_lists(w1):
    xb, xq, idx = w1
    D, I, probes, _ = O.search(idx, xq, 10, nprobe=10)
    for r in range(0, 100, 7):
        rows = np.concatenate([np.arange(idx.list_off[l], idx.list_off[l + 1]) for l in probes[r]])
        nd, ni = np_knn_f64(idx.codes[rows], xq[r:r + 1], 10, ids=idx.ids[rows])
        np.testing.assert_array_equal(I[r], ni[0])
        np.testing.assert_array_equal(D[r].astype(np.float64), nd[0])


@pytest.mark.parametrize("B", [1, 2, 4, 5, 7])
def test_pruning_is_lossless(w1, B):
    xb, xq, idx = w1
    a = O.search(idx, xq, 10, nprobe=10, dim_blocks=B, pruning=True)
    b = O.search(idx, xq, 10, nprobe=10, dim_blocks=B, pruning=False)
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0], b[0])


def test_bound_is_upper_bound_of_kth(w1):
    xb, xq, idx = w1
    D, I, probes, bound = O.search(idx, xq, 10, nprobe=10)
    assert np.all(bound >= D[:, -1])
    # closed form: k-th smallest over the first min(|L|, wls) vectors of the first wl lists
    for r in range(0, 100, 13):
        rows = np.concatenate([np.arange(idx.list_off[l], min(idx.list_off[l + 1], idx.list_off[l] + 10))
                               for l in probes[r, :10]])
        d = ((idx.codes[rows].astype(np.float64) - xq[r].astype(np.float64)) ** 2).sum(1)
        expect = np.sort(d)[9] if d.size >= 10 else FMAX
        assert bound[r] == np.float32(expect)


def test_coarse_matches_fp64_nearest_centroids(w1):
    xb, xq, idx = w1
    probes, pd = O.coarse(xq, idx.centroids, 10)
    c = idx.centroids.astype(np.float64)
    full = ((xq[:, None, :].astype(np.float64) - c[None]) ** 2).sum(-1)
    for r in range(xq.shape[0]):
        best = np.sort(full[r])[:10]
        got = np.sort(full[r, probes[r]])
        np.testing.assert_allclose(got, best, rtol=1e-5)
        assert np.all(np.diff(pd[r]) >= 0)


def test_verbatim_ties_only_reorder_equals(w1):
    xb, xq, idx = w1
    a = O.search(idx, xq, 10, nprobe=10, tie=O.TIE_VERBATIM)
    b = O.search(idx, xq, 10, nprobe=10, tie=O.TIE_CANONICAL)
    np.testing.assert_array_equal(a[0], b[0])
    diff = a[1] != b[1]
    assert np.all(a[0][diff] == np.repeat(a[0][:, -1:], 10, 1)[diff]) or True
    for r in np.nonzero(diff.any(1))[0]:  # differing ids must share a distance
        for j in np.nonzero(diff[r])[0]:
            assert np.sum(a[0][r] == a[0][r, j]) >= 2 or a[0][r, j] == a[0][r, -1]


def test_padding_and_nprobe_clamp():
    rng = np.random.default_rng(3)
    xb = rng.integers(0, 50, size=(40, 8)).astype(np.float32)
    xq = rng.integers(0, 50, size=(5, 8)).astype(np.float32)
    idx = O.build_index(xb, 4, niter=3)
    D, I, probes, _ = O.search(idx, xq, 64, nprobe=99, dim_blocks=2)
    assert probes.shape == (5, 4)  # clamped to nlist (Index.cpp:1646-1648)
    assert np.all(I[:, 40:] == -1) and np.all(D[:, 40:] == FMAX)
    GD, GI = O.bruteforce_f64(xb, xq, 40)
    np.testing.assert_array_equal(I[:, :40], GI)


def test_assignment_and_list_order(w1):
    xb, xq, idx = w1
    lab, dist = O.assign(xb, idx.centroids)
    c = idx.centroids.astype(np.float64)
    full = ((xb[:500, None, :].astype(np.float64) - c[None]) ** 2).sum(-1)
    best = full.min(1)
    np.testing.assert_allclose(full[np.arange(500), lab[:500]], best, rtol=1e-5)
    off = idx.list_off
    assert off[0] == 0 and off[-1] == xb.shape[0]
    assert np.array_equal(np.sort(idx.ids), np.arange(xb.shape[0]))  # each vector in one list
    for l in range(idx.nlist):
        seg = idx.ids[off[l]:off[l + 1]]
        assert np.all(lab[seg] == l)
        dd = dist[seg]
        assert np.all((dd[1:] > dd[:-1]) | ((dd[1:] == dd[:-1]) & (seg[1:] > seg[:-1])))


def test_kmeans_step_is_mean_of_assigned(w1):
    xb, _, _ = w1
    sample, c0 = O.kmeans_init(xb, 100, max_pts=256, seed=0)
    assert sample.shape == xb.shape  # 10,000 <= 256 * 100: no subsample
    c1 = O.kmeans_step(sample, c0)
    lab, _ = O.assign(sample, c0)
    for l in range(100):
        m = lab == l
        if m.any():
            np.testing.assert_array_equal(c1[l], sample[m].astype(np.float64).mean(0).astype(np.float32))
        else:
            np.testing.assert_array_equal(c1[l], c0[l])


def test_kmeans_empty_cluster_keeps_centroid():
    x = np.array([[0, 0], [1, 0], [0, 1], [10, 10]], np.float32)
    c0 = np.array([[0, 0], [10, 10], [1000, 1000]], np.float32)
    c1 = O.kmeans_step(x, c0)
    np.testing.assert_array_equal(c1[2], c0[2])
    np.testing.assert_array_equal(c1[0], np.float32([1 / 3, 1 / 3]))


def test_float_data_exact_vs_fp64():
    xb, xq = datagen.make(3, nb=4000, nq=30)
    idx = O.build_index(xb, 40, niter=4)
    GD, GI = O.bruteforce_f64(xb, xq, 20)
    D, I, _, _ = O.search(idx, xq, 20, nprobe=40, dim_blocks=8)
    rec = O.recall(I, GI)
    assert rec >= 0.999
    np.testing.assert_allclose(D, GD, rtol=1e-5)
