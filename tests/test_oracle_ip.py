"""The oracle's inner-product path (Index.cpp:408-418, 567-671, 1509-1551; Clustering.cpp:121-151)
pinned by closed forms: the search equals the FP64 brute force over the probed lists, the coarse
probes are the largest inner products (ties to the lower id), triangle pruning on unit vectors
prunes (vectors skipped) and changes nothing, assignment is the argmax, lists are ordered by
descending inner product with the centroid, k-means is the mean of the rows assigned by
normalized inner product, and its centroids end normalized."""
import numpy as np
import pytest

from oracle import oracle as O


def unit(x):
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


@pytest.fixture(scope="module")
def ipdata():
    rng = np.random.default_rng(11)
    centers = unit(rng.normal(size=(30, 32)))
    z = rng.integers(0, 30, size=3000)
    xb = unit(centers[z] + 0.08 * rng.normal(size=(3000, 32)).astype(np.float32))
    xq = unit(centers[rng.integers(0, 30, size=60)] + 0.08 * rng.normal(size=(60, 32)).astype(np.float32))
    idx = O.ip_build_index(xb, 40, niter=8, seed=0)
    return xb, xq, idx


def brute_ip(xb, xq, k, allowed=None):
    s = xq.astype(np.float64) @ xb.astype(np.float64).T
    if allowed is not None:
        s = np.where(allowed, s, -np.inf)
    I = np.argsort(-s, axis=1, kind="stable")[:, :k]
    return np.take_along_axis(s, I, axis=1), I


def test_ip_search_full_probe_is_brute_force(ipdata):
    xb, xq, idx = ipdata
    D, I, _ = O.ip_search(idx, xq, 10, nprobe=40)
    GD, GI = brute_ip(xb, xq, 10)
    np.testing.assert_array_equal(I, GI)
    np.testing.assert_allclose(D, GD, rtol=1e-5, atol=1e-6)
    assert np.all(np.diff(D, axis=1) <= 0)


@pytest.mark.parametrize("nprobe", [1, 3, 8])
def test_ip_search_is_brute_force_over_probed_lists(ipdata, nprobe):
    xb, xq, idx = ipdata
    D, I, probes = O.ip_search(idx, xq, 10, nprobe=nprobe)
    # probes: the largest <q, c>, ties to the lower id
    s = xq.astype(np.float64) @ idx.centroids.astype(np.float64).T
    np.testing.assert_array_equal(probes, np.argsort(-s, axis=1, kind="stable")[:, :nprobe])
    owner = np.empty(xb.shape[0], np.int64)
    for l in range(idx.nlist):
        owner[idx.ids[idx.list_off[l]:idx.list_off[l + 1]]] = l
    allowed = (owner[None, :] == probes[:, :, None]).any(1)
    GD, GI = brute_ip(xb, xq, 10, allowed)
    np.testing.assert_array_equal(I, GI)


@pytest.mark.parametrize("nprobe", [2, 6, 40])
def test_triangle_pruning_is_lossless_on_unit_vectors(ipdata, nprobe):
    xb, xq, idx = ipdata
    D0, I0, _ = O.ip_search(idx, xq, 10, nprobe=nprobe, triangle=False)
    D1, I1, _, skipped = O.ip_search(idx, xq, 10, nprobe=nprobe, triangle=True, return_skipped=True)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    if nprobe > 2:
        assert skipped > 0  # the band does cut


def test_triangle_band_is_undefined_for_long_queries(ipdata):
    """|q| = 3: <q, c> exceeds 1 for near centroids, sqrt(1 - c^2) is NaN and no cut happens."""
    xb, xq, idx = ipdata
    D0, I0, _ = O.ip_search(idx, 3 * xq, 10, nprobe=6, triangle=False)
    D1, I1, _, skipped = O.ip_search(idx, 3 * xq, 10, nprobe=6, triangle=True, return_skipped=True)
    np.testing.assert_array_equal(I1, I0)


def test_ip_assign_argmax_and_ties():
    c = np.array([[1, 0], [0, 1], [1, 0]], np.float32)  # rows 0 and 2 tie
    x = np.array([[2, 1], [0.5, 3], [-1, -1]], np.float32)
    lab, best = O.ip_assign(x, c)
    np.testing.assert_array_equal(lab, [0, 1, 0])
    np.testing.assert_array_equal(best, [2, 3, -1])


def test_ip_lists_order(ipdata):
    xb, xq, idx = ipdata
    for l in range(idx.nlist):
        a, b = idx.list_off[l], idx.list_off[l + 1]
        c2c = idx.c2c[a:b]
        ids = idx.ids[a:b]
        assert np.all((np.diff(c2c) < 0) | ((np.diff(c2c) == 0) & (np.diff(ids) > 0)))
        np.testing.assert_allclose(c2c, xb[ids].astype(np.float64) @ idx.centroids[l].astype(np.float64), rtol=1e-5,
                                   atol=1e-6)
    lab, _ = O.ip_assign(xb, idx.centroids)
    for l in range(idx.nlist):
        assert np.all(lab[idx.ids[idx.list_off[l]:idx.list_off[l + 1]]] == l)


def test_ip_kmeans_step_and_normalization():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(400, 8)).astype(np.float32)
    c1 = O.ip_kmeans(x, 5, niter=1, max_pts=0, seed=3)
    np.testing.assert_allclose(np.linalg.norm(c1.astype(np.float64), axis=1), 1.0, atol=1e-6)
    # the same step by hand: init rows of the seeded shuffle, assignment by normalized rows
    sample, c0 = O.kmeans_init(x, 5, max_pts=0, seed=3)
    lab, _ = O.ip_assign(O.ip_normalize_rows(sample), c0)
    means = np.stack([sample[lab == c].astype(np.float64).mean(0) if np.any(lab == c) else c0[c] for c in range(5)])
    np.testing.assert_allclose(c1, O.ip_normalize_centroids(means.astype(np.float32)), rtol=1e-6, atol=1e-7)


def test_ip_normalize_rows():
    x = np.array([[3, 4], [0, 0]], np.float32)
    np.testing.assert_array_equal(O.ip_normalize_rows(x), np.array([[0.6, 0.8], [0, 0]], np.float32))
