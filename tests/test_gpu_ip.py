"""The inner-product metric on the GPU against the oracle (Index.cpp:408-418, 567-671,
1440-1551; Clustering.cpp:121-151): k-means centroids, list membership, list order and
candidate2centroid values bit for bit; search results (ids and inner products) bit for bit at
several nprobe and k, with and without triangle pruning; padding when the probed lists hold fewer
than k vectors; duplicate vectors ordered by id; long queries (the band undefined); persistence in
the reference's index format; and the forms the IP search does not take are refused."""
import numpy as np
import pytest

from harmony_amd import _lib
from harmony_amd._lib import HV_METRIC_IP, HV_OPT_NONE, HV_OPT_TRIANGLE, HvError
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def say(msg):
    print("[ip] " + msg, flush=True)


def unit(x):
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


@pytest.fixture(scope="module")
def ipdata():
    rng = np.random.default_rng(5)
    d, nb, nq, nlist = 64, 12000, 200, 48
    centers = unit(rng.normal(size=(40, d)))
    xb = unit(centers[rng.integers(0, 40, size=nb)] + 0.08 * rng.normal(size=(nb, d)).astype(np.float32))
    xb[100:110] = xb[7]  # duplicates: equal inner products, ordered by id
    xq = unit(centers[rng.integers(0, 40, size=nq)] + 0.08 * rng.normal(size=(nq, d)).astype(np.float32))
    say("train/add %d x %d, nlist %d" % (nb, d, nlist))
    g = Index(d, nlist, nprobe=4, metric=HV_METRIC_IP, opt_level=HV_OPT_TRIANGLE)
    g.train(xb, niter=8, seed=0)
    g.add(xb)
    o = O.IpIndexData(g.centroids(), g.list_offsets(), g.list_ids(), g.list_codes(), g.list_c2c())
    return xb, xq, g, o


def test_train_matches_oracle(ipdata):
    xb, _, g, _ = ipdata
    c = O.ip_kmeans(xb, g.nlist, niter=8, max_pts=256, seed=0)
    np.testing.assert_array_equal(g.centroids(), c)
    np.testing.assert_allclose(np.linalg.norm(c.astype(np.float64), axis=1), 1.0, atol=1e-6)


def test_lists_match_oracle(ipdata):
    xb, _, g, o = ipdata
    lab, best = O.ip_assign(xb, g.centroids())
    off, order = O.ip_build_lists(lab, best, g.nlist)
    np.testing.assert_array_equal(g.list_offsets(), off)
    np.testing.assert_array_equal(g.list_ids(), order)
    np.testing.assert_array_equal(g.list_c2c(), best[order])
    np.testing.assert_array_equal(g.list_codes(), xb[order])


@pytest.mark.parametrize("triangle", [False, True])
@pytest.mark.parametrize("nprobe,k", [(1, 10), (4, 10), (4, 100), (16, 10), (48, 100)])
def test_search_matches_oracle(ipdata, nprobe, k, triangle):
    _, xq, g, o = ipdata
    say("search nprobe=%d k=%d triangle=%s" % (nprobe, k, triangle))
    g.nprobe = nprobe
    g.opt_level = HV_OPT_TRIANGLE if triangle else HV_OPT_NONE
    D, I = g.search(xq, k)
    oD, oI, oP = O.ip_search(o, xq, k, nprobe=nprobe, triangle=triangle)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    np.testing.assert_array_equal(g.last_probes(), oP)
    assert np.all(np.diff(D, axis=1) <= 0)


def test_triangle_is_lossless_and_full_probe_is_brute_force(ipdata):
    xb, xq, g, _ = ipdata
    g.nprobe = g.nlist
    g.opt_level = HV_OPT_TRIANGLE
    D1, I1 = g.search(xq, 20)
    g.opt_level = HV_OPT_NONE
    D0, I0 = g.search(xq, 20)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    s = xq.astype(np.float64) @ xb.astype(np.float64).T
    GI = np.argsort(-s, axis=1, kind="stable")[:, :20]
    # FP32 chains against FP64: ids agree wherever the FP64 gap to the next is clear of rounding
    gap = np.abs(np.diff(np.take_along_axis(s, GI, 1), axis=1)) > 1e-5
    ok = np.concatenate([gap, np.ones((len(xq), 1), bool)], 1) & np.concatenate([np.ones((len(xq), 1), bool), gap], 1)
    assert np.mean(I0[ok] == GI[ok]) == 1.0


@pytest.mark.parametrize("nprobe", [1, 8, 48])
def test_triangle_computes_fewer_inner_products(ipdata, nprobe):
    """Without the cut every probed vector is scored once: the sum of the probed list sizes. With
    it (nprobe > 1) the nearest list is scored whole and the others only over their band."""
    _, xq, g, _ = ipdata
    g.nprobe = nprobe
    g.opt_level = HV_OPT_NONE
    D0, I0, s0 = g.search(xq, 10, return_stats=True)
    sizes = np.diff(g.list_offsets())
    assert s0["scan_pairs"] == int(sizes[g.last_probes()].sum())
    g.opt_level = HV_OPT_TRIANGLE
    D1, I1, s1 = g.search(xq, 10, return_stats=True)
    np.testing.assert_array_equal(I1, I0)
    say("nprobe %d: %d -> %d inner products" % (nprobe, s0["scan_pairs"], s1["scan_pairs"]))
    if nprobe == 1:
        assert s1["scan_pairs"] == s0["scan_pairs"]
    else:
        assert s1["scan_pairs"] < s0["scan_pairs"]
        assert s1["scan_pairs"] >= int(sizes[g.last_probes()[:, 0]].sum())


def test_duplicates_are_ordered_by_id(ipdata):
    xb, _, g, o = ipdata
    g.nprobe = 4
    g.opt_level = HV_OPT_TRIANGLE
    D, I = g.search(xb[7:8], 12)
    assert list(I[0, :11]) == [7] + list(range(100, 110))
    oD, oI, _ = O.ip_search(o, xb[7:8], 12, nprobe=4, triangle=True)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_long_queries_leave_the_band_undefined(ipdata):
    _, xq, g, o = ipdata
    g.nprobe = 6
    g.opt_level = HV_OPT_TRIANGLE
    D, I = g.search(3 * xq, 10)
    oD, oI, _ = O.ip_search(o, 3 * xq, 10, nprobe=6, triangle=True)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)


def test_padding_when_lists_hold_fewer_than_k():
    rng = np.random.default_rng(3)
    xb = unit(rng.normal(size=(300, 16)))
    g = Index(16, 30, nprobe=1, metric=HV_METRIC_IP, opt_level=HV_OPT_TRIANGLE)
    g.train(xb, niter=4, seed=1)
    g.add(xb)
    o = O.IpIndexData(g.centroids(), g.list_offsets(), g.list_ids(), g.list_codes(), g.list_c2c())
    xq = unit(rng.normal(size=(17, 16)))
    for triangle in (False, True):
        g.opt_level = HV_OPT_TRIANGLE if triangle else HV_OPT_NONE
        D, I = g.search(xq, 64)
        oD, oI, _ = O.ip_search(o, xq, 64, nprobe=1, triangle=triangle)
        np.testing.assert_array_equal(I, oI)
        np.testing.assert_array_equal(D, oD)
        assert np.any(I == -1) and np.all(D[I == -1] == -FMAX)


def test_persistence_round_trip(ipdata, tmp_path):
    _, xq, g, o = ipdata
    p = str(tmp_path / "ip.index")
    g.save_index(p)
    h = Index(g.d, g.nlist, metric=HV_METRIC_IP)
    h.load_index(p)
    np.testing.assert_array_equal(h.list_c2c(), g.list_c2c())
    np.testing.assert_array_equal(h.centroids(), g.centroids())
    for idx in (g, h):
        idx.nprobe = 8
        idx.opt_level = HV_OPT_TRIANGLE
    D0, I0 = g.search(xq, 10)
    D1, I1 = h.search(xq, 10)
    np.testing.assert_array_equal(I1, I0)
    np.testing.assert_array_equal(D1, D0)
    # an L2 handle takes the file's metric
    m = Index(g.d, g.nlist)
    m.load_index(p)
    m.nprobe = 8
    m.opt_level = HV_OPT_TRIANGLE
    D2, I2 = m.search(xq, 10)
    np.testing.assert_array_equal(I2, I0)


def test_unsupported_forms_are_refused(ipdata):
    _, xq, g, _ = ipdata
    g.nprobe = 4
    g.opt_level = HV_OPT_NONE
    g.preSearch(precision=_lib.HV_PRECISION_BF16)
    with pytest.raises(HvError):
        g.search(xq, 10)
    g.preSearch(precision=_lib.HV_PRECISION_FP32)
    with pytest.raises(HvError):
        g.ground_truth_f64(xq, 10)
    g.set_query_k(np.full(len(xq), 5, np.int32))
    with pytest.raises(HvError):
        g.search(xq, 10)
    g.set_query_k(None)
    D, I = g.search(xq, 10)  # the handle still works
    assert I.shape == (len(xq), 10)
