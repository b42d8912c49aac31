"""The BF16 form's nearest-list pass: each query's nearest list scanned with the query's candidates
kept in LDS and reduced to the k-th smallest upper bound inside the same kernel (nearest_pass=1,
the default), against the oracle and against the two-kernel path it replaces (nearest_pass=0:
every pair of the nearest list written out, then compacted). The shapes exercise its branches:
lists of thousands of vectors (several in-kernel compactions per query), more than 16 queries per
list (several work items per list), lists shorter than k, per-query k, float data with d = 420
(the kernel's unpipelined form), massive ties that keep a query's candidates above the buffer's
headroom (the query spills to its survivor buffer and the compaction after the pass finishes it),
and the learned search's fused pass (every searched rank in one launch, each item appending and
lowering its query's bound)."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP32, datagen, learned
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FMAX = np.finfo(np.float32).max


def say(*a):
    print("[gpu-test nearest]", *a, flush=True)
    sys.stdout.flush()


def clustered(nb, nq, d, centers, seed, spread=12.0):
    """Integer-valued clipped Gaussian mixture (rows and queries from the same centers)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 128, size=(centers, d))
    rows = c[rng.integers(0, centers, size=nb)] + rng.normal(0, spread, (nb, d))
    qs = c[rng.integers(0, centers, size=nq)] + rng.normal(0, spread, (nq, d))
    f = lambda a: np.clip(np.rint(a), 0, 255).astype(np.float32)  # noqa: E731
    return f(rows), f(qs)


def build(xb, nlist, niter=6):
    g = Index(xb.shape[1], nlist, nprobe=1)
    g.train(xb, niter=niter, seed=0, max_points_per_centroid=256)
    g.add(xb)
    return g, O.IndexData.from_gpu_index(g)


def both_passes(g, xq, k, **kw):
    out = {}
    for near in (1, 0):
        g.preSearch(dim_blocks=4, precision=HV_PRECISION_BF16, nearest_pass=near, **kw)
        out[near] = g.search(xq, k, return_stats=True)
    g.preSearch(dim_blocks=4, precision=HV_PRECISION_FP32, nearest_pass=1)
    return out


@pytest.fixture(scope="module")
def long_lists():
    # 40,000 x 64 from 8 centers into 16 lists: lists of about 2,500 vectors (cut into items of
    # 1,024, several compactions per query); 400 queries: about 25 per list (two query tiles)
    say("40,000 x 64, 8 centers, nlist 16, 400 queries")
    xb, xq = clustered(40000, 400, 64, 8, seed=3)
    g, o = build(xb, 16)
    sizes = np.diff(g.list_offsets())
    assert sizes.max() > 1500
    return g, o, xq


@pytest.mark.parametrize("k,nprobe", [(100, 1), (100, 3), (10, 1), (10, 4), (128, 2), (1, 1)])
def test_long_lists_match_oracle(long_lists, k, nprobe):
    g, o, xq = long_lists
    g.nprobe = nprobe
    say("long lists k=%d nprobe=%d" % (k, nprobe))
    res = both_passes(g, xq, k)
    oD, oI, _, _ = O.search(o, xq, k, nprobe=nprobe, dim_blocks=4)
    for near, (D, I, st) in res.items():
        np.testing.assert_array_equal(I, oI, err_msg="nearest_pass=%d" % near)
        np.testing.assert_array_equal(D, oD, err_msg="nearest_pass=%d" % near)
    s1, s0 = res[1][2], res[0][2]
    say("survivors/query: pass %.1f, two kernels %.1f; overflowed %d / %d"
        % (s1["survivors"] / len(xq), s0["survivors"] / len(xq), s1["overflow_queries"], s0["overflow_queries"]))
    if s1["overflow_queries"] == 0 and s0["overflow_queries"] == 0:
        # the pass leaves {lb <= min(bound, k-th upper bound over the list)} (its upper bounds are
        # re-formed from an 8-byte entry, at most a BF16 rounding of the width above the pair's);
        # the two-kernel path the same for lists that fit its direct slots (4096), and for longer
        # ones keeps the tail's pairs against the k-th over the first 4096 only
        assert s1["survivors"] <= 1.02 * s0["survivors"]
        assert s1["survivors"] >= 0.98 * s0["survivors"] or np.diff(g.list_offsets()).max() > 4096
    if k <= 100:
        assert s1["overflow_queries"] == 0


def test_per_query_k(long_lists):
    g, o, xq = long_lists
    g.nprobe = 2
    kq = np.where(np.arange(len(xq)) % 3 == 0, 10, 100).astype(np.int32)
    g.set_query_k(kq)
    try:
        res = both_passes(g, xq, 100)
    finally:
        g.set_query_k(None)
    for kk in (10, 100):
        oD, oI, _, _ = O.search(o, xq, kk, nprobe=2, dim_blocks=4)
        rows = np.nonzero(kq == kk)[0]
        for near, (D, I, _) in res.items():
            np.testing.assert_array_equal(I[rows, :kk], oI[rows], err_msg="nearest_pass=%d k=%d" % (near, kk))
            np.testing.assert_array_equal(D[rows, :kk], oD[rows], err_msg="nearest_pass=%d k=%d" % (near, kk))
            assert np.all(I[rows, kk:] == -1) and np.all(D[rows, kk:] == FMAX)


def test_short_lists_and_large_k():
    """Lists shorter than k (the bound stays where it started; every pair is kept)."""
    xb, xq = clustered(3000, 50, 32, 20, seed=5)
    g, o = build(xb, 60, niter=4)
    g.nprobe = 2
    res = both_passes(g, xq, 100)
    oD, oI, _, _ = O.search(o, xq, 100, nprobe=2, dim_blocks=4)
    for near, (D, I, _) in res.items():
        np.testing.assert_array_equal(I, oI, err_msg="nearest_pass=%d" % near)
        np.testing.assert_array_equal(D, oD, err_msg="nearest_pass=%d" % near)


def test_ties_spill_to_survivor_buffer():
    """Thirty distinct rows repeated 200 times: a query's candidates tie in groups of 200, more
    than the LDS buffer keeps after a compaction, so the query spills to its survivor buffer and
    the compaction after the pass finishes it; ties are broken by id either way."""
    rng = np.random.default_rng(9)
    pts = rng.integers(0, 64, size=(30, 32)).astype(np.float32)
    xb = np.repeat(pts, 200, axis=0)
    xb = xb[rng.permutation(len(xb))]
    xq = (pts[:12] + rng.integers(-2, 3, size=(12, 32))).astype(np.float32)
    g, o = build(xb, 3, niter=3)
    g.nprobe = 1
    res = both_passes(g, xq, 100)
    oD, oI, _, _ = O.search(o, xq, 100, nprobe=1, dim_blocks=4)
    for near, (D, I, _) in res.items():
        np.testing.assert_array_equal(I, oI, err_msg="nearest_pass=%d" % near)
        np.testing.assert_array_equal(D, oD, err_msg="nearest_pass=%d" % near)
    s1, s0 = res[1][2], res[0][2]
    assert s1["overflow_queries"] == 0 and s0["overflow_queries"] == 0
    assert s1["survivors"] >= 200 * len(xq) and s0["survivors"] >= 200 * len(xq)


@pytest.mark.parametrize("nq,nlist,nprobe", [(60, 12, 3), (400, 4, 2)])
def test_float_wide_dims(nq, nlist, nprobe):
    """Float data, d = 420 (14 MFMA steps: the kernel's form without code prefetch), variance
    order, 8 blocks. 400 queries over 4 lists at nprobe 2: about 200 queries per list, so most
    items hold four query blocks and the scan stages their code blocks in LDS."""
    say("d=420: %d queries, nlist %d, nprobe %d" % (nq, nlist, nprobe))
    xb, xq = datagen.make(3, nb=6000, nq=nq)
    perm = datagen.variance_order(xb)
    g = Index(420, nlist, nprobe=nprobe, dim_order=perm)
    g.train(xb, niter=4, seed=0, max_points_per_centroid=256)
    g.add(xb)
    o = O.IndexData.from_gpu_index(g)
    o = O.IndexData(o.centroids[:, perm], o.list_off, o.ids, o.codes[:, perm])
    out = {}
    for near in (1, 0):
        g.preSearch(dim_blocks=8, precision=HV_PRECISION_BF16, nearest_pass=near)
        out[near] = g.search(xq, 50)
    oD, oI, _, _ = O.search(o, xq[:, perm], 50, nprobe=nprobe, dim_blocks=8)
    for near, (D, I) in out.items():
        np.testing.assert_array_equal(I, oI, err_msg="nearest_pass=%d" % near)
        np.testing.assert_array_equal(D, oD, err_msg="nearest_pass=%d" % near)


def test_fused_learned_pass_matches_oracle(long_lists):
    """The learned BF16 search runs its searched ranks [0, 4) in one nearest-list pass: items of
    several ranks of one query (and the parts of its split lists, about 2,500 vectors each) run in
    one launch, each keeping what its own k-th upper bound allows and lowering the query's bound
    by an atomic minimum. With every query at the cap the results are the oracle's at nprobe 4,
    with the predictor's seed, a seed far too low (every query redone exactly) and no seed."""
    g, o, xq = long_lists
    g.nprobe = 4
    say("fused learned pass: traces and a short training on the 400 queries")
    feats, targets, _ = learned.collect_traces(g, xq, k=100)
    learned.train(g, feats, targets, epochs=5, batch=128, lr=3e-3, seed=0)
    oD, oI, _, _ = O.search(o, xq, 100, nprobe=4, dim_blocks=4)
    try:
        for bscale in (1.0, 1e-6, 0.0):
            g.preSearch(learned=True, learned_nprobe_scale=1e6, learned_bound_scale=bscale,
                        precision=HV_PRECISION_BF16, dim_blocks=4)
            D, I, st = g.search(xq, 100, return_stats=True)
            say("bound scale %g: survivors/query %.1f, overflowed %d, redone %d"
                % (bscale, st["survivors"] / len(xq), st["overflow_queries"], st["fallback_queries"]))
            np.testing.assert_array_equal(I, oI, err_msg="bound scale %g" % bscale)
            np.testing.assert_array_equal(D, oD, err_msg="bound scale %g" % bscale)
            assert np.all(g.last_nprobe(len(xq)) == 4)
            assert st["overflow_queries"] == 0
            if bscale == 1e-6:
                assert st["fallback_queries"] == len(xq)
    finally:
        g.preSearch(learned=False, precision=HV_PRECISION_FP32)
