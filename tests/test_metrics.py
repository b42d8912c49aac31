"""recall@k and r2 against hand-computed values (calculate_recall, calculate_r2; utils.h:660-756).

Both restatements are checked: the oracle's (or_recall / or_r2, the checker) and the product's
harmony_amd.metrics (what bench.py reports). The cases are built so that a plausible slip fails
one of them: de-duplicating repeated ids, skipping past a -1 instead of stopping, taking the
ground-truth set from all gt_k columns instead of the first k, dropping the D[0] < 0 rule,
accepting k > gt_k, pairwise instead of sequential FP32 sums in r2.
"""
import numpy as np
import pytest

from harmony_amd import metrics
from oracle import oracle as O

# 3 queries, k = 4, gt_k = 6
I3 = np.array([[1, 2, 3, 4],      # 4 hits
               [5, -1, 10, 11],   # 1 hit: the -1 ends the row (10, 11 are in the GT set)
               [9, 9, 8, 30]],    # 9 twice + 8 = 3 hits (30 sits in GT column 5, outside the first k)
              dtype=np.int64)
GT3 = np.array([[4, 3, 2, 1, 50, 51],
                [5, 10, 11, 12, 52, 53],
                [9, 8, 13, 14, 15, 30]], dtype=np.int64)
HITS3 = 4 + 1 + 3


def both_recall(I, GT, D=None, GD=None, **kw):
    return O.recall(I, GT, D, GD, **kw), metrics.recall(I, GT, D, GD, **kw)


def test_recall_closed_form():
    want = float(np.float32(HITS3) / np.float32(12))
    for r in both_recall(I3, GT3):
        assert r == want, r
    # without the duplicate rule (np.intersect1d) it would read 7/12; with all gt_k columns, 9/12
    assert want != float(np.float32(7) / np.float32(12))


def test_recall_per_query_closed_form():
    np.testing.assert_array_equal(metrics.recall_per_query(I3, GT3), [1.0, 0.25, 0.75])


def test_recall_negative_first_distance_is_zero():
    D = np.ones((3, 4), np.float32)
    D[0, 0] = -1.0
    for r in both_recall(I3, GT3, D=D, GD=np.ones((3, 6), np.float32)):
        assert r == 0.0
    D[0, 0] = 0.0  # zero is not negative
    for r in both_recall(I3, GT3, D=D, GD=np.ones((3, 6), np.float32)):
        assert r > 0.0


def test_recall_k_above_gt_k_raises():
    with pytest.raises(ValueError):
        O.recall(I3, GT3[:, :3])
    with pytest.raises(ValueError):
        metrics.recall(I3, GT3[:, :3])


def test_recall_distance_sanity_count():
    # one query, k = 4; GT row ends in -1 so its last valid distance (topK) is GD[0, 2] = 4
    I = np.array([[7, 8, 9, 6]], np.int64)
    GT = np.array([[7, 8, 6, -1]], np.int64)
    GD = np.array([[1, 2, 4, np.finfo(np.float32).max]], np.float32)
    # 1 <= 4, 2 <= 4, 4.0001 within 1e-4 relative of 4, 5 is not (|5 - 4| / 5 = 0.2)
    D = np.array([[1.0, 2.0, 4.0001, 5.0]], np.float32)
    for r, correct in (O.recall(I, GT, D, GD, return_correct=True),
                       metrics.recall(I, GT, D, GD, return_correct=True)):
        assert correct == 3
        assert r == 0.75  # 7, 8, 6 are hits
    # inner product: "correct" means >= topK
    Dip = np.array([[5.0, 4.0, 3.99999, 1.0]], np.float32)
    for _, correct in (O.recall(I, GT, Dip, GD, metric_l2=False, return_correct=True),
                       metrics.recall(I, GT, Dip, GD, metric_l2=False, return_correct=True)):
        assert correct == 3


def test_recall_random_agrees():
    rng = np.random.default_rng(7)
    I = rng.integers(-1, 40, size=(50, 10))
    GT = rng.integers(-1, 40, size=(50, 16))
    D = rng.random((50, 10), dtype=np.float32)
    GD = rng.random((50, 16), dtype=np.float32)
    a = O.recall(I, GT, D, GD, return_correct=True)
    b = metrics.recall(I, GT, D, GD, return_correct=True)
    assert a == b


def test_r2_closed_form():
    I = np.array([[1, 2, -1], [3, 4, 5]], np.int64)
    D = np.array([[1, 2, 9], [3, 4, 5]], np.float32)
    GD = np.array([[1, 1, 1, 7], [2, 2, 2, 7]], np.float32)  # gt_k = 4 > k = 3
    # sum D over entries before the -1: 1 + 2 + 3 + 4 + 5 = 15; GD over the same slots: 2 + 6 = 8
    for r in (O.r2(I, D, GD), metrics.r2(I, D, GD)):
        assert r == 15.0 / 8.0 - 1.0
    with pytest.raises(ValueError):
        O.r2(I, D, GD[:, :2])
    with pytest.raises(ValueError):
        metrics.r2(I, D, GD[:, :2])


def test_r2_sequential_fp32_sums():
    # 2^24 followed by ones: a sequential FP32 sum never moves (1 is half an ulp, ties to even),
    # a pairwise or FP64 sum does
    k = 64
    D = np.ones((1, k), np.float32)
    D[0, 0] = 2.0 ** 24
    GD = np.full((1, k), 2.0 ** 24 / k, np.float32)
    I = np.arange(k, dtype=np.int64)[None, :]
    want = float(np.float32(2.0 ** 24) / np.float32(2.0 ** 24) - np.float32(1))
    for r in (O.r2(I, D, GD), metrics.r2(I, D, GD)):
        assert r == want == 0.0
