"""Workload 1 golden vectors (tests/golden/w1_golden.npz, written by make_golden.py): the data
generator reproduces the same arrays, and the oracle reproduces the stored outputs."""
import os

import numpy as np
import pytest

from harmony_amd import datagen
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "w1_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


def test_generator_reproduces_workload1(gold):
    xb, xq = datagen.make(1)
    assert datagen.sha256(xb) == str(gold["base_sha256"])
    assert datagen.sha256(xq) == str(gold["query_sha256"])


def test_oracle_reproduces_golden(gold):
    xb, xq = datagen.make(1)
    GD, GI = O.bruteforce_f64(xb, xq, 10)
    np.testing.assert_array_equal(GI, gold["gt_I"])
    np.testing.assert_array_equal(GD, gold["gt_D"])
    idx = O.build_index(xb, 100, niter=10, seed=0)
    assert datagen.sha256(idx.centroids) == str(gold["centroid_sha256"])
    D, I, probes, bound = O.search(idx, xq, 10, nprobe=10, dim_blocks=4, pruning=True)
    np.testing.assert_array_equal(I, gold["I"])
    np.testing.assert_array_equal(D, gold["D"])
    np.testing.assert_array_equal(probes, gold["probes"])
    np.testing.assert_array_equal(bound, gold["bound"])


def test_golden_recall_is_meaningful(gold):
    # the IVF result at nprobe 10 of 100 lists finds most of the exact neighbours
    r = O.recall(gold["I"], gold["gt_I"])
    assert 0.9 <= r <= 1.0
