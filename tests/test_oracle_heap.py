"""Pins the oracle's heap restatement (heap.hpp:13-112) against the reference's own heap.hpp,
compiled from its source into oracle/_ref, and the canonical (distance, id) heap against a
full sort. Mirrors the invariant of src/tests/heap_test.cpp:7-91 with seeded data and ties."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run_oracle_heap(vals, k, tie):
    L = O.lib()
    v = np.empty(k, np.float32)
    ids = np.empty(k, np.int64)
    L.or_heap_init(k, _p(v), _p(ids))
    for i, x in enumerate(vals):
        L.or_heap_push(tie, k, _p(v), _p(ids), float(x), i)
    L.or_heap_sort(tie, k, _p(v), _p(ids))
    return v, ids


def run_ref_heap(vals, k):
    R = O.ref_heap()
    v = np.empty(k, np.float32)
    ids = np.empty(k, np.int64)
    R.ref_init_result(k, _p(v), _p(ids))
    for i, x in enumerate(vals):
        if np.float32(x) < v[0]:  # the L2 callers' insert guard (IVFScan.hpp:267, node.cpp:654)
            R.ref_heap_replace_top(k, _p(v), _p(ids), float(x), i)
    R.ref_sort_result(k, _p(v), _p(ids))
    return v, ids


@pytest.mark.parametrize("seed,n,k,levels", [(0, 100, 10, 1000), (1, 500, 37, 20), (2, 60, 100, 5),
                                             (3, 1000, 1, 50), (4, 7, 16, 3)])
def test_verbatim_heap_matches_reference_source(seed, n, k, levels):
    if O.ref_heap() is None:
        pytest.skip("reference heap not built (oracle/_ref absent)")
    vals = np.random.default_rng(seed).integers(0, levels, size=n).astype(np.float32)
    ov, oi = run_oracle_heap(vals, k, O.TIE_VERBATIM)
    rv, ri = run_ref_heap(vals, k)
    np.testing.assert_array_equal(ov, rv)
    np.testing.assert_array_equal(oi, ri)  # ties included: same sift order, same survivors


@pytest.mark.parametrize("seed,n,k,levels", [(0, 100, 10, 1000), (1, 500, 37, 20), (2, 60, 100, 5),
                                             (5, 3000, 64, 7)])
def test_canonical_heap_is_sorted_prefix(seed, n, k, levels):
    vals = np.random.default_rng(seed).integers(0, levels, size=n).astype(np.float32)
    v, ids = run_oracle_heap(vals, k, O.TIE_CANONICAL)
    order = sorted(range(n), key=lambda i: (vals[i], i))[:k]
    m = len(order)
    np.testing.assert_array_equal(ids[:m], order)
    np.testing.assert_array_equal(v[:m], vals[order])
    assert np.all(ids[m:] == -1) and np.all(v[m:] == np.finfo(np.float32).max)


def test_verbatim_and_canonical_agree_without_ties():
    vals = np.random.default_rng(9).permutation(5000).astype(np.float32)
    a = run_oracle_heap(vals, 50, O.TIE_VERBATIM)
    b = run_oracle_heap(vals, 50, O.TIE_CANONICAL)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_verbatim_keeps_same_distance_set_under_ties():
    vals = np.random.default_rng(11).integers(0, 4, size=300).astype(np.float32)
    a = run_oracle_heap(vals, 40, O.TIE_VERBATIM)
    b = run_oracle_heap(vals, 40, O.TIE_CANONICAL)
    np.testing.assert_array_equal(a[0], b[0])  # distances equal; ids may differ among equals
