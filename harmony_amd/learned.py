"""Learned probe/bound predictor: traces from the plain search, GPU training, calibration.

The predictor (a 65 -> 128 -> 2 MLP run by the library's own HIP kernels) reads a query's
64 smallest centroid distances (the coarse step's upper bounds of them) and its norm, and predicts the k-th neighbour distance and
the number of lists needed to reach a target recall. Serving uses the prediction to trim each
query's probe list and to seed its pruning bound. The margin test widens a query's probe count to
every list (of its first nprobe) that the triangle inequality cannot rule out within the seeded
bound, and a query whose k-th result lands above its seed is redone exactly over all nprobe lists
(the plain path), so with the default margin every result is the plain search's at nprobe. No
counterpart in the reference, which probes a fixed nprobe for every query and takes its bound
from a warm-up scan.

Labels come from the plain search itself: a search over every list gives each query's exact
top-k (the plain path's arithmetic), hence the k-th distance; the minimal nprobe is the smallest
p such that at least ceil(target * k) of those neighbours lie in the query's p nearest lists
(a search over those p lists returns all of them: they are the best candidates it sees).
"""
import math
import time

import numpy as np

from . import _lib, metrics
from ._lib import HV_PRECISION_BF16
from .index import DeviceArray


def id_to_list(index):
    """list number of every vector id."""
    off = index.list_offsets()
    ids = index.list_ids()
    out = np.empty(int(ids.max()) + 1 if ids.size else 0, dtype=np.int32)
    for l in range(len(off) - 1):
        out[ids[off[l]:off[l + 1]]] = l
    return out


def search_dev(index, q, k):
    """(D, I) of a search with the queries and results in device memory (hv_search_device)."""
    q = np.ascontiguousarray(q, dtype=np.float32)
    n = q.shape[0]
    qd, Dd, Id = DeviceArray.from_host(q), DeviceArray(n * k * 4), DeviceArray(n * k * 8)
    index.search_device(qd, n, k, Dd, Id)
    return Dd.to_host(np.float32, (n, k)), Id.to_host(np.int64, (n, k))


def collect_traces(index, queries, k, target_recall=0.95, chunk=10000):
    """(features [n][65], targets [n][2], exact ids [n][k]) of the plain search: targets are
    log1p(k-th distance) and log(minimal nprobe for target_recall)."""
    q = np.ascontiguousarray(queries, dtype=np.float32)
    n = q.shape[0]
    nlist = index.nlist
    saved = (index.nprobe, index._params.learned, index._params.precision)
    list_of = id_to_list(index)
    need = int(math.ceil(target_recall * k))
    feats = np.empty((n, _lib.HV_MLP_IN), np.float32)
    targets = np.empty((n, 2), np.float32)
    exact = np.empty((n, k), np.int64)
    try:
        index.nprobe = nlist
        index.preSearch(learned=False, precision=HV_PRECISION_BF16)
        for c0 in range(0, n, chunk):
            qc = q[c0:c0 + chunk]
            m = qc.shape[0]
            feats[c0:c0 + m] = index.query_features(qc)
            D, I = search_dev(index, qc, k)
            probes = index.last_probes(m)[:, :nlist]
            rank = np.empty((m, nlist), np.int32)
            np.put_along_axis(rank, probes, np.arange(nlist, dtype=np.int32)[None, :].repeat(m, 0), axis=1)
            valid = I >= 0
            lists = list_of[np.where(valid, I, 0)]
            r = np.where(valid, np.take_along_axis(rank, lists, axis=1), nlist)
            r.sort(axis=1)
            nmin = np.minimum(r[:, need - 1] + 1, nlist)
            targets[c0:c0 + m, 0] = np.log1p(np.maximum(D[:, k - 1], 0.0))
            targets[c0:c0 + m, 1] = np.log(nmin.astype(np.float64))
            exact[c0:c0 + m] = I
    finally:
        index.nprobe = saved[0]
        index.preSearch(learned=bool(saved[1]), precision=saved[2])
    return feats, targets, exact


def standardization(feats, targets):
    """(fmean, fstd, tmean, tstd) of a training set (std floored so constant columns pass)."""
    f = np.asarray(feats, np.float64)
    t = np.asarray(targets, np.float64)
    return (f.mean(0), np.maximum(f.std(0), 1e-3), t.mean(0), np.maximum(t.std(0), 1e-3))


def train(index, feats, targets, epochs=30, batch=1024, lr=2e-3, seed=0):
    """Fresh network on the index, trained on the GPU; returns per-epoch losses."""
    index.mlp_set_norm(*standardization(feats, targets))
    index.mlp_init(seed)
    return index.mlp_train(feats, targets, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)


def recall_per_query(I, exact):
    """Each query's hits / k against its exact ids (calculate_recall's counting, utils.h:676-685)."""
    return metrics.recall_per_query(I, exact)


def calibrate(index, feats, targets, queries, exact, nprobe, k, target_recall=0.95, margin=0.01,
              bound_quantile=1.0, bound_margins=(1.25, 1.5, 2.0),
              scales=(0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0, 2.5, 3.0, 4.0),
              widen_margins=(0.0, 0.5, 1.0), reps=5):
    """Serving settings from held-out queries. A bound scale is a margin times the bound_quantile
    of true / predicted k-th distance, so that nearly every query keeps its seeded bound (a query
    whose seed is below its k-th distance is redone exactly over all nprobe lists, while a looser
    seed costs little: the compaction after each scan phase lowers the bound to the k-th upper
    bound anyway). For each bound margin and each widening margin (0: the predicted counts alone;
    1: counts widened to every list the triangle and Voronoi bounds cannot exclude within the
    seed, so results equal the plain search's at nprobe), the smallest of `scales` whose mean
    recall against the exact ids reaches target_recall + margin qualifies; the qualifying setting
    with the shortest measured search time on these queries (the mean of `reps` searches after
    one) is chosen. Leaves the index in learned mode with it; returns (nprobe_scale, bound_scale,
    recall, mean probes); the chosen widening margin is in index._params.learned_margin."""
    pred = index.mlp_predict(feats)
    ratio = np.expm1(targets[:, 0].astype(np.float64)) / np.maximum(np.expm1(pred[:, 0].astype(np.float64)), 1e-30)
    q_ratio = float(np.quantile(ratio, bound_quantile))
    index.nprobe = nprobe
    q = np.ascontiguousarray(queries, dtype=np.float32)
    n = q.shape[0]
    qd, Dd, Id = DeviceArray.from_host(q), DeviceArray(n * k * 4), DeviceArray(n * k * 8)

    def apply(s, b, wm):
        index.preSearch(learned=True, learned_nprobe_scale=s, learned_bound_scale=b, learned_margin=wm)

    qualified, fallback = [], None
    for bm in bound_margins:
        bound_scale = float(max(1.0, bm * q_ratio))
        for wm in widen_margins:
            for s in scales:
                apply(s, bound_scale, wm)
                index.search_device(qd, n, k, Dd, Id)
                I = Id.to_host(np.int64, (n, k))
                rec = float(recall_per_query(I, exact).mean())
                cand = (s, bound_scale, wm, rec, float(index.last_nprobe().mean()))
                fallback = cand if fallback is None or rec > fallback[3] else fallback
                if rec >= target_recall + margin:
                    qualified.append(cand)
                    break
    best, best_t = None, None
    for cand in qualified:
        apply(cand[0], cand[1], cand[2])
        index.search_device(qd, n, k, Dd, Id)
        t0 = time.perf_counter()
        for _ in range(reps):
            index.search_device(qd, n, k, Dd, Id)
        t = (time.perf_counter() - t0) / reps
        if best_t is None or t < best_t:
            best, best_t = cand, t
    s, bound_scale, wm, rec, probes = best if best is not None else fallback
    apply(s, bound_scale, wm)
    return s, bound_scale, rec, probes
