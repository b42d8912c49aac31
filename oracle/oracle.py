"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so) and of the reference's own heap
compiled from source (oracle/_ref/libref_heap.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TIE_CANONICAL = 0
TIE_VERBATIM = 1

_P = ctypes.c_void_p
_I = ctypes.c_int64
_C = ctypes.c_int

_SIG = {
    "or_heap_init": (None, [_I, _P, _P]),
    "or_heap_push": (None, [_C, _I, _P, _P, ctypes.c_float, _I]),
    "or_heap_replace_top": (None, [_C, _I, _P, _P, ctypes.c_float, _I]),
    "or_heap_sort": (None, [_C, _I, _P, _P]),
    "or_l2": (ctypes.c_float, [_P, _P, _I]),
    "or_l2_blocked": (ctypes.c_float, [_P, _P, _I, _I]),
    "or_coarse": (None, [_I, _P, _I, _I, _P, _I, _C, _C, _P, _P]),
    "or_warmup_bound": (None, [_I, _P, _I, _P, _P, _I, _I, _P, _I, _I, _I, _C, _C, _P]),
    "or_search": (_C, [_I, _P, _I, _I, _P, _P, _P, _P, _I, _I, _I, _C, _I, _I, _C, _C, _C, _P, _P, _P, _P]),
    "or_bruteforce_f64": (None, [_I, _P, _I, _P, _I, _I, _C, _P, _P]),
    "or_assign": (None, [_I, _P, _I, _I, _P, _C, _P, _P]),
    "or_build_lists": (None, [_I, _P, _P, _I, _P, _P]),
    "or_kmeans_init": (None, [_I, _P, _I, _I, _I, ctypes.c_uint64, _P, _P, _P]),
    "or_kmeans_step": (None, [_I, _P, _I, _I, _P, _C]),
    "or_kmeans": (None, [_I, _P, _I, _I, _I, _I, ctypes.c_uint64, _C, _P]),
    "or_ip_normalize_rows": (None, [_I, _P, _I, _P]),
    "or_ip_normalize_centroids": (None, [_I, _I, _P]),
    "or_ip_assign": (None, [_I, _P, _I, _I, _P, _C, _P, _P]),
    "or_ip_kmeans": (None, [_I, _P, _I, _I, _I, _I, ctypes.c_uint64, _C, _P]),
    "or_ip_build_lists": (None, [_I, _P, _P, _I, _P, _P]),
    "or_ip_search": (None, [_I, _P, _I, _I, _P, _P, _P, _P, _P, _I, _I, _C, _C, _P, _P, _P, _P]),
    "or_recall": (ctypes.c_float, [_I, _I, _P, _P, _P, _P, _I, _C, _P, _P]),
    "or_r2": (ctypes.c_float, [_I, _I, _P, _P, _P, _I, _P]),
}

_cache = {}


def _load(name):
    if name not in _cache:
        path = os.path.join(HERE, "build", name)
        if not os.path.isfile(path):
            raise RuntimeError("%s not built: run `python oracle/build.py`" % path)
        L = ctypes.CDLL(path)
        for fn, (res, args) in _SIG.items():
            f = getattr(L, fn)
            f.restype = res
            f.argtypes = args
        _cache[name] = L
    return _cache[name]


def lib(fast=False):
    """The exact oracle, or with fast=True the build with the reference's Release flags (the
    timed CPU baseline), AVX-512 flavour when the host has avx512f."""
    if not fast:
        return _load("liboracle.so")
    flags = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = line
                    break
    except OSError:
        pass
    return _load("liboracle_fast_v4.so" if " avx512f " in flags + " " else "liboracle_fast_v3.so")


def ref_heap():
    """The reference's heap.hpp compiled from its own source, or None when not built."""
    path = os.path.join(HERE, "_ref", "libref_heap.so")
    if not os.path.isfile(path):
        return None
    if "ref" not in _cache:
        L = ctypes.CDLL(path)
        L.ref_init_result.argtypes = [_I, _P, _P]
        L.ref_heap_replace_top.argtypes = [_I, _P, _P, ctypes.c_float, _I]
        L.ref_sort_result.argtypes = [_I, _P, _P]
        for f in (L.ref_init_result, L.ref_heap_replace_top, L.ref_sort_result):
            f.restype = None
        _cache["ref"] = L
    return _cache["ref"]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class IndexData:
    """An IVF index as flat arrays: centroids [nlist][d], list_off [nlist+1], ids [nb] and codes
    [nb][d] in list order (each list in its within-list order)."""

    def __init__(self, centroids, list_off, ids, codes):
        self.centroids = _f32(centroids)
        self.list_off = _i64(list_off)
        self.ids = _i64(ids)
        self.codes = _f32(codes)
        self.nlist, self.d = self.centroids.shape

    @classmethod
    def from_gpu_index(cls, index):
        return cls(index.centroids(), index.list_offsets(), index.list_ids(), index.list_codes())


def coarse(q, centroids, nprobe, tie=TIE_CANONICAL, nthreads=0, fast=False):
    q, c = _f32(q), _f32(centroids)
    nq, d = q.shape
    probes = np.empty((nq, nprobe), np.int64)
    pd = np.empty((nq, nprobe), np.float32)
    lib(fast).or_coarse(nq, _p(q), d, c.shape[0], _p(c), nprobe, tie, nthreads, _p(probes), _p(pd))
    return probes, pd


def search(idx, q, k, nprobe, dim_blocks=4, pruning=True, warmup_list=10, warmup_list_size=0,
           tie=TIE_CANONICAL, bound_blocked=True, nthreads=0, fast=False):
    """(D, I, probes, bound) of the per-node search path (see or_search)."""
    q = _f32(q)
    nq, d = q.shape
    npr = min(nprobe, idx.nlist)
    D = np.empty((nq, k), np.float32)
    I = np.empty((nq, k), np.int64)
    probes = np.empty((nq, npr), np.int64)
    bound = np.empty(nq, np.float32)
    rc = lib(fast).or_search(nq, _p(q), d, idx.nlist, _p(idx.centroids), _p(idx.list_off), _p(idx.ids),
                             _p(idx.codes), k, nprobe, dim_blocks, 1 if pruning else 0, warmup_list,
                             warmup_list_size, tie, 1 if bound_blocked else 0, nthreads, _p(D), _p(I), _p(probes),
                             _p(bound))
    if rc != 0:
        raise ValueError("or_search rejected its arguments")
    return D, I, probes, bound


def bruteforce_f64(xb, q, k, nthreads=0):
    xb, q = _f32(xb), _f32(q)
    nq, d = q.shape
    D = np.empty((nq, k), np.float64)
    I = np.empty((nq, k), np.int64)
    lib().or_bruteforce_f64(xb.shape[0], _p(xb), nq, _p(q), d, k, nthreads, _p(D), _p(I))
    return D, I


def assign(x, centroids, nthreads=0):
    x, c = _f32(x), _f32(centroids)
    n, d = x.shape
    lab = np.empty(n, np.int64)
    dist = np.empty(n, np.float32)
    lib().or_assign(n, _p(x), d, c.shape[0], _p(c), nthreads, _p(lab), _p(dist))
    return lab, dist


def build_lists(labels, dist, nlist):
    labels, dist = _i64(labels), _f32(dist)
    n = labels.size
    off = np.empty(nlist + 1, np.int64)
    order = np.empty(n, np.int64)
    lib().or_build_lists(n, _p(labels), _p(dist), nlist, _p(off), _p(order))
    return off, order


def kmeans(x, nlist, niter=10, max_pts=256, seed=0, nthreads=0):
    x = _f32(x)
    n, d = x.shape
    c = np.empty((nlist, d), np.float32)
    lib().or_kmeans(n, _p(x), d, nlist, niter, max_pts, seed, nthreads, _p(c))
    return c


def kmeans_init(x, nlist, max_pts=256, seed=0):
    x = _f32(x)
    n, d = x.shape
    cap = nlist * max_pts if (max_pts > 0 and n > nlist * max_pts) else n
    sample = np.empty((cap, d), np.float32)
    ns = ctypes.c_int64(0)
    c = np.empty((nlist, d), np.float32)
    lib().or_kmeans_init(n, _p(x), d, nlist, max_pts, seed, _p(sample), ctypes.byref(ns), _p(c))
    return sample[: ns.value], c


def kmeans_step(x, centroids, nthreads=0):
    x = _f32(x)
    c = np.array(centroids, dtype=np.float32, copy=True)
    lib().or_kmeans_step(x.shape[0], _p(x), x.shape[1], c.shape[0], _p(c), nthreads)
    return c


def build_index(xb, nlist, niter=10, max_pts=256, seed=0, nthreads=0):
    """The oracle's own train + add (k-means, assignment, list order)."""
    c = kmeans(xb, nlist, niter, max_pts, seed, nthreads)
    lab, dist = assign(xb, c, nthreads)
    off, order = build_lists(lab, dist, nlist)
    return IndexData(c, off, order, _f32(xb)[order])


def recall(I, GT, D=None, GD=None, metric_l2=True, return_correct=False):
    """calculate_recall (utils.h:660-732) restated in or_recall: k = I's row length, gt_k = GT's.
    D / GD may be None (then the D[0] < 0 rule and the distance sanity count are skipped).
    Raises ValueError when k > gt_k (the reference throws std::invalid_argument)."""
    I, GT = _i64(I), _i64(GT)
    nq, k = I.shape
    Dp = _f32(D) if D is not None else None
    GDp = _f32(GD) if GD is not None else None
    correct = ctypes.c_int64(0)
    status = ctypes.c_int(0)
    r = lib().or_recall(nq, k, _p(I), _p(Dp) if Dp is not None else None, _p(GT),
                        _p(GDp) if GDp is not None else None, GT.shape[1], 1 if metric_l2 else 0,
                        ctypes.byref(correct), ctypes.byref(status))
    if status.value != 0:
        raise ValueError("k should be less than or equal to gt_k.")
    return (float(r), int(correct.value)) if return_correct else float(r)


def r2(I, D, GD):
    """calculate_r2 (utils.h:734-756) restated in or_r2: k = I's row length, gt_k = GD's."""
    I, D, GD = _i64(I), _f32(D), _f32(GD)
    nq, k = I.shape
    status = ctypes.c_int(0)
    r = lib().or_r2(nq, k, _p(I), _p(D), _p(GD), GD.shape[1], ctypes.byref(status))
    if status.value != 0:
        raise ValueError("k should be less than or equal to gt_k.")
    return float(r)


# ---------------------------------------------------------------- inner-product metric
def ip_normalize_rows(x):
    x = _f32(x)
    out = np.empty_like(x)
    lib().or_ip_normalize_rows(x.shape[0], _p(x), x.shape[1], _p(out))
    return out


def ip_normalize_centroids(c):
    c = np.array(c, dtype=np.float32, copy=True)
    lib().or_ip_normalize_centroids(c.shape[0], c.shape[1], _p(c))
    return c


def ip_assign(x, centroids, nthreads=0):
    x, c = _f32(x), _f32(centroids)
    n, d = x.shape
    lab = np.empty(n, np.int64)
    best = np.empty(n, np.float32)
    lib().or_ip_assign(n, _p(x), d, c.shape[0], _p(c), nthreads, _p(lab), _p(best))
    return lab, best


def ip_kmeans(x, nlist, niter=10, max_pts=256, seed=0, nthreads=0):
    x = _f32(x)
    n, d = x.shape
    c = np.empty((nlist, d), np.float32)
    lib().or_ip_kmeans(n, _p(x), d, nlist, niter, max_pts, seed, nthreads, _p(c))
    return c


def ip_build_lists(labels, best, nlist):
    labels, best = _i64(labels), _f32(best)
    off = np.empty(nlist + 1, np.int64)
    order = np.empty(labels.size, np.int64)
    lib().or_ip_build_lists(labels.size, _p(labels), _p(best), nlist, _p(off), _p(order))
    return off, order


class IpIndexData(IndexData):
    """An inner-product index: IndexData plus each vector's candidate2centroid (list order)."""

    def __init__(self, centroids, list_off, ids, codes, c2c):
        super().__init__(centroids, list_off, ids, codes)
        self.c2c = _f32(c2c)


def ip_build_index(xb, nlist, niter=10, max_pts=256, seed=0, nthreads=0):
    """The oracle's own IP train + add."""
    c = ip_kmeans(xb, nlist, niter, max_pts, seed, nthreads)
    lab, best = ip_assign(xb, c, nthreads)
    off, order = ip_build_lists(lab, best, nlist)
    return IpIndexData(c, off, order, _f32(xb)[order], best[order])


def ip_search(idx, q, k, nprobe, triangle=False, nthreads=0, return_skipped=False):
    """(D, I, probes) of the IP search (descending inner products, (-FLT_MAX, -1) padding)."""
    q = _f32(q)
    nq, d = q.shape
    npr = min(nprobe, idx.nlist)
    D = np.empty((nq, k), np.float32)
    I = np.empty((nq, k), np.int64)
    probes = np.empty((nq, npr), np.int64)
    skipped = ctypes.c_int64(0)
    lib().or_ip_search(nq, _p(q), d, idx.nlist, _p(idx.centroids), _p(idx.list_off), _p(idx.ids), _p(idx.codes),
                       _p(idx.c2c), k, nprobe, 1 if triangle else 0, nthreads, _p(D), _p(I), _p(probes),
                       ctypes.byref(skipped))
    return (D, I, probes, skipped.value) if return_skipped else (D, I, probes)
