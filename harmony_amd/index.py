"""Python mirror of harmony::Index (src/Index.h:13-149) over the C-ABI.

Same names and argument meaning as the reference class where they exist:
  Index(d, nlist, nprobe, metric, opt_level)      Index.h:15-25
  train(x, niter, seed, max_points_per_centroid)  Index.h:69 / Clustering.h:13-18
  add(x)                                          Index.h:80 (ids are positions 0..n-1)
  preSearch(warmup_list, warmup_list_size, ...)   Index.h:90 (the warm-up / pruning knobs)
  search(queries, k) -> (D, I)                    Index.h:82
  nprobe, opt_level                               mutable public members (query.cpp:683,686)
Errors raise HvError (the reference throws std::runtime_error / exits; Index.cpp:1642-1645).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (HV_BOUND_EXTERNAL, HV_BOUND_WARMUP, HV_METRIC_L2, HV_OPT_NONE, HvError, KmeansParams,
                   SearchParams, Stats, check, lib)


def _f32(a, shape_last=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape_last is not None and (a.ndim != 2 or a.shape[1] != shape_last):
        raise ValueError("expected an [n][%d] float32 array, got %r" % (shape_last, a.shape))
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceArray:
    """A raw device allocation owned by the library's allocator (no GPU framework needed)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(None, lib().hv_device_malloc(ctypes.byref(p), self.nbytes))
        self.ptr = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        check(None, lib().hv_memcpy_htod(d.ptr, _ptr(a), a.nbytes))
        return d

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        check(None, lib().hv_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().hv_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Index:
    def __init__(self, d, nlist, nprobe=1, metric=HV_METRIC_L2, opt_level=HV_OPT_NONE, dim_order=None):
        self._h = None
        L = lib()
        h = L.hv_create(int(d), int(nlist), int(metric), int(opt_level))
        if not h:
            raise HvError(_lib.HV_ERR_NO_DEVICE, (L.hv_last_error(None) or b"").decode())
        self._h = ctypes.c_void_p(h)
        self.d = int(d)
        self.nlist = int(nlist)
        self.metric = metric
        self._params = SearchParams()
        L.hv_default_search_params(ctypes.byref(self._params))
        self._params.nprobe = int(nprobe)
        self._params.opt_level = int(opt_level)
        if dim_order is not None:
            perm = np.ascontiguousarray(dim_order, dtype=np.int32)
            check(self._h, L.hv_set_dim_order(self._h, _ptr(perm)))
        self._push()

    # -- lifetime
    def close(self):
        if self._h:
            lib().hv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- knobs
    def _push(self):
        check(self._h, lib().hv_set_search_params(self._h, ctypes.byref(self._params)))

    @property
    def nprobe(self):
        return self._params.nprobe

    @nprobe.setter
    def nprobe(self, v):
        self._params.nprobe = int(v)
        self._push()

    @property
    def opt_level(self):
        return self._params.opt_level

    @opt_level.setter
    def opt_level(self, v):
        self._params.opt_level = int(v)
        self._push()

    def preSearch(self, warmup_list=10, warmup_list_size=0, pruning=True, dim_blocks=4, bound_mode=HV_BOUND_WARMUP,
                  survivor_cap=0, refine_bound=True, tile_shape=0, tile_shape_first=0, precision=None,
                  learned=None, learned_nprobe_scale=None, learned_bound_scale=None, learned_min_nprobe=None,
                  rerank_r=None, graph=None, scan_vsplit=None, learned_margin=None, nearest_pass=None):
        """Warm-up and pruning knobs (Index::preSearch, Index.h:90; Param::pruning, Index.h:59).

        precision: HV_PRECISION_FP32 (default), HV_PRECISION_BF16 (MFMA filter + FP32 distances
        of the survivors; same results) or HV_PRECISION_FP8 (FP8 MFMA estimates, the k * rerank_r
        smallest per query re-ranked in FP32; approximate). None keeps the current setting.
        nearest_pass (BF16): 1 (the default) filters and reduces each query's nearest list in one
        kernel; 0 writes every pair of the nearest list out and compacts it in a second; the
        results are the same. None keeps the current setting.

        learned (None keeps the current settings): per-query probe count and pruning bound from
        the trained predictor (mlp_*): probes = ceil(predicted nprobe * learned_nprobe_scale),
        capped at nprobe; bound = predicted k-th distance * learned_bound_scale (<= 0: none).
        learned_margin (FP32 and BF16 forms): a query's probe count also grows to every list of
        its first nprobe whose triangle lower bound (coarse distance minus the list's radius) is
        within sqrt(margin * seeded bound); a query whose k-th result lands above its seeded bound
        is redone exactly over all nprobe lists. With margin >= 1 (the default) every learned
        result is the plain search's at nprobe; 0 keeps the predicted counts alone."""
        p = self._params
        if precision is not None:
            p.precision = int(precision)
        if learned is not None:
            p.learned = 1 if learned else 0
        if learned_nprobe_scale is not None:
            p.learned_nprobe_scale = float(learned_nprobe_scale)
        if learned_bound_scale is not None:
            p.learned_bound_scale = float(learned_bound_scale)
        if learned_min_nprobe is not None:
            p.learned_min_nprobe = int(learned_min_nprobe)
        if learned_margin is not None:
            p.learned_margin = float(learned_margin)
        if rerank_r is not None:
            p.rerank_r = int(rerank_r)
        if graph is not None:
            p.graph = 1 if graph else 0
        if scan_vsplit is not None:
            p.scan_vsplit = int(scan_vsplit)
        if nearest_pass is not None:
            p.nearest_pass = 1 if nearest_pass else 0
        p.refine_bound = 1 if refine_bound else 0
        p.tile_shape = int(tile_shape)
        p.tile_shape_first = int(tile_shape_first)
        p.warmup_list = int(warmup_list)
        p.warmup_list_size = int(warmup_list_size)
        p.pruning = 1 if pruning else 0
        p.dim_blocks = int(dim_blocks)
        p.bound_mode = int(bound_mode)
        p.survivor_cap = int(survivor_cap)
        self._push()

    def set_external_bounds(self, bounds):
        b = np.ascontiguousarray(bounds, dtype=np.float32)
        check(self._h, lib().hv_set_external_bounds(self._h, b.size, _ptr(b)))

    def set_query_nprobe(self, nprobe_per_query):
        """Per-query probe counts for the next searches (None clears): query i scans its first
        nprobe_per_query[i] lists (clamped to [1, nprobe])."""
        a = np.zeros(0, np.int32) if nprobe_per_query is None else np.ascontiguousarray(nprobe_per_query, np.int32)
        check(self._h, lib().hv_set_query_nprobe(self._h, a.size, _ptr(a)))

    def set_query_k(self, k_per_query):
        """Per-query k for the next searches (None clears): row i holds query i's k_i results,
        then (FLT_MAX, -1) padding (k_i clamped to [1, k] of the search)."""
        a = np.zeros(0, np.int32) if k_per_query is None else np.ascontiguousarray(k_per_query, np.int32)
        check(self._h, lib().hv_set_query_k(self._h, a.size, _ptr(a)))

    # -- build
    def train(self, x, niter=10, seed=0, max_points_per_centroid=256):
        x = _f32(x, self.d)
        kp = KmeansParams(int(niter), int(seed), int(max_points_per_centroid))
        check(self._h, lib().hv_train(self._h, x.shape[0], _ptr(x), ctypes.byref(kp)))

    def set_centroids(self, c):
        c = _f32(c, self.d)
        check(self._h, lib().hv_set_centroids(self._h, _ptr(c)))

    def add(self, x):
        x = _f32(x, self.d)
        check(self._h, lib().hv_add(self._h, x.shape[0], _ptr(x)))

    # -- search
    def search(self, queries, k, return_stats=False):
        q = _f32(queries, self.d)
        n = q.shape[0]
        D = np.empty((n, k), dtype=np.float32)
        I = np.empty((n, k), dtype=np.int64)
        st = Stats()
        check(self._h, lib().hv_search(self._h, n, _ptr(q), int(k), _ptr(D), _ptr(I), ctypes.byref(st)))
        return (D, I, st.as_dict()) if return_stats else (D, I)

    def search_device(self, q_dev, n, k, D_dev, I_dev):
        """Queries and results already in HBM (DeviceArray). Returns the stats dict."""
        st = Stats()
        check(self._h, lib().hv_search_device(self._h, int(n), q_dev.ptr, int(k), D_dev.ptr, I_dev.ptr,
                                              ctypes.byref(st)))
        return st.as_dict()

    def search_device_noalloc(self, q_dev, n, k, D_dev, I_dev):
        """search_device with the library's fixed per-GPU scratch block: allocates nothing."""
        st = Stats()
        check(self._h, lib().hv_search_device_noalloc(self._h, int(n), q_dev.ptr, int(k), D_dev.ptr, I_dev.ptr,
                                                      ctypes.byref(st)))
        return st.as_dict()

    def ground_truth_f64(self, queries, k):
        q = _f32(queries, self.d)
        n = q.shape[0]
        D = np.empty((n, k), dtype=np.float64)
        I = np.empty((n, k), dtype=np.int64)
        check(self._h, lib().hv_ground_truth_f64(self._h, n, _ptr(q), int(k), _ptr(D), _ptr(I)))
        return D, I

    # -- persistence (Index::save_index / load_index, Index.h:83-85)
    def save_index(self, path):
        check(self._h, lib().hv_save(self._h, str(path).encode()))

    def load_index(self, path):
        check(self._h, lib().hv_load(self._h, str(path).encode()))
        L = lib()
        self.d = int(L.hv_dim(self._h))
        self.nlist = int(L.hv_nlist(self._h))
        # the load resets the search opt level (Index.cpp:1743): keep the mirror in step
        check(self._h, L.hv_get_search_params(self._h, ctypes.byref(self._params)))

    # -- introspection
    @property
    def ntotal(self):
        return lib().hv_ntotal(self._h)

    def centroids(self):
        out = np.empty((self.nlist, self.d), dtype=np.float32)
        check(self._h, lib().hv_get_centroids(self._h, _ptr(out)))
        return out

    def list_offsets(self):
        out = np.empty(self.nlist + 1, dtype=np.int64)
        check(self._h, lib().hv_get_list_offsets(self._h, _ptr(out)))
        return out

    def list_ids(self):
        out = np.empty(self.ntotal, dtype=np.int64)
        check(self._h, lib().hv_get_list_ids(self._h, _ptr(out)))
        return out

    def list_c2c(self):
        """METRIC_IP: each vector's inner product with its centroid, list order."""
        out = np.empty(self.ntotal, dtype=np.float32)
        check(self._h, lib().hv_get_list_c2c(self._h, _ptr(out)))
        return out

    def list_codes(self):
        out = np.empty((self.ntotal, self.d), dtype=np.float32)
        check(self._h, lib().hv_get_list_codes(self._h, _ptr(out)))
        return out

    def last_n(self):
        """Queries of the last search."""
        return int(lib().hv_last_n(self._h))

    def last_probes(self, n=None):
        """Probe lists of the last search ([n][stride], stride = nprobe).
        n defaults to the last search's query count; a smaller n is rejected by the library."""
        n = self.last_n() if n is None else int(n)
        stride = int(lib().hv_last_probe_stride(self._h)) or min(self.nprobe, self.nlist)
        out = np.empty((n, stride), dtype=np.int64)
        check(self._h, lib().hv_get_last_probes(self._h, _ptr(out), out.size))
        return out

    def last_nprobe(self, n=None):
        """Per-query probe counts of the last search."""
        n = self.last_n() if n is None else int(n)
        out = np.empty(n, dtype=np.int32)
        check(self._h, lib().hv_get_last_nprobe(self._h, _ptr(out), out.size))
        return out

    # -- learned probe/bound predictor
    def query_features(self, queries):
        """[n][65]: log1p of the 64 smallest upper bounds of the squared centroid distances that the
        coarse step forms (ascending), log1p(|q|^2): the learned search's features."""
        q = _f32(queries, self.d)
        out = np.empty((q.shape[0], _lib.HV_MLP_IN), dtype=np.float32)
        check(self._h, lib().hv_query_features(self._h, q.shape[0], _ptr(q), _ptr(out)))
        return out

    def mlp_set_norm(self, fmean, fstd, tmean, tstd):
        norm = np.concatenate([np.asarray(a, np.float32).ravel() for a in (fmean, fstd, tmean, tstd)])
        assert norm.size == 2 * _lib.HV_MLP_IN + 2 * _lib.HV_MLP_OUT
        check(self._h, lib().hv_mlp_set_norm(self._h, _ptr(norm)))

    def mlp_get_norm(self):
        out = np.empty(2 * _lib.HV_MLP_IN + 2 * _lib.HV_MLP_OUT, dtype=np.float32)
        check(self._h, lib().hv_mlp_get_norm(self._h, _ptr(out)))
        i, o = _lib.HV_MLP_IN, _lib.HV_MLP_OUT
        return out[:i], out[i:2 * i], out[2 * i:2 * i + o], out[2 * i + o:]

    def mlp_init(self, seed=0):
        check(self._h, lib().hv_mlp_init(self._h, int(seed)))

    def mlp_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        assert p.size == _lib.HV_MLP_PARAMS
        check(self._h, lib().hv_mlp_set_params(self._h, _ptr(p)))

    def mlp_get_params(self):
        out = np.empty(_lib.HV_MLP_PARAMS, dtype=np.float32)
        check(self._h, lib().hv_mlp_get_params(self._h, _ptr(out)))
        return out

    def mlp_grad(self, feat, target):
        """(loss, gradient) of 1/(2n) sum (y - t)^2 over the standardized batch."""
        x = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(_lib.HV_MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_float(0.0)
        check(self._h, lib().hv_mlp_grad(self._h, x.shape[0], _ptr(x), _ptr(t), _ptr(g), ctypes.byref(loss)))
        return float(loss.value), g

    def mlp_train(self, feat, target, epochs=20, batch=1024, lr=1e-3, seed=0):
        """Adam on the GPU; returns each epoch's mean minibatch loss."""
        x = np.ascontiguousarray(feat, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        losses = np.empty(int(epochs), dtype=np.float32)
        check(self._h, lib().hv_mlp_train(self._h, x.shape[0], _ptr(x), _ptr(t), int(epochs), int(batch),
                                          float(lr), int(seed), _ptr(losses)))
        return losses

    def mlp_predict(self, feat):
        """[n][2] raw targets: log1p(k-th distance), log(nprobe)."""
        x = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.empty((x.shape[0], 2), dtype=np.float32)
        check(self._h, lib().hv_mlp_predict(self._h, x.shape[0], _ptr(x), _ptr(out)))
        return out

    def last_bounds(self, n=None):
        n = self.last_n() if n is None else int(n)
        out = np.empty(n, dtype=np.float32)
        check(self._h, lib().hv_get_last_bounds(self._h, _ptr(out), out.size))
        return out
