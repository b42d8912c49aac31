"""ctypes binding of libharmony_amd.so (the C-ABI declared in include/harmony_ivf.h).

The library is loaded from this package directory only; there is no fallback. A missing
library raises HipExtensionMissing with the command that builds it.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libharmony_amd.so")

HV_OK = 0
HV_ERR_INVALID_ARG = -1
HV_ERR_STATE = -2
HV_ERR_DEVICE = -3
HV_ERR_IO = -4
HV_ERR_CAPACITY = -5
HV_ERR_NO_DEVICE = -6

HV_METRIC_IP = 0
HV_METRIC_L2 = 1
HV_OPT_NONE = 0
HV_OPT_TRIANGLE = 1
HV_OPT_SUBNN_L2 = 2
HV_OPT_SUBNN_IP = 4
HV_OPT_ALL = 7
HV_BOUND_WARMUP = 0
HV_BOUND_EXTERNAL = 1
HV_PRECISION_FP32 = 0
HV_PRECISION_BF16 = 1
HV_PRECISION_FP8 = 2
HV_MLP_IN = 65
HV_MLP_HIDDEN = 128
HV_MLP_OUT = 2
HV_MLP_PARAMS = HV_MLP_IN * HV_MLP_HIDDEN + HV_MLP_HIDDEN + HV_MLP_HIDDEN * HV_MLP_OUT + HV_MLP_OUT


class HipExtensionMissing(RuntimeError):
    pass


class HvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("harmony_amd error %d: %s" % (code, msg))
        self.code = code


class KmeansParams(ctypes.Structure):
    _fields_ = [("niter", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("max_points_per_centroid", ctypes.c_int64)]


class SearchParams(ctypes.Structure):
    _fields_ = [("nprobe", ctypes.c_int64), ("opt_level", ctypes.c_int32), ("dim_blocks", ctypes.c_int32),
                ("pruning", ctypes.c_int32), ("bound_mode", ctypes.c_int32), ("warmup_list", ctypes.c_int64),
                ("warmup_list_size", ctypes.c_int64), ("precision", ctypes.c_int32),
                ("refine_bound", ctypes.c_int32), ("survivor_cap", ctypes.c_int64), ("tile_shape", ctypes.c_int32),
                ("tile_shape_first", ctypes.c_int32), ("learned", ctypes.c_int32),
                ("learned_nprobe_scale", ctypes.c_float), ("learned_bound_scale", ctypes.c_float),
                ("learned_min_nprobe", ctypes.c_int32), ("rerank_r", ctypes.c_int32),
                ("graph", ctypes.c_int32), ("scan_vsplit", ctypes.c_int32), ("learned_margin", ctypes.c_float),
                ("nearest_pass", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ms_total", "ms_upload", "ms_coarse", "ms_warmup", "ms_group",
                                               "ms_scan", "ms_select", "ms_download")] + \
               [(n, ctypes.c_int64) for n in ("candidates", "block_evals", "block_evals_max", "survivors",
                                              "overflow_queries", "work_items")] + \
               [("ms_main", ctypes.c_double), ("main_pairs", ctypes.c_int64), ("fallback_queries", ctypes.c_int64),
                ("ms_rerank", ctypes.c_double), ("rerank_pairs", ctypes.c_int64),
                ("ms_scan_kernels", ctypes.c_double), ("scan_launches", ctypes.c_int64),
                ("scan_pairs", ctypes.c_int64), ("widened_queries", ctypes.c_int64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "hv_create": (_P, [_I64, _I64, _I32, _I32]),
    "hv_destroy": (None, [_P]),
    "hv_last_error": (ctypes.c_char_p, [_P]),
    "hv_set_dim_order": (ctypes.c_int, [_P, _P]),
    "hv_default_kmeans_params": (None, [ctypes.POINTER(KmeansParams)]),
    "hv_train": (ctypes.c_int, [_P, _I64, _P, ctypes.POINTER(KmeansParams)]),
    "hv_set_centroids": (ctypes.c_int, [_P, _P]),
    "hv_add": (ctypes.c_int, [_P, _I64, _P]),
    "hv_default_search_params": (None, [ctypes.POINTER(SearchParams)]),
    "hv_set_search_params": (ctypes.c_int, [_P, ctypes.POINTER(SearchParams)]),
    "hv_get_search_params": (ctypes.c_int, [_P, ctypes.POINTER(SearchParams)]),
    "hv_set_external_bounds": (ctypes.c_int, [_P, _I64, _P]),
    "hv_set_query_nprobe": (ctypes.c_int, [_P, _I64, _P]),
    "hv_set_query_k": (ctypes.c_int, [_P, _I64, _P]),
    "hv_search": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, ctypes.POINTER(Stats)]),
    "hv_search_device": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, ctypes.POINTER(Stats)]),
    "hv_search_device_noalloc": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, ctypes.POINTER(Stats)]),
    "hv_ground_truth_f64": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P]),
    "hv_save": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "hv_load": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "hv_ntotal": (_I64, [_P]),
    "hv_dim": (_I64, [_P]),
    "hv_nlist": (_I64, [_P]),
    "hv_get_centroids": (ctypes.c_int, [_P, _P]),
    "hv_get_list_offsets": (ctypes.c_int, [_P, _P]),
    "hv_get_list_ids": (ctypes.c_int, [_P, _P]),
    "hv_get_list_c2c": (ctypes.c_int, [_P, _P]),
    "hv_get_list_codes": (ctypes.c_int, [_P, _P]),
    "hv_last_n": (_I64, [_P]),
    "hv_get_last_probes": (ctypes.c_int, [_P, _P, _I64]),
    "hv_get_last_bounds": (ctypes.c_int, [_P, _P, _I64]),
    "hv_get_last_nprobe": (ctypes.c_int, [_P, _P, _I64]),
    "hv_last_probe_stride": (_I64, [_P]),
    "hv_query_features": (ctypes.c_int, [_P, _I64, _P, _P]),
    "hv_mlp_set_norm": (ctypes.c_int, [_P, _P]),
    "hv_mlp_get_norm": (ctypes.c_int, [_P, _P]),
    "hv_mlp_init": (ctypes.c_int, [_P, ctypes.c_uint64]),
    "hv_mlp_set_params": (ctypes.c_int, [_P, _P]),
    "hv_mlp_get_params": (ctypes.c_int, [_P, _P]),
    "hv_mlp_grad": (ctypes.c_int, [_P, _I64, _P, _P, _P, _P]),
    "hv_mlp_train": (ctypes.c_int, [_P, _I64, _P, _P, _I32, _I32, ctypes.c_float, ctypes.c_uint64, _P]),
    "hv_mlp_predict": (ctypes.c_int, [_P, _I64, _P, _P]),
    "hv_device_count": (ctypes.c_int, []),
    "hv_device_malloc": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), _I64]),
    "hv_device_free": (ctypes.c_int, [_P]),
    "hv_memcpy_htod": (ctypes.c_int, [_P, _P, _I64]),
    "hv_memcpy_dtoh": (ctypes.c_int, [_P, _P, _I64]),
    "hv_memcpy_dtod": (ctypes.c_int, [_P, _P, _I64]),
}

_lib = None


def lib():
    """Load the library once; raise HipExtensionMissing when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise HipExtensionMissing(
                "libharmony_amd.so is not built; run `python -m harmony_amd.build` (or "
                "__graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(h, rc):
    if rc != HV_OK:
        msg = lib().hv_last_error(h)
        raise HvError(rc, msg.decode() if msg else "")
    return rc
