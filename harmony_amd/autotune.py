"""Autotuner for the search's tile shapes, timed on the caller's index and a sample of queries.

The knobs change only speed, never results (every form returns the same ids and distances for
any setting; the FP8 form's shortlist is the same too):
  * tile_shape_first / tile_shape: the FP32 list scan's compiled tiles for the nearest-list phase
    and the remaining-lists phase (queries per tile x vectors per tile, 7 shapes; 0 = the
    library's rule by queries per list);
  * scan_vsplit: vectors per work item of the MFMA filter scans (BF16 / FP8 forms);
  * graph: one captured HIP graph per search instead of ~40 launches.
The dimension-block count B is not a knob here: it fixes the order of the FP32 partial sums
(node.cpp:639-646), hence the last bit of every distance. A coordinate search: each knob in
turn over its candidates with the others fixed at their best so far, each candidate timed as
the median of `reps` device-resident searches (hv_search_device) after two warm-up calls (in
graph mode the first sizes the buffers and the second captures). No counterpart in the reference (its per-node scan has no
tiles).
"""
import numpy as np

from ._lib import HV_PRECISION_FP32
from .index import DeviceArray

SHAPES = {
    0: "automatic", 1: "64 queries x 128 vectors", 2: "64 x 64", 3: "32 x 128", 4: "32 x 64",
    5: "16 x 256", 6: "16 x 128", 7: "64 x 128 (8 x 4 per thread)",
}


def _time(index, bufs, k, reps):
    q, n, D, I = bufs
    index.search_device(q, n, k, D, I)
    index.search_device(q, n, k, D, I)
    ts = []
    for _ in range(reps):
        st = index.search_device(q, n, k, D, I)
        ts.append(st["ms_total"])
    return float(np.median(ts))


def autotune(index, queries, k, sample=4096, reps=3, vsplits=(128, 256, 512, 1024), try_graph=True, log=None):
    """Tune the index's search parameters for its current form (precision, learned, nprobe) on
    up to `sample` of `queries`; leaves the best setting applied and returns
    (best parameters dict, {setting: ms} of every trial)."""
    q = np.ascontiguousarray(queries[:sample], dtype=np.float32)
    n = q.shape[0]
    # timed on device-resident queries and results (the path a serving caller keeps hot)
    bufs = (DeviceArray.from_host(q), n, DeviceArray(n * k * 4), DeviceArray(n * k * 8))
    p = index._params
    fp32 = p.precision == HV_PRECISION_FP32
    best = {"tile_shape_first": int(p.tile_shape_first), "tile_shape": int(p.tile_shape),
            "scan_vsplit": int(p.scan_vsplit), "graph": int(p.graph)}
    trials = {}

    def apply(cfg):
        p.tile_shape_first = cfg["tile_shape_first"]
        p.tile_shape = cfg["tile_shape"]
        p.scan_vsplit = cfg["scan_vsplit"]
        p.graph = cfg["graph"]
        index._push()

    def measure(cfg):
        key = tuple(sorted(cfg.items()))
        if key not in trials:
            apply(cfg)
            trials[key] = _time(index, bufs, k, reps)
            if log:
                log("autotune %s: %.3f ms" % (dict(key), trials[key]))
        return trials[key]

    knobs = []
    if fp32:
        knobs += [("tile_shape_first", range(0, 8)), ("tile_shape", range(0, 8))]
    if not fp32:
        knobs.append(("scan_vsplit", vsplits))
    if try_graph:
        knobs.append(("graph", (0, 1)))
    best_ms = measure(best)
    for name, values in knobs:
        for v in values:
            cfg = dict(best, **{name: int(v)})
            ms = measure(cfg)
            if ms < best_ms:
                best, best_ms = cfg, ms
    apply(best)
    return dict(best, ms=best_ms), {str(dict(k_)): v for k_, v in trials.items()}
