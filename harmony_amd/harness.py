"""File formats of the reference's query harness, restated for callers of this library.

  load_fvecs / load_bvecs / load_txt / load_xvecs   src/utils.h:187-407 (loadFvecs, loadBvecs,
                                                    loadBvecs2Fvecs, loadTvecs, loadXvecs)
  xvecs_info                                        src/utils.h:187-203, 239-306 (loadXvecsInfo)
  write_results / load_results                      src/utils.h:565-620 (writeResultsToFile,
                                                    loadResults): the groundtruth_<k>.<fmt> files
  LOG_COLUMNS / log_row / append_log                src/stats.h:122-134 (Stats::myToCsv) written
                                                    the way CsvWriter writes (src/utils.h:41-136)
  harness_paths                                     src/query.cpp:281-300, 368-382

Semantics kept from the reference (so files and numbers line up with runs it made):
  * fvecs: records of int32 d + d float32; the file's first int is d, n = file size // (4 + 4 d)
    (a trailing partial record is ignored); bounds = (a, b), 1-based and inclusive, b = 0 meaning
    the last record (utils.h:205-238).
  * bvecs: records of int32 d + d uint8, converted to float32. The reference's loadBvecs reads
    each vector 4 bytes past its record's data (it seeks to 4 + (a-1)(4+d), then skips another 4
    bytes per record, utils.h:284-300) and counts n = (size - 4) // (4 + d), one record fewer than
    a file that ends on a record boundary holds. load_bvecs(reference_offset=True) reproduces those
    reads and that count exactly; the default reads the records where they are.
  * txt: one vector per line, comma separated, the first field dropped (a label); loadXvecs
    ignores the bounds for txt (utils.h:394-397), so does load_xvecs.
  * results: binary = int64 ids [nq][k] then float32 distances [nq][k]; txt = per query line
    "id dist id dist ..." with 6 decimals (utils.h:565-593). The reference decides by the path's
    ending: "txt" -> text, anything else -> binary.
  * CSV: a header line when the file is new (or not appended to), values comma separated, floating
    point in fixed notation with digits10 + 1 decimals (7 for float, 16 for double), bools as 0/1.
"""
import math
import os

import numpy as np


def _bounds(bounds, nmax):
    a, b = int(bounds[0]), int(bounds[1])
    b = nmax if b == 0 else b
    if not (a >= 1 and b <= nmax and b >= a):  # the reference asserts (utils.h:227, 293, 385)
        raise ValueError("bounds (%d, %d) outside 1..%d" % (a, b, nmax))
    return a, b


def xvecs_info(path):
    """(n, d) of a .fvecs / .bvecs / .txt file (loadXvecsInfo, utils.h:321-331)."""
    path = str(path)
    if path.endswith(".fvecs") or path.endswith(".bvecs"):
        size = os.path.getsize(path)
        d = int(np.fromfile(path, dtype="<i4", count=1)[0])
        if path.endswith(".fvecs"):
            return size // (4 + 4 * d), d
        return size // (4 + d), d  # loadBvecsInfo counts (size / (4 + d)), utils.h:258-272
    if path.endswith(".txt"):
        m, d = 0, 0
        with open(path) as f:
            for line in f:
                line = line.rstrip("\n")
                if m == 0:
                    d = max(0, len(line.split(",")) - 1)
                m += 1
        return m, d
    raise ValueError("no support file")


def load_fvecs(path, bounds=(1, 0)):
    """float32 [n][d] of records a..b (1-based, inclusive; b = 0: the last)."""
    raw = np.fromfile(str(path), dtype=np.uint8)
    d = int(raw[:4].view("<i4")[0])
    rec = 4 + 4 * d
    a, b = _bounds(bounds, raw.size // rec)
    block = raw[(a - 1) * rec: b * rec].reshape(b - a + 1, rec)
    return np.ascontiguousarray(block[:, 4:]).view("<f4").reshape(b - a + 1, d).astype(np.float32)


def load_bvecs(path, bounds=(1, 0), reference_offset=False):
    """uint8 records as float32 [n][d] (loadBvecs2Fvecs). reference_offset=True reads what the
    reference's loadBvecs reads (4 bytes past each record's data)."""
    raw = np.fromfile(str(path), dtype=np.uint8)
    d = int(raw[:4].view("<i4")[0])
    rec = 4 + d
    if not reference_offset:
        a, b = _bounds(bounds, raw.size // rec)
        block = raw[(a - 1) * rec: b * rec].reshape(b - a + 1, rec)
        return block[:, 4:].astype(np.float32)
    a, b = _bounds(bounds, (raw.size - 4) // rec)
    start = 4 + (a - 1) * rec  # seekg(4 + (a - 1) * vecSizeof), then seekg(4, cur) per record
    n = b - a + 1
    end = start + 4 + (n - 1) * rec + d
    if end > raw.size:
        raise ValueError("the reference's bvecs reads run past the end of the file for record %d" % b)
    out = np.empty((n, d), np.float32)
    for i in range(n):
        p = start + i * rec + 4
        out[i] = raw[p:p + d]
    return out


def load_txt(path):
    """float32 [m][d]: one vector per line, comma separated, first field dropped (loadTvecs)."""
    rows = []
    with open(str(path)) as f:
        for line in f:
            parts = line.rstrip("\n").split(",")
            rows.append([float(v) for v in parts[1:]])
    d = len(rows[0]) if rows else 0
    out = np.zeros((len(rows), d), np.float32)
    for i, r in enumerate(rows):  # rows are copied d values each, as utils.h:387-391 does
        out[i, :min(d, len(r))] = np.asarray(r[:d], np.float32)
    return out


def load_xvecs(path, bounds=(1, 0)):
    """By the path's suffix (loadXvecs, utils.h:393-404); txt ignores the bounds as the reference does."""
    path = str(path)
    if path.endswith(".fvecs"):
        return load_fvecs(path, bounds)
    if path.endswith(".bvecs"):
        return load_bvecs(path, bounds)
    if path.endswith(".txt"):
        return load_txt(path)
    raise ValueError("no support file")


def write_fvecs(path, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    rec = np.empty((n, 1 + d), dtype="<f4")
    rec[:, 0] = np.frombuffer(np.int32(d).tobytes(), dtype="<f4")[0]
    rec[:, 1:] = x
    rec.tofile(str(path))


def write_bvecs(path, x):
    x = np.ascontiguousarray(x, dtype=np.uint8)
    n, d = x.shape
    rec = np.empty((n, 4 + d), np.uint8)
    rec[:, :4] = np.frombuffer(np.int32(d).tobytes(), np.uint8)
    rec[:, 4:] = x
    rec.tofile(str(path))


def write_results(path, I, D):
    """groundtruth / result file (writeResultsToFile, utils.h:565-593); creates the directory."""
    path = str(path)
    I = np.ascontiguousarray(I, dtype=np.int64)
    D = np.ascontiguousarray(D, dtype=np.float32)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    if path.endswith("txt"):
        with open(path, "w") as f:
            for i in range(I.shape[0]):
                f.write(" ".join("%d %.6f" % (int(a), float(b)) for a, b in zip(I[i], D[i])) + "\n")
    else:
        with open(path, "wb") as f:
            f.write(I.astype("<i8").tobytes())
            f.write(D.astype("<f4").tobytes())


def load_results(path, nq, k):
    """(I int64 [nq][k], D float32 [nq][k]) of a result file (loadResults, utils.h:595-620)."""
    path = str(path)
    if path.endswith("txt"):
        vals = open(path).read().split()
        if len(vals) < 2 * nq * k:
            raise ValueError("result file holds fewer than nq * k entries")
        I = np.array([int(v) for v in vals[0:2 * nq * k:2]], np.int64).reshape(nq, k)
        D = np.array([float(v) for v in vals[1:2 * nq * k:2]], np.float32).reshape(nq, k)
        return I, D
    raw = open(path, "rb").read()
    if len(raw) < nq * k * 12:
        raise ValueError("result file holds fewer than nq * k entries")
    I = np.frombuffer(raw, "<i8", nq * k).reshape(nq, k).astype(np.int64)
    D = np.frombuffer(raw, "<f4", nq * k, nq * k * 8).reshape(nq, k).astype(np.float32)
    return I, D


# Stats::myToCsv (stats.h:122-134): column, C++ type of the value written
LOG_COLUMNS = [
    ("dataset", "str"), ("mode", "str"), ("nb", "size_t"), ("nq", "size_t"), ("d", "size_t"),
    ("nlist", "size_t"), ("nprobe", "size_t"), ("k", "size_t"), ("divideIVF", "bool"),
    ("orderOptimize", "bool"), ("UnblockSend", "bool"), ("cut", "bool"), ("block", "size_t"),
    ("worker", "size_t"), ("group", "size_t"), ("team", "size_t"), ("time_speedup", "float"),
    ("query_time", "double"), ("original_time", "double"), ("recall", "double"), ("r2", "double"),
    ("variance", "double"), ("train_time", "double"), ("add_time", "double"), ("preSearch_time", "double"),
    ("ratio", "double"), ("teamRatio", "double"), ("brute_ratio", "double"),
]


def _fmt(value, kind):
    if kind == "str":
        return str(value)
    if kind == "bool":
        return "1" if value else "0"
    if kind == "size_t":
        return "%d" % int(value)
    if kind == "float":  # std::fixed, setprecision(digits10 + 1) of a float
        return "%.7f" % float(np.float32(value))
    return "%.16f" % float(value)


def log_row(values):
    """One CSV line of myToCsv's columns (values: dict by column name)."""
    return ",".join(_fmt(values[name], kind) for name, kind in LOG_COLUMNS)


def append_log(path, values, append=True):
    """Write one row as CsvWriter(path, headers, append) + myToCsv do: the header when the file
    is created (or when not appending), then the row."""
    path = str(path)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fresh = not (append and os.path.exists(path))
    with open(path, "w" if fresh else "a") as f:
        if fresh:
            f.write(",".join(name for name, _ in LOG_COLUMNS) + "\n")
        f.write(log_row(values) + "\n")


def time_speedup(faiss_query_time, query_time):
    """Stats::summary's time_speedup (stats.h:90-97): 100 * faiss / own time, 0 without a time."""
    return 0.0 if query_time == 0 else 100.0 * faiss_query_time / query_time


def probe_variance(probes, nlist):
    """The spread of per-list probe counts query.cpp:729-741 logs as "variance": the standard
    deviation of how often each list was probed."""
    counts = np.bincount(np.asarray(probes).ravel(), minlength=nlist).astype(np.float64)
    avg = probes.size / float(nlist)
    return math.sqrt(float(((counts - avg) ** 2).sum()) / nlist)


def harness_paths(benchmarks_path, dataset, input_format="fvecs", output_format="bin", k=100, nlist=0, opt=0,
                  sub_nprobe_ratio=1.0):
    """The file layout query.cpp reads and writes (query.cpp:281-300, 368-382)."""
    root = os.path.join(str(benchmarks_path), dataset)
    # std::format("{}") of a float: its shortest round-trip form ("1", "0.5")
    ratio = np.format_float_positional(np.float32(sub_nprobe_ratio), trim="-")
    return {
        "base": os.path.join(root, "origin", "%s_base.%s" % (dataset, input_format)),
        "query": os.path.join(root, "origin", "%s_query.%s" % (dataset, input_format)),
        "groundtruth": os.path.join(root, "result", "groundtruth_%d.%s" % (k, output_format)),
        "log": os.path.join(root, "result", "log.csv"),
        "index": os.path.join(root, "index", "index_nlist_%d_opt_%d_subNprobeRatio_%s.index" % (nlist, opt, ratio)),
    }
