"""Quality metrics of a search result, restated from src/utils.h:660-756.

recall(I, GT, D, GD) follows calculate_recall (utils.h:660-732): 0 when D[0] < 0 (662-665);
k (the row length of I) must not exceed gt_k (the row length of GT, 671-673); the ground truth
of a query is the set of the first k ids of its GT row (676); the query's ids count until the
first -1 (678-680), and every one found in the set is a hit, so an id that appears twice counts
twice. The result is the FP32 quotient (float)hits / (nq * k) (731). The distance sanity count
of 687-729 is available with return_correct=True.

r2(I, D, GD) follows calculate_r2 (utils.h:734-756): sequential FP32 sums of D and GD over each
row's entries before the first -1 in I, then sum / g_sum - 1 in FP32.
"""
import numpy as np

FEPS = 1e-4  # utils.h:630


def _hits(I, GT):
    """[nq][k] bool: entry (i, j) lies before row i's first -1 and is among GT[i, :k]."""
    I = np.asarray(I, dtype=np.int64)
    GT = np.asarray(GT, dtype=np.int64)
    nq, k = I.shape
    if k > GT.shape[1]:
        raise ValueError("k should be less than or equal to gt_k.")
    valid = np.cumprod(I != -1, axis=1).astype(bool)
    # one sorted key space for all rows: (row << 34) | (id + 1); ids are < 2^33
    row = np.arange(nq, dtype=np.int64)[:, None] << 34
    member = np.isin(row + (I + 1), row + (GT[:, :k] + 1))
    return valid & member


def recall(I, GT, D=None, GD=None, metric_l2=True, return_correct=False):
    """calculate_recall (utils.h:660-732); D / GD optional (D[0] < 0 -> 0 needs D, the distance
    sanity count needs both)."""
    I = np.asarray(I, dtype=np.int64)
    nq, k = I.shape
    if D is not None and np.asarray(D).size and float(np.asarray(D, dtype=np.float32).flat[0]) < 0:
        return (0.0, 0) if return_correct else 0.0
    hits = int(_hits(I, GT).sum())
    r = float(np.float32(hits) / np.float32(nq * k))
    if not return_correct:
        return r
    correct = 0
    if D is not None and GD is not None:
        GT = np.asarray(GT, dtype=np.int64)
        D = np.asarray(D, dtype=np.float32)
        GD = np.asarray(GD, dtype=np.float32)
        valid = np.cumprod(I != -1, axis=1).astype(bool)
        # index of the last GT id other than -1 among the first k (0 when there is none)
        nz = GT[:, :k] != -1
        last = np.where(nz.any(1), k - 1 - np.argmax(nz[:, ::-1], axis=1), 0)
        top = GD[np.arange(nq), last][:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(D == 0, np.abs(top), np.abs((D - top) / D)).astype(np.float32)
        within = D <= top if metric_l2 else D >= top
        correct = int((valid & (within | (rel.astype(np.float64) < FEPS))).sum())
    return r, correct


def recall_per_query(I, GT):
    """Each query's share of recall(I, GT): hits / k per row (the same counting rule)."""
    I = np.asarray(I, dtype=np.int64)
    return _hits(I, GT).sum(1) / float(I.shape[1])


def r2(I, D, GD):
    """calculate_r2 (utils.h:734-756): sequential FP32 sums over each row's entries before the
    first -1 in I (row-major order), sum / g_sum - 1 in FP32."""
    I = np.asarray(I, dtype=np.int64)
    nq, k = I.shape
    GD = np.asarray(GD, dtype=np.float32)
    if k > GD.shape[1]:
        raise ValueError("k should be less than or equal to gt_k.")
    valid = np.cumprod(I != -1, axis=1).astype(bool)
    d = np.asarray(D, dtype=np.float32)[valid]
    g = GD[:, :k][valid]
    s = np.cumsum(d, dtype=np.float32)[-1] if d.size else np.float32(0)
    gs = np.cumsum(g, dtype=np.float32)[-1] if g.size else np.float32(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float32(s) / np.float32(gs) - np.float32(1))


def distance_error(D, GD, I):
    """Max and mean relative error of the returned distances against FP64 distances of the
    same rows (valid entries only)."""
    I = np.asarray(I)
    mask = I >= 0
    d = np.asarray(D, dtype=np.float64)[mask]
    g = np.asarray(GD, dtype=np.float64)[mask]
    rel = np.abs(d - g) / np.maximum(np.abs(g), 1e-30)
    return float(rel.max(initial=0.0)), float(rel.mean() if rel.size else 0.0)
