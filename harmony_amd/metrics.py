"""Quality metrics of a search result (restated from src/utils.h:660-756)."""
import numpy as np


def recall(I, GT, k=None):
    """|I_q ∩ GT_q| summed over queries / (nq * k); a -1 in I ends its row
    (calculate_recall, utils.h:660-688)."""
    I = np.asarray(I)
    GT = np.asarray(GT)
    nq, kk = I.shape
    k = kk if k is None else k
    hit = 0
    for i in range(nq):
        row = I[i, :k]
        stop = np.nonzero(row == -1)[0]
        if stop.size:
            row = row[: stop[0]]
        hit += np.intersect1d(row, GT[i, :k], assume_unique=False).size
    return hit / float(nq * k)


def recall_per_query(I, GT, k=None):
    I = np.asarray(I)
    GT = np.asarray(GT)
    nq, kk = I.shape
    k = kk if k is None else k
    out = np.empty(nq)
    for i in range(nq):
        row = I[i, :k]
        out[i] = np.intersect1d(row[row >= 0], GT[i, :k]).size / float(k)
    return out


def r2(I, D, GD):
    """sum(D) / sum(GD) - 1 over the non-padded entries (calculate_r2, utils.h:734-756)."""
    I = np.asarray(I)
    mask = np.cumprod(I != -1, axis=1).astype(bool)
    return float(np.asarray(D, dtype=np.float64)[mask].sum() / np.asarray(GD, dtype=np.float64)[mask].sum() - 1.0)


def distance_error(D, GD, I):
    """Max and mean relative error of the returned distances against FP64 distances of the
    same rows (valid entries only)."""
    I = np.asarray(I)
    mask = I >= 0
    d = np.asarray(D, dtype=np.float64)[mask]
    g = np.asarray(GD, dtype=np.float64)[mask]
    rel = np.abs(d - g) / np.maximum(np.abs(g), 1e-30)
    return float(rel.max(initial=0.0)), float(rel.mean() if rel.size else 0.0)
