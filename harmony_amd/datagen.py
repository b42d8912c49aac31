"""Synthetic workloads of the IVF-flat search path (seeded, reproducible).

Random streams come from numpy's PCG64 generator. Stream seeds: centers = S, base = S+1,
queries = S+2, extra query sets S+3, S+4, ... where S is the workload number. Normal draws
are taken in fixed row chunks, so the arrays depend only on the seed and the sizes.

Workloads (sizes as named):
  1  "int-mixture-10k":   50 centers ~ U[0,128)^128; 10,000 base + 100 queries, each
                          clip(round(c_z + N(0, 12^2)), 0, 255), z uniform; integer-valued fp32.
                          Index: nlist=100 (k-means 10 iters, seed 0), nprobe=10, k=10, B=4.
  2  "int-mixture-1m":    the same recipe with 1,000 centers; 1,000,000 base + 10,000 queries.
                          Index: nlist=1024; nprobe in {16, 64, 256}; k=100; B=4.
  3  "skewed-420d":       d=420, per-dim std s_i = exp(N(0,1)); 2,000 centers N(0,1)*s; points
                          mu_z + N(0, 0.3^2)*s; 1,000,000 base + 10,000 queries; nlist=2048.
  4  "zipf-96d-10m":      d=96; 10,000 unit centers; P(z=j) ~ (j+1)^-0.8; points
                          normalize(mu_z + N(0, 0.1^2 I)); 10,000,000 base + 20,000 queries.
Integer-valued workloads keep every partial and full L2 sum below 2^24, so FP32 distances are
exact in any summation order.
"""
import hashlib

import numpy as np

CHUNK = 131072


def _rng(seed):
    return np.random.Generator(np.random.PCG64(int(seed)))


def int_mixture(n_centers, n, d, seed_centers, seed_points, sigma=12.0):
    """clip(round(c_z + N(0, sigma^2)), 0, 255) with centers ~ U[0,128)^d (stream seed_centers)."""
    centers = _rng(seed_centers).uniform(0.0, 128.0, size=(n_centers, d))
    r = _rng(seed_points)
    z = r.integers(0, n_centers, size=n)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        v = centers[z[s:e]] + r.normal(0.0, sigma, size=(e - s, d))
        out[s:e] = np.clip(np.rint(v), 0.0, 255.0)
    return out


def skewed_mixture(n_centers, n, d, seed_centers, seed_points, noise=0.3):
    rc = _rng(seed_centers)
    std = np.exp(rc.normal(0.0, 1.0, size=d))
    centers = rc.normal(0.0, 1.0, size=(n_centers, d)) * std
    r = _rng(seed_points)
    z = r.integers(0, n_centers, size=n)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        out[s:e] = centers[z[s:e]] + r.normal(0.0, noise, size=(e - s, d)) * std
    return out


def zipf_unit_mixture(n_centers, n, d, seed_centers, seed_points, sigma=0.1, alpha=0.8):
    rc = _rng(seed_centers)
    c = rc.normal(0.0, 1.0, size=(n_centers, d))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    w = (np.arange(n_centers) + 1.0) ** (-alpha)
    w /= w.sum()
    r = _rng(seed_points)
    z = r.choice(n_centers, size=n, p=w)
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, CHUNK):
        e = min(n, s + CHUNK)
        v = c[z[s:e]] + r.normal(0.0, sigma, size=(e - s, d))
        out[s:e] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return out


WORKLOADS = {
    1: dict(name="int-mixture-10k", d=128, nb=10_000, nq=100, centers=50, nlist=100, nprobe=[10], k=10,
            dim_blocks=4, kind="int"),
    2: dict(name="int-mixture-1m", d=128, nb=1_000_000, nq=10_000, centers=1000, nlist=1024,
            nprobe=[16, 64, 256], k=100, dim_blocks=4, kind="int"),
    3: dict(name="skewed-420d", d=420, nb=1_000_000, nq=10_000, centers=2000, nlist=2048, nprobe=[64], k=100,
            dim_blocks=8, kind="skewed"),
    4: dict(name="zipf-96d-10m", d=96, nb=10_000_000, nq=20_000, centers=10_000, nlist=4096, nprobe=[8, 512],
            k=100, dim_blocks=4, kind="zipf"),
}


def make(workload, nb=None, nq=None, query_set=0):
    """(base, queries) of a workload; nb / nq override the sizes (same streams, prefix rows
    are NOT the same as the full-size arrays' rows when n changes: the center draw z is drawn
    for all rows first)."""
    w = WORKLOADS[workload]
    S = workload
    nb = w["nb"] if nb is None else nb
    nq = w["nq"] if nq is None else nq
    gen = {"int": int_mixture, "skewed": skewed_mixture, "zipf": zipf_unit_mixture}[w["kind"]]
    base = gen(w["centers"], nb, w["d"], S, S + 1)
    queries = gen(w["centers"], nq, w["d"], S, S + 2 + query_set)
    return base, queries


def uneven_plan(workload, nq, seed_offset=5):
    """Per-query probe counts round(exp(U[ln 8, ln 512])) and k in {10, 100} (stream S+5) of
    the uneven-nprobe workload (zipf-96d-10m)."""
    r = _rng(workload + seed_offset)
    nprobe = np.rint(np.exp(r.uniform(np.log(8.0), np.log(512.0), size=nq))).astype(np.int32)
    k = np.where(r.integers(0, 2, size=nq) == 0, 10, 100).astype(np.int32)
    return nprobe, k


def sha256(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def variance_order(x):
    """Dimensions by descending variance (the block order of the d=420 workload)."""
    v = np.asarray(x, dtype=np.float64).var(axis=0)
    return np.argsort(-v, kind="stable").astype(np.int32)
