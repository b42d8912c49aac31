"""Writes tests/golden/w1_golden.npz: workload 1 (10,000 x 128 integer mixture, 100 queries)
expected outputs from the CPU oracle — the FP64 exact top-10 (pinned against numpy in
tests/test_oracle_search.py) and the IVF search result at nlist 100, nprobe 10, k 10, B=4 with
pruning on — plus SHA-256 hashes of the generated arrays and of the oracle's centroids.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from harmony_amd import datagen  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    xb, xq = datagen.make(1)
    GD, GI = O.bruteforce_f64(xb, xq, 10)
    idx = O.build_index(xb, 100, niter=10, seed=0)
    D, I, probes, bound = O.search(idx, xq, 10, nprobe=10, dim_blocks=4, pruning=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "w1_golden.npz")
    np.savez_compressed(out, base_sha256=np.array(datagen.sha256(xb)), query_sha256=np.array(datagen.sha256(xq)),
                        centroid_sha256=np.array(datagen.sha256(idx.centroids)), gt_D=GD, gt_I=GI, D=D, I=I,
                        probes=probes, bound=bound, list_off=idx.list_off)
    print("wrote", out)


if __name__ == "__main__":
    main()
