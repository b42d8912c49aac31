"""GPU parity of the coarse assignment: the probe lists (nearest centroids in the FP32
direct-difference chain's order, ties to the lower centroid id) against the oracle's coarse
step, for every shape the library dispatches: one wavefront per query (list counts up to 512,
1024 and 2048 per row), one workgroup per query (more probes per query or longer rows), and the
unfused path beyond. Duplicated centroids and queries placed on centroids make exact ties."""
import sys

import numpy as np
import pytest

from harmony_amd import HV_PRECISION_BF16, HV_PRECISION_FP32
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test coarse]", *a, flush=True)
    sys.stdout.flush()


def tied_case(nlist, d, nq, seed):
    rng = np.random.default_rng(seed)
    cent = rng.integers(0, 64, size=(nlist, d)).astype(np.float32)
    cent[1::7] = cent[0::7][: cent[1::7].shape[0]]  # duplicated centroids: equal distances
    xq = rng.integers(0, 64, size=(nq, d)).astype(np.float32)
    xq[::5] = cent[rng.integers(0, nlist, size=xq[::5].shape[0])]  # queries on centroids
    xb = cent[rng.integers(0, nlist, size=4 * nlist)] + rng.integers(-2, 3, size=(4 * nlist, d)).astype(np.float32)
    return cent, xb, xq


@pytest.mark.parametrize("nlist,nprobe", [(300, 16), (1000, 64), (1500, 16), (2000, 256), (1000, 300),
                                          (3000, 32), (700, 700), (5000, 8)])
@pytest.mark.parametrize("precision", [HV_PRECISION_FP32, HV_PRECISION_BF16])
def test_probe_lists_match_oracle(nlist, nprobe, precision):
    say("nlist", nlist, "nprobe", nprobe, "precision", precision)
    d, nq = 64, 200
    cent, xb, xq = tied_case(nlist, d, nq, seed=nlist + nprobe)
    g = Index(d, nlist, nprobe=nprobe)
    g.set_centroids(cent)
    g.add(xb)
    g.preSearch(precision=precision)
    g.search(xq, 1)
    probes, _ = O.coarse(xq, cent, nprobe)
    np.testing.assert_array_equal(g.last_probes(nq), probes)
    g.close()
