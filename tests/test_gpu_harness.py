"""The reference's own benchmark data through this library: the first 2,000 base vectors and the
100 queries of benchmarks/sift10k (integer-valued SIFT descriptors, byte copies under
tests/golden/harness) loaded by the harness loaders, indexed on the GPU and searched in the
per-node mode (4 dimension blocks, pruning) and the unpruned mode, bit for bit against the oracle
on the same index; at nprobe = nlist the ids equal the FP64 brute force. Then the CLI end to end:
index file, ground-truth file and log.csv rows in the reference's layout, and a --cache rerun."""
import os

import numpy as np
import pytest

from harmony_amd import harness as H
from harmony_amd import query
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "harness")


@pytest.fixture(scope="module")
def sift():
    xb = H.load_xvecs(os.path.join(GOLD, "sift10k_base_2000.fvecs"))
    xq = H.load_xvecs(os.path.join(GOLD, "sift10k_query.fvecs"))
    nlist = int(np.sqrt(xb.shape[0]))  # query.cpp:323-325
    g = Index(128, nlist, nprobe=4)
    g.train(xb, niter=10, seed=0)
    g.add(xb)
    return xb, xq, g, O.IndexData.from_gpu_index(g)


@pytest.mark.parametrize("nprobe", [1, 4, 44])
@pytest.mark.parametrize("pruning", [True, False])
def test_sift10k_matches_oracle(sift, nprobe, pruning):
    xb, xq, g, o = sift
    g.nprobe = nprobe
    g.preSearch(warmup_list=10, warmup_list_size=0, pruning=pruning, dim_blocks=4)
    D, I = g.search(xq, 10)
    oD, oI, _, _ = O.search(o, xq, 10, nprobe=nprobe, dim_blocks=4, pruning=pruning)
    np.testing.assert_array_equal(I, oI)
    np.testing.assert_array_equal(D, oD)
    if nprobe == 44:
        GD, GI = O.bruteforce_f64(xb, xq, 10)
        np.testing.assert_array_equal(I, GI)
        np.testing.assert_array_equal(D.astype(np.float64), GD)


def test_cli_end_to_end(tmp_path):
    bench = tmp_path / "bench"
    origin = bench / "sift10k" / "origin"
    origin.mkdir(parents=True)
    (origin / "sift10k_base.fvecs").write_bytes(open(os.path.join(GOLD, "sift10k_base_2000.fvecs"), "rb").read())
    (origin / "sift10k_query.fvecs").write_bytes(open(os.path.join(GOLD, "sift10k_query.fvecs"), "rb").read())
    argv = ["--benchmarks_path", str(bench), "--dataset", "sift10k", "--k", "10", "--nprobes", "4", "0",
            "--mode", "base", "--kmeans_niter", "10", "--kmeans_seed", "0", "--loop", "2"]
    rows = query.run(argv)
    assert [r["nprobe"] for r in rows] == [4, 44] and rows[-1]["recall"] == 1.0
    assert os.path.exists(bench / "sift10k" / "index" / "index_nlist_44_opt_0_subNprobeRatio_1.index")
    gt = bench / "sift10k" / "result" / "groundtruth_10.bin"
    assert os.path.getsize(gt) == 100 * 10 * 12
    GI, GD = H.load_results(gt, 100, 10)
    xb = H.load_xvecs(str(origin / "sift10k_base.fvecs"))
    xq = H.load_xvecs(str(origin / "sift10k_query.fvecs"))
    _, BI = O.bruteforce_f64(xb, xq, 10)
    np.testing.assert_array_equal(GI, BI)
    lines = (bench / "sift10k" / "result" / "log.csv").read_text().splitlines()
    assert lines[0].startswith("dataset,mode,nb,nq,d,nlist,nprobe,k,") and len(lines) == 3
    assert lines[2].startswith("sift10k,Baseline,2000,100,128,44,44,10,0,1,1,1,0,0,1,1,")
    # --cache: the saved index is loaded, the ground truth file is read back; same recall
    rows2 = query.run(argv + ["--cache"])
    assert [r["recall"] for r in rows2] == [r["recall"] for r in rows]
    # the unpruned mode over every list: the exact kNN
    rows3 = query.run(argv + ["--cache", "--mode", "original", "--nprobes", "44"])
    assert rows3[0]["recall"] == 1.0 and rows3[0]["variance"] == 0.0


def test_cli_inner_product(tmp_path):
    """--metric ip with OPT_TRIANGLE: index built and saved with candidate2centroid, ground truth
    from the full-probe scan, recall 1 at nprobe = nlist, and the same recall with the cut off."""
    bench = tmp_path / "bench"
    origin = bench / "unit" / "origin"
    origin.mkdir(parents=True)
    rng = np.random.default_rng(4)
    xb = rng.normal(size=(3000, 32)).astype(np.float32)
    xb /= np.linalg.norm(xb, axis=1, keepdims=True)
    xq = xb[rng.integers(0, 3000, size=50)] + 0.05 * rng.normal(size=(50, 32)).astype(np.float32)
    xq /= np.linalg.norm(xq, axis=1, keepdims=True)
    H.write_fvecs(origin / "unit_base.fvecs", xb)
    H.write_fvecs(origin / "unit_query.fvecs", xq)
    argv = ["--benchmarks_path", str(bench), "--dataset", "unit", "--k", "10", "--metric", "ip", "--nprobes", "3",
            "0", "--opt_levels", "OPT_TRIANGLE", "--kmeans_niter", "6", "--kmeans_seed", "1"]
    rows = query.run(argv)
    assert [r["nprobe"] for r in rows] == [3, 54] and rows[-1]["recall"] == 1.0
    ipath = bench / "unit" / "index" / "index_nlist_54_opt_1_subNprobeRatio_1.index"
    g = Index(32, 54)
    g.load_index(str(ipath))
    c2c = g.list_c2c()
    ids = g.list_ids()
    off = g.list_offsets()
    cent = g.centroids()
    for l in range(54):
        a, b = off[l], off[l + 1]
        np.testing.assert_allclose(c2c[a:b], xb[ids[a:b]] @ cent[l], rtol=1e-5, atol=1e-6)
    GI, _ = H.load_results(bench / "unit" / "result" / "groundtruth_10.bin", 50, 10)
    s = xq.astype(np.float64) @ xb.astype(np.float64).T
    assert np.mean(GI[:, 0] == np.argmax(s, axis=1)) == 1.0
    rows2 = query.run(argv + ["--cache", "--opt_levels", "0"])  # the opt-1 file, searched without the cut
    assert [r["recall"] for r in rows2] == [r["recall"] for r in rows]
