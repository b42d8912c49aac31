"""The harness file formats (src/utils.h:187-407, 565-620; src/stats.h:122-134) on CPU: loaders
against hand-built files and the reference's own benchmark files (tests/golden/harness, byte
copies made by tests/golden/make_harness_fixtures.py), result files, the log.csv row, the path
layout and the CLI's --dataset_info leg (no GPU needed)."""
import os
import struct

import numpy as np
import pytest

from harmony_amd import harness as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "harness")
REF = "/root/reference/benchmarks"


def test_fvecs_round_trip_and_bounds(tmp_path):
    x = np.arange(5 * 3, dtype=np.float32).reshape(5, 3) * 0.5
    p = tmp_path / "a.fvecs"
    H.write_fvecs(p, x)
    assert os.path.getsize(p) == 5 * (4 + 12)
    assert H.xvecs_info(str(p)) == (5, 3)
    np.testing.assert_array_equal(H.load_fvecs(p), x)
    np.testing.assert_array_equal(H.load_fvecs(p, (2, 4)), x[1:4])   # 1-based, inclusive
    np.testing.assert_array_equal(H.load_fvecs(p, (5, 0)), x[4:])    # b = 0: the last record
    with pytest.raises(ValueError):
        H.load_fvecs(p, (0, 2))
    with pytest.raises(ValueError):
        H.load_fvecs(p, (2, 6))
    with open(p, "ab") as f:  # a trailing partial record is ignored
        f.write(b"\x01\x02")
    assert H.load_fvecs(p).shape == (5, 3)


def test_bvecs_records_and_reference_offset(tmp_path):
    rows = np.array([[1, 2, 3, 4, 5, 6], [7, 8, 9, 10, 11, 12], [13, 14, 15, 16, 17, 18]], np.uint8)
    p = tmp_path / "b.bvecs"
    H.write_bvecs(p, rows)
    np.testing.assert_array_equal(H.load_bvecs(p), rows.astype(np.float32))
    assert H.xvecs_info(str(p)) == (3, 6)
    # the reference's reads: n = (30 - 4) // 10 = 2 records, each 4 bytes past its data
    # (record 0 reads bytes 8..13 = its data 5, 6 then record 1's header int 6 = 6, 0, 0, 0)
    ref = H.load_bvecs(p, reference_offset=True)
    np.testing.assert_array_equal(ref, np.array([[5, 6, 6, 0, 0, 0], [11, 12, 6, 0, 0, 0]], np.float32))


def test_txt_drops_first_field_and_ignores_bounds(tmp_path):
    p = tmp_path / "t.txt"
    p.write_text("7,1.5,2,3\n8,4,5,6.25\n")
    assert H.xvecs_info(str(p)) == (2, 3)
    np.testing.assert_array_equal(H.load_xvecs(str(p), (2, 2)), np.array([[1.5, 2, 3], [4, 5, 6.25]], np.float32))
    with pytest.raises(ValueError):
        H.load_xvecs(str(tmp_path / "x.ivecs"))


def test_reference_malformed_test_file():
    """benchmarks/test/origin/test_base.fvecs holds an extra int after d: the loader reads what
    loadFvecs reads from it (5 records of 20 values, each starting with that int's bits)."""
    x = H.load_fvecs(os.path.join(GOLD, "test_base.fvecs"))
    assert x.shape == (5, 20)
    assert x[0, 0].view(np.int32) == 20
    np.testing.assert_array_equal(x[0, 1:], np.arange(1, 20, dtype=np.float32))


def test_reference_sift10k_and_iris_fixtures():
    b = H.load_fvecs(os.path.join(GOLD, "sift10k_base_2000.fvecs"))
    q = H.load_fvecs(os.path.join(GOLD, "sift10k_query.fvecs"))
    assert b.shape == (2000, 128) and q.shape == (100, 128)
    assert np.all(b == np.round(b)) and b.min() >= 0  # integer-valued SIFT descriptors
    raw = np.fromfile(os.path.join(GOLD, "sift10k_base_2000.fvecs"), np.float32).reshape(2000, 129)
    np.testing.assert_array_equal(b, raw[:, 1:])
    assert np.all(raw[:, 0].view(np.int32) == 128)
    iris = H.load_fvecs(os.path.join(GOLD, "iris_base.fvecs"))
    assert iris.shape == (150, 4)
    if os.path.isdir(REF):  # the slice is the reference's own first 2,000 records
        full = H.load_fvecs(os.path.join(REF, "sift10k", "origin", "sift10k_base.fvecs"))
        np.testing.assert_array_equal(full[:2000], b)


def test_results_files_round_trip(tmp_path):
    I = np.array([[3, 1], [0, -1]], np.int64)
    D = np.array([[0.5, 2.25], [1.0, np.finfo(np.float32).max]], np.float32)
    pb = tmp_path / "r" / "groundtruth_2.bin"
    H.write_results(pb, I, D)
    assert os.path.getsize(pb) == 2 * 2 * 12
    raw = open(pb, "rb").read()
    assert struct.unpack("<4q", raw[:32]) == (3, 1, 0, -1)
    assert struct.unpack("<4f", raw[32:]) == (0.5, 2.25, 1.0, float(np.finfo(np.float32).max))
    I2, D2 = H.load_results(pb, 2, 2)
    np.testing.assert_array_equal(I2, I)
    np.testing.assert_array_equal(D2, D)
    pt = tmp_path / "groundtruth_2.txt"
    H.write_results(pt, I[:1], D[:1])
    assert pt.read_text() == "3 0.500000 1 2.250000\n"
    I3, D3 = H.load_results(pt, 1, 2)
    np.testing.assert_array_equal(I3, I[:1])


def test_log_row_format(tmp_path):
    row = {"dataset": "sift10k", "mode": "Baseline", "nb": 2000, "nq": 100, "d": 128, "nlist": 44, "nprobe": 4,
           "k": 10, "divideIVF": False, "orderOptimize": True, "UnblockSend": True, "cut": True, "block": 0,
           "worker": 0, "group": 1, "team": 1, "time_speedup": 0.0, "query_time": 0.001, "original_time": 0.0,
           "recall": float(np.float32(0.95)), "r2": 0.0, "variance": 1.5, "train_time": 0.25, "add_time": 0.125,
           "preSearch_time": 0.0, "ratio": 0.0, "teamRatio": 0.0, "brute_ratio": 1.0}
    line = H.log_row(row)
    assert line == ("sift10k,Baseline,2000,100,128,44,4,10,0,1,1,1,0,0,1,1,0.0000000,0.0010000000000000,"
                    "0.0000000000000000,0.9499999880790710,0.0000000000000000,1.5000000000000000,"
                    "0.2500000000000000,0.1250000000000000,0.0000000000000000,0.0000000000000000,"
                    "0.0000000000000000,1.0000000000000000")
    p = tmp_path / "res" / "log.csv"
    H.append_log(p, row)
    H.append_log(p, row)
    lines = p.read_text().splitlines()
    assert lines[0].split(",") == [c for c, _ in H.LOG_COLUMNS] and len(lines) == 3


def test_paths_and_probe_variance():
    p = H.harness_paths("/b", "sift10k", "fvecs", "bin", 100, nlist=100, opt=1, sub_nprobe_ratio=1.0)
    assert p["base"] == "/b/sift10k/origin/sift10k_base.fvecs"
    assert p["groundtruth"] == "/b/sift10k/result/groundtruth_100.bin"
    assert p["index"] == "/b/sift10k/index/index_nlist_100_opt_1_subNprobeRatio_1.index"
    assert H.harness_paths("/b", "x", sub_nprobe_ratio=0.5)["index"].endswith("_subNprobeRatio_0.5.index")
    # lists probed 3, 1, 0, 0 times: mean 1, variance (4 + 0 + 1 + 1) / 4 = 1.5
    assert H.probe_variance(np.array([[0, 0], [0, 1]]), 4) == pytest.approx(1.5 ** 0.5)


def test_cli_dataset_info(tmp_path):
    from harmony_amd import query
    root = tmp_path / "bench" / "sift10k" / "origin"
    root.mkdir(parents=True)
    for src, dst in (("sift10k_base_2000.fvecs", "sift10k_base.fvecs"), ("sift10k_query.fvecs", "sift10k_query.fvecs")):
        (root / dst).write_bytes(open(os.path.join(GOLD, src), "rb").read())
    out = tmp_path / "info.csv"
    assert query.run(["--benchmarks_path", str(tmp_path / "bench"), "--dataset", "sift10k", "--dataset_info",
                      "--csv", str(out)]) == 0
    assert out.read_text() == "nb, nq, d\n2000, 100, 128\n"
    with pytest.raises(SystemExit):
        query.run(["--mode", "group"])  # no single-GPU meaning: not offered
