"""Copies the reference's own benchmark data files (data, not code) into tests/golden/harness/:
the first 2,000 records of benchmarks/sift10k/origin/sift10k_base.fvecs, all 100 records of
sift10k_query.fvecs (byte-identical record slices, so the fvecs loader reads the reference's own
bytes), and the whole iris and test datasets. Run in a checkout that has the reference at
/root/reference: python tests/golden/make_harness_fixtures.py"""
import os

SRC = "/root/reference/benchmarks"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness")


def copy(rel, dst_name, max_bytes=None):
    with open(os.path.join(SRC, rel), "rb") as f:
        data = f.read()
    if max_bytes is not None:
        data = data[:max_bytes]
    with open(os.path.join(OUT, dst_name), "wb") as f:
        f.write(data)
    print(dst_name, len(data))


if __name__ == "__main__":
    os.makedirs(OUT, exist_ok=True)
    rec = 4 + 4 * 128
    copy("sift10k/origin/sift10k_base.fvecs", "sift10k_base_2000.fvecs", 2000 * rec)
    copy("sift10k/origin/sift10k_query.fvecs", "sift10k_query.fvecs")
    copy("iris/origin/iris_base.fvecs", "iris_base.fvecs")
    copy("iris/origin/iris_query.fvecs", "iris_query.fvecs")
    copy("test/origin/test_base.fvecs", "test_base.fvecs")
