"""The reference index file format (Index::save_index / load_index, Index.cpp:1708-1770;
IVF::save_IVF / load_IVF, IVF.cpp:76-134) through hv_save / hv_load: files are parsed and
written here with numpy from the layout those functions define."""
import struct

import numpy as np
import pytest

from harmony_amd import datagen
from harmony_amd._lib import HvError
from harmony_amd.index import Index
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def read_reference_index(path):
    buf = open(path, "rb").read()
    pos = 0

    def take(fmt):
        nonlocal pos
        v = struct.unpack_from(fmt, buf, pos)
        pos += struct.calcsize(fmt)
        return v

    d, nlist = take("<QQ")
    metric, opt = take("<ii")
    take("<QQQ")
    cent = np.frombuffer(buf, np.float32, nlist * d, pos).reshape(nlist, d)
    pos += 4 * nlist * d
    lists = {}
    while pos < len(buf):
        (lid,) = take("<Q")
        n, dd, sk = take("<QQQ")
        (lopt,) = take("<i")
        ids = np.frombuffer(buf, np.uint64, n, pos).astype(np.int64)
        pos += 8 * n
        codes = np.frombuffer(buf, np.float32, n * dd, pos).reshape(n, dd)
        pos += 4 * n * dd
        norms = np.frombuffer(buf, np.float32, n, pos)
        pos += 4 * n
        assert lopt == 0
        lists[lid] = (ids, codes, norms)
    return d, nlist, metric, opt, cent, lists


def write_reference_index(path, cent, lists, opt=1):
    """As the reference writes it, with the TRIANGLE per-list arrays (opt bit 1) present."""
    nlist, d = cent.shape
    with open(path, "wb") as f:
        f.write(struct.pack("<QQii QQQ".replace(" ", ""), d, nlist, 1, opt, 0, 1, 1))
        f.write(np.ascontiguousarray(cent, np.float32).tobytes())
        for lid in sorted(lists):
            ids, codes = lists[lid]
            n = ids.size
            f.write(struct.pack("<QQQQi", lid, n, d, 0, opt))
            f.write(ids.astype(np.uint64).tobytes())
            f.write(np.ascontiguousarray(codes, np.float32).tobytes())
            f.write((codes.astype(np.float32) ** 2).sum(1).astype(np.float32).tobytes())
            if opt & 1:
                f.write(np.ones(n, np.float32).tobytes())  # candidate2centroid
                f.write(np.ones(n, np.float32).tobytes())  # sqrt_candidate2centroid


@pytest.fixture(scope="module")
def built():
    xb, xq = datagen.make(1)
    g = Index(128, 100, nprobe=10)
    g.train(xb, niter=10, seed=0)
    g.add(xb)
    return xb, xq, g


def test_save_writes_reference_layout(built, tmp_path):
    xb, xq, g = built
    p = tmp_path / "w1.index"
    g.save_index(p)
    d, nlist, metric, opt, cent, lists = read_reference_index(p)
    assert (d, nlist, metric) == (128, 100, 1)
    np.testing.assert_array_equal(cent, g.centroids())
    off, ids, codes = g.list_offsets(), g.list_ids(), g.list_codes()
    assert sorted(lists) == [l for l in range(100) if off[l + 1] > off[l]]
    for l, (lids, lcodes, norms) in lists.items():
        np.testing.assert_array_equal(lids, ids[off[l]:off[l + 1]])
        np.testing.assert_array_equal(lcodes, codes[off[l]:off[l + 1]])
        np.testing.assert_array_equal(norms, (lcodes.astype(np.float64) ** 2).sum(1).astype(np.float32))


def test_load_round_trip_searches_identically(built, tmp_path):
    xb, xq, g = built
    p = tmp_path / "w1.index"
    g.save_index(p)
    h = Index(8, 3)  # load replaces d and nlist with the file's
    h.load_index(p)
    assert (h.d, h.nlist, h.ntotal) == (128, 100, 10000)
    for idx in (g, h):
        idx.nprobe = 10
        idx.preSearch(pruning=True, dim_blocks=4)
    D1, I1 = g.search(xq, 10)
    D2, I2 = h.search(xq, 10)
    np.testing.assert_array_equal(I1, I2)
    np.testing.assert_array_equal(D1, D2)


def test_load_reference_written_file_with_optional_arrays(tmp_path):
    rng = np.random.default_rng(4)
    cent = rng.integers(0, 100, size=(6, 16)).astype(np.float32)
    lists = {}
    nxt = 0
    for l in (0, 2, 3, 5):  # lists 1 and 4 empty: not stored
        n = int(rng.integers(1, 30))
        lists[l] = (np.arange(nxt, nxt + n), rng.integers(0, 100, size=(n, 16)).astype(np.float32))
        nxt += n
    p = tmp_path / "ref.index"
    write_reference_index(p, cent, lists, opt=1)
    h = Index(16, 6)
    h.load_index(p)
    assert h.ntotal == nxt
    off = h.list_offsets()
    for l in range(6):
        n = lists[l][0].size if l in lists else 0
        assert off[l + 1] - off[l] == n
    q = rng.integers(0, 100, size=(7, 16)).astype(np.float32)
    h.nprobe = 6
    h.preSearch(pruning=True, dim_blocks=2)
    D, I = h.search(q, 5)
    xb = np.zeros((nxt, 16), np.float32)
    for l, (ids, codes) in lists.items():
        xb[ids] = codes
    GD, GI = O.bruteforce_f64(xb, q, 5)
    np.testing.assert_array_equal(I, GI)


def test_load_rejects_truncated_file(built, tmp_path):
    xb, xq, g = built
    p = tmp_path / "w1.index"
    g.save_index(p)
    data = open(p, "rb").read()
    t = tmp_path / "trunc.index"
    open(t, "wb").write(data[: len(data) - 1000])
    h = Index(128, 100)
    with pytest.raises(HvError):
        h.load_index(t)
    with pytest.raises(HvError):
        h.load_index(tmp_path / "missing.index")


def test_load_rejects_truncated_optional_arrays(tmp_path):
    """A file cut inside the last list's optional arrays (IVF.cpp:117-122) is an I/O error, not a
    silent seek past the end."""
    rng = np.random.default_rng(5)
    cent = rng.integers(0, 100, size=(4, 8)).astype(np.float32)
    lists = {l: (np.arange(10 * l, 10 * l + 10), rng.integers(0, 100, size=(10, 8)).astype(np.float32))
             for l in range(4)}
    p = tmp_path / "opt.index"
    write_reference_index(p, cent, lists, opt=1)
    data = open(p, "rb").read()
    for cut in (4, 40, 79):  # inside sqrt_candidate2centroid / candidate2centroid of list 3
        t = tmp_path / ("cut%d.index" % cut)
        open(t, "wb").write(data[: len(data) - cut])
        h = Index(8, 4)
        with pytest.raises(HvError, match="truncated"):
            h.load_index(t)


def test_failed_load_leaves_handle_intact(built, tmp_path):
    xb, xq, g = built
    p = tmp_path / "w1.index"
    g.save_index(p)
    h = Index(128, 100)
    h.load_index(p)
    h.nprobe = 10
    h.preSearch(pruning=True, dim_blocks=4)
    D0, I0 = h.search(xq, 10)
    data = open(p, "rb").read()
    t = tmp_path / "trunc.index"
    open(t, "wb").write(data[: len(data) // 2])
    with pytest.raises(HvError):
        h.load_index(t)
    assert (h.ntotal, h.nlist) == (10000, 100)
    D1, I1 = h.search(xq, 10)
    np.testing.assert_array_equal(I0, I1)
    np.testing.assert_array_equal(D0, D1)


def test_load_resets_search_opt_level(tmp_path):
    """load_index sets opt_level = OPT_NONE and keeps the file's added_opt_level (Index.cpp:1742-1743):
    a search at OPT_NONE passes the subset check, hv_save writes the file's level back."""
    rng = np.random.default_rng(6)
    cent = rng.integers(0, 100, size=(3, 8)).astype(np.float32)
    lists = {l: (np.arange(5 * l, 5 * l + 5), rng.integers(0, 100, size=(5, 8)).astype(np.float32))
             for l in range(3)}
    p = tmp_path / "opt.index"
    write_reference_index(p, cent, lists, opt=1)
    h = Index(8, 3)
    h.opt_level = 1
    h.load_index(p)
    assert h.opt_level == 0
    h.nprobe = 3
    h.preSearch(pruning=False)
    h.search(cent, 2)
    q = tmp_path / "again.index"
    h.save_index(q)
    assert read_reference_index(q)[3] == 1
