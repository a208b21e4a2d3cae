"""Point-cloud readers (xyz, PCD, binvox) through the C ABI against the oracle and
against hand-built files whose points are known in closed form.  Host code only."""
import os
import shutil
import struct

import numpy as np
import pytest

from oracle import files as F

HERE = os.path.dirname(os.path.abspath(__file__))
MODELS = os.path.join(HERE, "golden", "models")
_seq = [0]


@pytest.fixture
def work():
    # format detection searches the whole path for "xyz" / "pcd" / "binvox", so the
    # scratch directory gets a name free of them (pytest's tmp_path embeds test names)
    _seq[0] += 1
    d = f"/tmp/gvx_files_{os.getpid()}_{_seq[0]}"
    os.makedirs(d, exist_ok=True)
    yield d
    shutil.rmtree(d, ignore_errors=True)


@pytest.fixture(scope="module")
def g():
    import gvx
    return gvx


def _put(path, data):
    with open(path, "wb") as f:
        f.write(data if isinstance(data, bytes) else data.encode())
    return path


def _same(a, b):
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b, equal_nan=True)


# ---------------------------------------------------------------- closed forms
def test_binvox_voxel_order(g, work):
    # dim d h w = 2 3 4, scale 8 -> 8 / w = 2, one voxel at index 19 = x 1, z 1, y 3
    # (y fastest, then z, then x); a voxel of value 2 is not a point; trailing byte ignored
    body = bytes([0, 19, 1, 1, 0, 4]) + b"\x07"
    p = _put(f"{work}/one.binvox", b"#binvox 1\ndim 2 3 4\ntranslate 0.5 -1 3\nscale 8\ndata\n" + body)
    want = np.array([[2 * 1 + 0.5, 2 * 3 - 1, 2 * 1 + 3]], np.float32)
    assert _same(F.read_binvox(p), want)
    assert _same(g.load_point_cloud(p), want)
    p2 = _put(f"{work}/two.binvox", b"#binvox 1\ndim 1 1 3\nfoo bar baz\ntranslate 0 0 0\nscale 3\ndata\n" +
              bytes([2, 1, 1, 2]))
    want2 = np.array([[0, 1, 0], [0, 2, 0]], np.float32)
    assert _same(F.read_binvox(p2), want2)
    assert _same(g.load_point_cloud(p2), want2)


def test_binvox_runs_past_end_rejected(g, work):
    p = _put(f"{work}/bad.binvox", b"#binvox 1\ndim 1 1 2\ntranslate 0 0 0\nscale 1\ndata\n" + bytes([1, 3]))
    with pytest.raises(g.GvxError):
        g.load_point_cloud(p)
    with pytest.raises(ValueError):
        F.read_binvox(p)


def test_text_triples_per_line(g, work):
    p = _put(f"{work}/pts.xyz", "1 2 3 4 5 6 7\n8 9 10\n\n1e-1 -2.5 +3\n0.25 0.5\n")
    want = np.array([[1, 2, 3], [4, 5, 6], [8, 9, 10], [np.float32(0.1), -2.5, 3]], np.float32)
    assert _same(F.read_xyz(p), want)
    assert _same(g.load_point_cloud(p), want)


def _pcd_head(fields, sizes, types, n, mode):
    return (f"# .PCD v0.7\nVERSION 0.7\nFIELDS {fields}\nSIZE {sizes}\nTYPE {types}\nCOUNT {' '.join(['1'] * len(fields.split()))}\n"
            f"WIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\nDATA {mode}\n").encode()


def test_pcd_ascii_binary_compressed_by_hand(g, work):
    pts = np.array([[1.5, -2.25, 3.0], [np.nan, np.nan, np.nan]], np.float32)
    rgba = [4278190080, 7]
    a = _put(f"{work}/a.pcd", _pcd_head("rgba x y z", "4 4 4 4", "U F F F", 2, "ascii") +
             b"4278190080 1.5 -2.25 3\n7 nan nan nan\n")
    b = _put(f"{work}/b.pcd", _pcd_head("x y z rgba", "4 4 4 4", "F F F U", 2, "binary") +
             b"".join(struct.pack("<fffI", *p, c) for p, c in zip(pts.tolist(), rgba)))
    for p in (a, b):
        assert _same(F.read_pcd(p), pts)
        assert _same(g.load_point_cloud(p), pts)
    # binary_compressed: fields stored one after the other, then LZF; 24 zero bytes as
    # one literal byte and one back-reference (length 23 = 7 + 14 + 2, offset 0)
    raw_len = 2 * 3 * 4
    stream = bytes([0x00, 0x00, 0xE0, 14, 0x00])
    c = _put(f"{work}/c.pcd", _pcd_head("x y z", "4 4 4", "F F F", 2, "binary_compressed") +
             struct.pack("<II", len(stream), raw_len) + stream)
    zeros = np.zeros((2, 3), np.float32)
    assert _same(F.read_pcd(c), zeros)
    assert _same(g.load_point_cloud(c), zeros)
    # fields of other widths: x as float64, y as int16, z as uint8
    d = _put(f"{work}/d.pcd", _pcd_head("x y z", "8 2 1", "F I U", 1, "binary") + struct.pack("<dhB", 0.125, -3, 200))
    want = np.array([[0.125, -3, 200]], np.float32)
    assert _same(F.read_pcd(d), want)
    assert _same(g.load_point_cloud(d), want)


def _head_raw(fields, sizes, types, counts, points, mode):
    return (f"VERSION 0.7\nFIELDS {fields}\nSIZE {sizes}\nTYPE {types}\nCOUNT {counts}\nWIDTH 1\nHEIGHT 1\n"
            f"POINTS {points}\nDATA {mode}\n").encode()


@pytest.mark.parametrize("case", [
    # POINTS near 2^64 / 12: n * stride and 3 * n wrap around 64 bits
    ("x y z", "4 4 4", "F F F", "1 1 1", str((1 << 64) // 12 + 1), "binary"),
    ("x y z", "4 4 4", "F F F", "1 1 1", str((1 << 64) // 3 + 5), "binary"),
    ("x y z", "4 4 4", "F F F", "1 1 1", str((1 << 62)), "ascii"),
    ("x y z", "4 4 4", "F F F", "1 1 1", str((1 << 64) // 12 + 1), "binary_compressed"),
    # zero / negative sizes and counts
    ("x y z", "0 4 4", "F F F", "1 1 1", "1", "binary"),
    ("x y z", "4 -4 4", "F F F", "1 1 1", "1", "binary"),
    ("x y z", "4 4 4", "F F F", "1 0 1", "1", "binary"),
    ("x y z", "4 4 4", "F F F", "1 1 -3", "1", "binary"),
    # a huge COUNT that makes the stride wrap
    ("x y z", "4 4 8", "F F F", "1 1 2305843009213693952", "1", "binary"),
    # more points than the file can hold
    ("x y z", "4 4 4", "F F F", "1 1 1", "1000", "binary"),
])
def test_pcd_malformed_headers_rejected(g, work, case):
    """Header values are checked before anything is allocated or read: sizes and counts must be
    positive, n * stride must not wrap and must fit in the rest of the file.  Each file carries
    one real point (12 bytes) after the header."""
    p = _put(f"{work}/m.pcd", _head_raw(*case) + struct.pack("<fff", 1, 2, 3) + struct.pack("<II", 12, 1 << 31))
    with pytest.raises(g.GvxError):
        g.load_point_cloud(p)


def test_binvox_huge_dims_rejected(g, work):
    # dims whose product the run-length data cannot cover (2 pairs = 510 voxels at most)
    p = _put(f"{work}/h.binvox", b"#binvox 1\ndim 100000 100000 100000\ntranslate 0 0 0\nscale 1\ndata\n" +
             bytes([1, 255, 0, 255]))
    with pytest.raises(g.GvxError):
        g.load_point_cloud(p)


def test_lzf_round_trip_of_the_oracle():
    rng = np.random.default_rng(5)
    for n in (0, 1, 31, 33, 1000, 70000):
        src = bytes(rng.integers(0, 4, n, dtype=np.uint8))  # repetitive: many back-references
        assert F.lzf_decompress(F.lzf_compress(src), n) == src


# ---------------------------------------------------------------- writers and options
@pytest.mark.parametrize("mode", ["ascii", "binary", "binary_compressed"])
def test_pcd_encodings_round_trip(g, work, mode):
    rng = np.random.default_rng(7)
    pts = rng.normal(0, 3, (5000, 3)).astype(np.float32)
    pts[17] = np.nan
    p = f"{work}/r_{mode}.pcd"
    F.write_pcd(p, pts, mode, extra=rng.integers(0, 2 ** 32, 5000, dtype=np.uint64).astype(np.uint32))
    assert _same(F.read_pcd(p), pts)
    assert _same(g.load_point_cloud(p), pts)


def test_binvox_writer_round_trip(g, work):
    rng = np.random.default_rng(8)
    occ = (rng.random((9, 7, 5)) < 0.3).astype(np.uint8)   # [x, z, y]
    p = f"{work}/r.binvox"
    F.write_binvox(p, occ, translate=(0.25, -1.5, 2.0), scale=10.0)
    x, z, y = np.nonzero(occ)
    s = np.float32(10.0) / np.float32(5)
    want = np.stack([s * x.astype(np.float32) + np.float32(0.25), s * y.astype(np.float32) + np.float32(-1.5),
                     s * z.astype(np.float32) + np.float32(2.0)], axis=1).astype(np.float32)
    order = np.lexsort((y, z, x))
    assert _same(g.load_point_cloud(p), want[order])
    assert _same(F.read_binvox(p), want[order])


def test_shift_scale_offset_and_model_path(g, work, monkeypatch):
    p = _put(f"{work}/o.xyz", "1 5 -2\n3 -4 7\n")
    got = g.load_point_cloud(p, shift_to_zero=True, offset=(0.5, 1.0, -1.0), scaling=2.0)
    base = np.array([[1, 5, -2], [3, -4, 7]], np.float32)
    want = (np.float32(2.0) * (base - base.min(axis=0)) + np.array([0.5, 1.0, -1.0], np.float32)).astype(np.float32)
    assert _same(got, want)
    assert _same(F.load_point_cloud(p, shift_to_zero=True, offset=(0.5, 1.0, -1.0), scaling=2.0), want)
    monkeypatch.setenv("GPU_VOXELS_MODEL_PATH", work)
    assert _same(g.load_point_cloud("o.xyz", use_model_path=True), base)
    with pytest.raises(g.GvxError):
        g.load_point_cloud(f"{work}/o.unknown")
    with pytest.raises(g.GvxError):
        g.load_point_cloud(f"{work}/missing.binvox")


def test_format_chosen_by_first_substring(g, work):
    # a PCD file under a directory whose name contains "xyz" is read by the text reader,
    # as the reference's substring test does: only all-numeric lines give points
    d = f"{work}/xyz_dir"
    os.makedirs(d)
    p = f"{d}/c.pcd"
    F.write_pcd(p, np.array([[1, 2, 3], [4, 5, 6]], np.float32), "ascii")
    got = g.load_point_cloud(p)
    assert _same(got, F.load_point_cloud(p))
    assert _same(got, np.array([[1, 2, 3], [4, 5, 6]], np.float32))


# ---------------------------------------------------------------- the reference's model files
@pytest.mark.parametrize("name", ["arm_6_link.binvox", "coordinate_system_100.binvox", "arm_6_link.xyz",
                                  "hollie_from_pointcloud2.pcd"])
def test_reference_models(g, name):
    p = os.path.join(MODELS, name)
    got = g.load_point_cloud(p)
    assert _same(got, F.load_point_cloud(p))
    n = {"arm_6_link.binvox": 250, "coordinate_system_100.binvox": 11777, "arm_6_link.xyz": 610,
         "hollie_from_pointcloud2.pcd": 3151}[name]
    assert got.shape == (n, 3)
