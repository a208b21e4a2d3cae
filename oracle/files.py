"""Oracle for the file formats on the path (test infrastructure only: imported by
tests/, never by the library).

A numpy restatement of
  * PointcloudFileHandler::loadPointCloud   helpers/PointcloudFileHandler.cpp:69-137
    (format by the first of "xyz"/"pcd"/"binvox" in the path, optional shift of the
    minimum to zero, p = scaling * p + offset, all in float32),
  * XyzFileReader::readPointCloud           helpers/XyzFileReader.cpp:29-56
  * BinvoxFileReader::readPointCloud        helpers/BinvoxFileReader.cpp:42-163
  * PcdFileReader::readPointCloud           helpers/PcdFileReader.cpp:31-54 (pcl::PointXYZ
    from ascii, binary and binary_compressed PCD data),
  * the map and list dump layouts of TemplateVoxelMap::writeToDisk
    (voxelmap/TemplateVoxelMap.hpp:906-954) and TemplateVoxelList::writeToDisk
    (voxellist/TemplateVoxelList.hpp:625-658),
plus writers used by the tests to build inputs (binvox run-length coding, PCD in its
three encodings with an LZF compressor).  Decimal numbers are converted with the C
library's strtof, as the C++ stream extraction does.
"""
import ctypes
import os
import re
import struct

import numpy as np

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]
_NUM = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$")

MT_BITVECTOR_VOXELMAP, MT_BITVECTOR_VOXELLIST, MT_PROBAB_VOXELMAP, MT_DISTANCE_VOXELMAP = 0, 1, 4, 9


def strtof(tok):
    return np.float32(_libc.strtof(tok.encode(), None))


# ------------------------------------------------------------------ readers
def read_xyz(path):
    """Numbers of a line are taken three at a time; one or two left over are dropped."""
    pts = []
    with open(path, "rb") as f:
        for raw in f.read().decode("latin-1").split("\n"):
            vals = []
            for tok in raw.split():
                if not _NUM.match(tok):
                    break
                vals.append(strtof(tok))
            for k in range(len(vals) // 3):
                pts.append(vals[3 * k:3 * k + 3])
    return np.array(pts, np.float32).reshape(-1, 3)


def read_binvox(path):
    b = open(path, "rb").read()
    pos = 0

    def token():
        nonlocal pos
        while pos < len(b) and b[pos:pos + 1].isspace():
            pos += 1
        s = pos
        while pos < len(b) and not b[pos:pos + 1].isspace():
            pos += 1
        return b[s:pos].decode("latin-1")

    if token() != "#binvox":
        raise ValueError("not a binvox file")
    int(token())
    dims = None
    t = np.zeros(3, np.float32)
    scale = np.float32(0)
    while True:
        k = token()
        if k == "data":
            break
        if k == "":
            raise ValueError("binvox header without data")
        if k == "dim":
            dims = [int(token()) for _ in range(3)]
        elif k == "translate":
            t = np.array([strtof(token()) for _ in range(3)], np.float32)
        elif k == "scale":
            scale = np.float32(strtof(token()) / np.float32(dims[2] if dims else 0))
        else:
            nl = b.find(b"\n", pos)
            pos = len(b) if nl < 0 else nl + 1
    d, h, w = dims
    size = d * h * w
    data = np.frombuffer(b[pos + 1:], np.uint8)
    vals, cnts = data[0:(len(data) // 2) * 2:2], data[1:(len(data) // 2) * 2:2].astype(np.int64)
    ends = np.cumsum(cnts)
    keep = np.searchsorted(ends, size, side="left") + 1  # pairs up to the one reaching size
    vals, cnts = vals[:keep], cnts[:keep]
    if cnts.sum() > size:
        raise ValueError("binvox run past the end")
    vox = np.zeros(size, np.uint8)
    vox[:cnts.sum()] = np.repeat(vals, cnts)
    idx = np.nonzero(vox == 1)[0]
    x, z, y = idx // (w * h), (idx // w) % h, idx % w       # y fastest, then z, then x
    pts = np.stack([scale * x.astype(np.float32) + t[0], scale * y.astype(np.float32) + t[1],
                    scale * z.astype(np.float32) + t[2]], axis=1)
    return pts.astype(np.float32)


def lzf_decompress(src, n_out):
    out = bytearray()
    i = 0
    while i < len(src):
        c = src[i]
        i += 1
        if c < 32:
            out += src[i:i + c + 1]
            i += c + 1
        else:
            ln = c >> 5
            if ln == 7:
                ln += src[i]
                i += 1
            ref = len(out) - ((c & 0x1F) << 8) - src[i] - 1
            i += 1
            for _ in range(ln + 2):
                out.append(out[ref])
                ref += 1
    if len(out) != n_out:
        raise ValueError("LZF size mismatch")
    return bytes(out)


_PCD_NP = {("F", 4): "<f4", ("F", 8): "<f8", ("I", 1): "i1", ("I", 2): "<i2", ("I", 4): "<i4", ("I", 8): "<i8",
           ("U", 1): "u1", ("U", 2): "<u2", ("U", 4): "<u4", ("U", 8): "<u8"}


def read_pcd(path):
    b = open(path, "rb").read()
    hdr = {}
    pos = 0
    while True:
        nl = b.find(b"\n", pos)
        line = b[pos:nl].decode("latin-1").strip()
        pos = nl + 1
        if not line or line.startswith("#"):
            continue
        k, *v = line.split()
        hdr[k] = v
        if k == "DATA":
            break
    fields = hdr["FIELDS"]
    sizes = [int(s) for s in hdr["SIZE"]]
    types = hdr["TYPE"]
    counts = [int(c) for c in hdr.get("COUNT", ["1"] * len(fields))]
    n = int(hdr["POINTS"][0]) if "POINTS" in hdr else int(hdr["WIDTH"][0]) * int(hdr.get("HEIGHT", ["1"])[0])
    mode = hdr["DATA"][0]
    fi = [fields.index(a) for a in "xyz"]
    if mode == "ascii":
        col = [sum(counts[:f]) for f in fi]
        pts = []
        for line in b[pos:].decode("latin-1").split("\n"):
            tok = line.split()
            if not tok:
                continue
            pts.append([np.float32(np.nan) if tok[c].lower() == "nan" else
                        (strtof(tok[c]) if types[f] == "F" else np.float32(float(tok[c]))) for c, f in zip(col, fi)])
            if len(pts) == n:
                break
        return np.array(pts, np.float32).reshape(-1, 3)
    dt = np.dtype([(f"f{i}", _PCD_NP[(types[i], sizes[i])], (counts[i],)) for i in range(len(fields))])
    if mode == "binary":
        rec = np.frombuffer(b[pos:pos + n * dt.itemsize], dt, count=n)
        return np.stack([rec[f"f{f}"][:, 0].astype(np.float32) for f in fi], axis=1)
    if mode == "binary_compressed":
        csize, usize = struct.unpack("<II", b[pos:pos + 8])
        raw = lzf_decompress(b[pos + 8:pos + 8 + csize], usize)
        cols, off = [], 0
        for i in range(len(fields)):
            a = np.frombuffer(raw[off:off + n * sizes[i] * counts[i]], _PCD_NP[(types[i], sizes[i])]).reshape(n, counts[i])
            cols.append(a[:, 0])
            off += n * sizes[i] * counts[i]
        return np.stack([cols[f].astype(np.float32) for f in fi], axis=1)
    raise ValueError("unknown PCD DATA " + mode)


def load_point_cloud(path, use_model_path=False, shift_to_zero=False, offset=(0.0, 0.0, 0.0), scaling=1.0):
    if use_model_path and os.environ.get("GPU_VOXELS_MODEL_PATH"):
        base = os.environ["GPU_VOXELS_MODEL_PATH"]
        path = base + path if (base.endswith("/") or path.startswith("/")) else base + "/" + path
    if "xyz" in path:
        pts = read_xyz(path)
    elif "pcd" in path:
        pts = read_pcd(path)
    elif "binvox" in path:
        pts = read_binvox(path)
    else:
        raise ValueError("no known file format")
    if shift_to_zero and len(pts):
        pts = pts - pts.min(axis=0)
    return (np.float32(scaling) * pts + np.asarray(offset, np.float32)).astype(np.float32)


# ------------------------------------------------------------------ dumps
def map_dump(path):
    """(map type, side, dims, raw voxel bytes) of a map dump."""
    b = open(path, "rb").read()
    t, side, dx, dy, dz = struct.unpack("<ifIII", b[:20])
    return t, np.float32(side), (dx, dy, dz), b[20:]


def make_map_dump(path, mtype, side, dims, raw):
    with open(path, "wb") as f:
        f.write(struct.pack("<ifIII", mtype, side, *dims))
        f.write(bytes(raw))


def list_dump(path):
    """(map type, dims, side, ids, coords, voxel bytes) of a voxel-list dump."""
    b = open(path, "rb").read()
    t, dx, dy, dz, side, n = struct.unpack("<iIIIfI", b[:24])
    o = 24
    ids = np.frombuffer(b[o:o + 4 * n], np.uint32)
    o += 4 * n
    coords = np.frombuffer(b[o:o + 12 * n], np.uint32).reshape(n, 3)
    o += 12 * n
    bits = np.frombuffer(b[o:o + 32 * n], np.uint8).reshape(n, 32)
    return t, (dx, dy, dz), np.float32(side), ids, coords, bits


def make_list_dump(path, dims, side, ids, coords, bits, mtype=MT_BITVECTOR_VOXELLIST):
    with open(path, "wb") as f:
        f.write(struct.pack("<iIIIfI", mtype, *dims, side, len(ids)))
        f.write(np.ascontiguousarray(ids, np.uint32).tobytes())
        f.write(np.ascontiguousarray(coords, np.uint32).tobytes())
        f.write(np.ascontiguousarray(bits, np.uint8).tobytes())


# ------------------------------------------------------------------ writers (test inputs)
def write_binvox(path, occ, translate=(0.0, 0.0, 0.0), scale=1.0):
    """occ[x, z, y] (d x h x w) of 0/1 as binvox: dim d h w, runs of at most 255."""
    d, h, w = occ.shape
    flat = np.ascontiguousarray(occ, np.uint8).ravel()
    out = bytearray()
    i = 0
    while i < len(flat):
        v = flat[i]
        j = i
        while j < len(flat) and flat[j] == v and j - i < 255:
            j += 1
        out += bytes([v, j - i])
        i = j
    with open(path, "wb") as f:
        f.write(f"#binvox 1\ndim {d} {h} {w}\ntranslate {translate[0]} {translate[1]} {translate[2]}\n"
                f"scale {scale}\ndata\n".encode())
        f.write(bytes(out))


def lzf_compress(src):
    """Greedy LZF: 3-byte hash matches up to 8192 back, literal runs of at most 32."""
    out, lit = bytearray(), bytearray()
    table = {}
    i, n = 0, len(src)

    def flush():
        while lit:
            chunk = lit[:32]
            out.append(len(chunk) - 1)
            out.extend(chunk)
            del lit[:32]

    while i < n:
        key = src[i:i + 3]
        ref = table.get(key) if len(key) == 3 else None
        table[key] = i
        if ref is not None and i - ref - 1 < 8192:
            ln = 0
            while i + ln < n and ln < 264 and src[ref + ln] == src[i + ln]:
                ln += 1
            if ln >= 3:
                flush()
                off = i - ref - 1
                l2 = ln - 2
                if l2 < 7:
                    out.append((l2 << 5) | (off >> 8))
                else:
                    out.append((7 << 5) | (off >> 8))
                    out.append(l2 - 7)
                out.append(off & 0xFF)
                i += ln
                continue
        lit.append(src[i])
        i += 1
    flush()
    return bytes(out)


def write_pcd(path, xyz, mode="binary", extra=None):
    """x y z (float32) plus optional extra uint32 field 'rgba', in ascii / binary /
    binary_compressed encoding."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    n = len(xyz)
    fields = ["x", "y", "z"] + (["rgba"] if extra is not None else [])
    sizes = ["4"] * len(fields)
    types = ["F", "F", "F"] + (["U"] if extra is not None else [])
    head = (f"# .PCD v0.7 - Point Cloud Data file format\nVERSION 0.7\nFIELDS {' '.join(fields)}\n"
            f"SIZE {' '.join(sizes)}\nTYPE {' '.join(types)}\nCOUNT {' '.join(['1'] * len(fields))}\n"
            f"WIDTH {n}\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS {n}\nDATA {mode}\n").encode()
    with open(path, "wb") as f:
        f.write(head)
        if mode == "ascii":
            for i in range(n):
                vals = [repr(float(v)) for v in xyz[i]] + ([str(int(extra[i]))] if extra is not None else [])
                f.write((" ".join(vals) + "\n").encode())
        elif mode == "binary":
            rec = np.zeros(n, np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")] +
                                       ([("rgba", "<u4")] if extra is not None else [])))
            rec["x"], rec["y"], rec["z"] = xyz[:, 0], xyz[:, 1], xyz[:, 2]
            if extra is not None:
                rec["rgba"] = extra
            f.write(rec.tobytes())
        elif mode == "binary_compressed":
            raw = xyz[:, 0].tobytes() + xyz[:, 1].tobytes() + xyz[:, 2].tobytes()
            if extra is not None:
                raw += np.ascontiguousarray(extra, np.uint32).tobytes()
            comp = lzf_compress(raw)
            f.write(struct.pack("<II", len(comp), len(raw)))
            f.write(comp)
        else:
            raise ValueError(mode)
