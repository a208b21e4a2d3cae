"""Map and voxel-list dumps (writeToDisk / readFromDisk) and insertPointCloudFromFile on
the GPU, against the dump layouts and loaders of the oracle."""
import os
import shutil

import numpy as np
import pytest

from oracle import files as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu
MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


@pytest.fixture
def work():
    d = f"/tmp/gvx_dumps_{os.getpid()}"
    os.makedirs(d, exist_ok=True)
    yield d
    shutil.rmtree(d, ignore_errors=True)


def _random_map_bytes(rng, mtype, n):
    if mtype == O.DIST_MAP:
        return rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    per = 32 if mtype == O.BITVECTOR_MAP else 1
    return rng.integers(0, 256, n * per, dtype=np.uint16).astype(np.uint8).view(
        np.int8 if mtype == O.PROB_MAP else np.uint8)


@pytest.mark.parametrize("mtype", [O.BITVECTOR_MAP, O.PROB_MAP, O.DIST_MAP])
def test_map_dump_round_trip(gvx_lib, work, mtype):
    gvx = gvx_lib
    rng = np.random.default_rng(mtype)
    dims, side = (40, 30, 21), 0.02
    n = dims[0] * dims[1] * dims[2]
    m = gvx.Map(mtype, dims, side)
    m.write(_random_map_bytes(rng, mtype, n))
    p = f"{work}/m.dump"
    m.write_to_disk(p)
    t, fside, fdims, raw = F.map_dump(p)
    assert (t, fside, fdims) == (mtype, np.float32(side), dims)
    assert raw == m.read().tobytes()
    m2 = gvx.Map(mtype, dims, side)
    assert m2.read_from_disk(p) is True
    assert np.array_equal(m2.read(), m.read())
    # a dump made by the oracle with another side length: loaded, side reported as different
    other = _random_map_bytes(rng, mtype, n)
    F.make_map_dump(f"{work}/o.dump", mtype, 0.05, dims, other.tobytes())
    assert m2.read_from_disk(f"{work}/o.dump") is False
    assert m2.read().tobytes() == other.tobytes()
    # wrong type or dimensions are refused and leave the map alone
    F.make_map_dump(f"{work}/t.dump", 9 if mtype != 9 else 0, side, dims, b"")
    with pytest.raises(gvx.GvxError):
        m2.read_from_disk(f"{work}/t.dump")
    F.make_map_dump(f"{work}/d.dump", mtype, side, (dims[0], dims[1], dims[2] + 1), b"")
    with pytest.raises(gvx.GvxError):
        m2.read_from_disk(f"{work}/d.dump")
    F.make_map_dump(f"{work}/s.dump", mtype, side, dims, other.tobytes()[:-1])   # truncated
    with pytest.raises(gvx.GvxError):
        m2.read_from_disk(f"{work}/s.dump")


def test_list_dump_round_trip(gvx_lib, work):
    gvx = gvx_lib
    rng = np.random.default_rng(11)
    dims, side = (57, 33, 21), 0.03
    lst = gvx.VoxelList(dims, side)
    pts = rng.uniform(0, 0.6, size=(20000, 3)).astype(np.float32)
    lst.insert_points(pts, sizes=np.array([12000, 8000], np.uint32), meanings=np.array([5, 77], np.uint32))
    p = f"{work}/l.dump"
    lst.write_to_disk(p)
    t, fdims, fside, ids, coords, bits = F.list_dump(p)
    gids, gcoords, gbits = lst.read()
    assert (t, fdims, fside) == (F.MT_BITVECTOR_VOXELLIST, dims, np.float32(side))
    assert np.array_equal(ids, gids) and np.array_equal(coords, gcoords) and np.array_equal(bits, gbits)
    l2 = gvx.VoxelList(dims, side)
    l2.read_from_disk(p)
    for a, b in zip(l2.read(), lst.read()):
        assert np.array_equal(a, b)
    # a dump made by the oracle, then the refusals (type, dims, side)
    o = O.List(dims, side)
    o.insert_points(pts[:500], meaning=9)
    F.make_list_dump(f"{work}/o.dump", dims, side, o.ids, o.coords, o.bits)
    l2.read_from_disk(f"{work}/o.dump")
    got = l2.read()
    assert np.array_equal(got[0], o.ids) and np.array_equal(got[1], o.coords) and np.array_equal(got[2], o.bits)
    F.make_list_dump(f"{work}/e.dump", dims, side, o.ids[:0], o.coords[:0], o.bits[:0])
    l2.read_from_disk(f"{work}/e.dump")
    assert len(l2) == 0
    for kw in ({"mtype": 0}, {"dims": (57, 33, 22)}, {"side": 0.031}):
        a = dict(dims=dims, side=side, mtype=F.MT_BITVECTOR_VOXELLIST)
        a.update(kw)
        F.make_list_dump(f"{work}/x.dump", a["dims"], a["side"], o.ids, o.coords, o.bits, mtype=a["mtype"])
        with pytest.raises(gvx.GvxError):
            l2.read_from_disk(f"{work}/x.dump")


@pytest.mark.parametrize("name", ["arm_6_link.binvox", "hollie_from_pointcloud2.pcd"])
def test_insert_point_cloud_from_file(gvx_lib, name):
    # GpuVoxelsMap::insertPointCloudFromFile = loadPointCloud + insertPointCloud
    gvx = gvx_lib
    pts = F.load_point_cloud(os.path.join(MODELS, name), shift_to_zero=True, offset=(0.1, 0.2, 0.05))
    loaded = gvx.load_point_cloud(os.path.join(MODELS, name), shift_to_zero=True, offset=(0.1, 0.2, 0.05))
    assert np.array_equal(loaded, pts, equal_nan=True)
    dims, side = (128, 128, 128), 0.01
    g = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
    g.insert_points(loaded, 1)                       # eBVM_OCCUPIED
    om = O.new_map(O.BITVECTOR_MAP, dims)
    O.map_insert_points(O.BITVECTOR_MAP, om, dims, side, pts, 1)
    assert np.array_equal(g.read().ravel(), om.ravel())
