"""GPU parity of BitVectorVoxelList: insert, radix sort + dedup, merge, collisions."""
import numpy as np
import pytest

from oracle import oracle as O
from refgen import DIM_X, DIM_Y, DIM_Z, bits_of, create_box_of_points, set_bits

pytestmark = pytest.mark.gpu
DIMS = (DIM_X, DIM_Y, DIM_Z)


def _same(gl, ol):
    ids, coords, bits = gl.read()
    assert np.array_equal(ids, ol.ids)
    assert np.array_equal(coords, ol.coords)
    assert np.array_equal(bits, ol.bits)


def _pts(rng, n, dims, side):
    hi = np.array(dims) * side
    p = rng.uniform(-0.1 * hi, 1.1 * hi, size=(n, 3)).astype(np.float32)
    p[rng.random(n) < 0.005] = np.nan
    return p


@pytest.mark.parametrize("n", [0, 1, 63, 4097, 300000])
def test_insert_parity(gvx_lib, n):
    gvx = gvx_lib
    rng = np.random.default_rng(n)
    dims, side = (57, 33, 21), 0.03
    g = gvx.VoxelList(dims, side)
    o = O.List(dims, side)
    p = _pts(rng, n, dims, side)
    g.insert_points(p, meaning=17)
    o.insert_points(p, meaning=17)
    _same(g, o)
    # second insert with per-cloud meanings (duplicates across calls are OR-merged)
    sizes = np.array([n // 3, n - n // 3], np.uint32)
    mn = np.array([4, 200], np.uint32)
    g.insert_points(p[::-1].copy(), sizes=sizes, meanings=mn)
    o.insert_points(p[::-1].copy(), sizes=sizes, meanings=mn)
    _same(g, o)


def test_large_ids_multi_pass_sort(gvx_lib):
    # 512^3 reference map: ids need four 8-bit passes
    gvx = gvx_lib
    rng = np.random.default_rng(1)
    dims, side = (512, 512, 512), 0.01
    p = rng.uniform(0, 5.12, size=(200000, 3)).astype(np.float32)
    p[:50000] = p[:50000] // np.float32(0.05) * np.float32(0.05)    # many duplicates
    g = gvx.VoxelList(dims, side)
    o = O.List(dims, side)
    g.insert_points(p, meaning=9)
    o.insert_points(p, meaning=9)
    _same(g, o)


def test_insert_coords_parity(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(2)
    dims = (20, 30, 40)
    c = rng.integers(0, 45, size=(10000, 3)).astype(np.uint32)
    g = gvx.VoxelList(dims, 1.0)
    o = O.List(dims, 1.0)
    g.insert_coords(c, 33)
    o.insert_coords(c, 33)
    _same(g, o)


@pytest.mark.parametrize("offset,meaning", [(None, -1), ((2, -1, 3), -1), ((0, 0, 0), 77), ((-5, 4, -2), 3)])
def test_merge_list_parity(gvx_lib, offset, meaning):
    gvx = gvx_lib
    rng = np.random.default_rng(3)
    dims, side = (30, 30, 30), 0.1
    a, b = gvx.VoxelList(dims, side), gvx.VoxelList(dims, side)
    oa, ob = O.List(dims, side), O.List(dims, side)
    pa, pb = _pts(rng, 5000, dims, side), _pts(rng, 7000, dims, side)
    a.insert_points(pa, meaning=10); oa.insert_points(pa, meaning=10)
    b.insert_points(pb, meaning=20); ob.insert_points(pb, meaning=20)
    a.merge_list(b, offset=offset, meaning=meaning)
    oa.merge_list(ob, offset=(0, 0, 0) if offset is None else offset, meaning=None if meaning < 0 else meaning)
    _same(a, oa)


def test_merge_map_parity(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(4)
    dims = (25, 19, 13)
    om = O.new_map(O.BITVECTOR_MAP, dims)
    c = rng.integers(0, (25, 19, 13), size=(800, 3)).astype(np.uint32)
    O.map_insert_coords(O.BITVECTOR_MAP, om, dims, c, 40)
    O.map_insert_coords(O.BITVECTOR_MAP, om, dims, c[:200], 0)
    m = gvx.Map(O.BITVECTOR_MAP, dims, 1.0)
    m.write(om)
    for off, mn in ((None, -1), ((1, 2, -1), 5)):
        g = gvx.VoxelList(dims, 1.0)
        o = O.List(dims, 1.0)
        pre = rng.integers(0, 13, size=(300, 3)).astype(np.uint32)
        g.insert_coords(pre, 99); o.insert_coords(pre, 99)
        g.merge_map(m, offset=off, meaning=mn)
        o.merge_map(om, offset=(0, 0, 0) if off is None else off, meaning=None if mn < 0 else mn)
        _same(g, o)


@pytest.mark.parametrize("mt", [O.BITVECTOR_MAP, O.PROB_MAP])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_list_collide_map_parity(gvx_lib, mt, mode):
    gvx = gvx_lib
    rng = np.random.default_rng(5 + mode)
    dims, side = (40, 35, 30), 0.05
    om = O.new_map(mt, dims)
    env = _pts(rng, 30000, dims, side)
    O.map_insert_points(mt, om, dims, side, env, 1 if mt == O.PROB_MAP else 1)
    if mt == O.PROB_MAP:
        O.map_insert_points(mt, om, dims, side, env[:10000], 0)
    m = gvx.Map(mt, dims, side)
    m.write(om)
    g, o = gvx.VoxelList(dims, side), O.List(dims, side)
    p = _pts(rng, 20000, dims, side)
    sizes = np.array([5000, 15000], np.uint32)
    g.insert_points(p, sizes=sizes, meanings=[30, 31]); o.insert_points(p, sizes=sizes, meanings=[30, 31])
    tm = set_bits(31)
    for off in (None, (1, 0, -2), (0, 5, 0)):
        for thr in (0.3, 1.0, 60.0):
            n, mask = g.collide_map(m, thr, offset=off, mode=mode, typemask=tm)
            n2, mask2 = o.collide_map(mt, om, thr, offset=(0, 0, 0) if off is None else off, mode=mode, typemask=tm)
            assert n == n2
            assert np.array_equal(mask, mask2)
            _same(g, o)


def test_list_collide_list_parity(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(6)
    dims, side = (30, 30, 30), 0.1
    a, b = gvx.VoxelList(dims, side), gvx.VoxelList(dims, side)
    oa, ob = O.List(dims, side), O.List(dims, side)
    pa, pb = _pts(rng, 9000, dims, side), _pts(rng, 4000, dims, side)
    a.insert_points(pa, meaning=12); oa.insert_points(pa, meaning=12)
    b.insert_points(pb, meaning=13); ob.insert_points(pb, meaning=13)
    for x, y, ox, oy in ((a, b, oa, ob), (b, a, ob, oa)):
        n, mask = x.collide_list(y)
        n2, mask2 = ox.collide_list(oy)
        assert n == n2 and np.array_equal(mask, mask2)


# ------------------------------------------------------------ reference known answers on the GPU
def _two_boxes():
    c1 = create_box_of_points((1.1, 1.1, 1.1), (3.9, 3.9, 3.9), 0.1)
    c2 = create_box_of_points((2.1, 2.1, 2.1), (4.9, 4.9, 4.9), 0.1)
    return np.concatenate([c1, c2]), np.array([len(c1), len(c2)], np.uint32)


def test_known_answers_lists(gvx_lib):
    gvx = gvx_lib
    pts, sizes = _two_boxes()
    L = gvx.VoxelList(DIMS, 1.0)
    L.insert_points(pts, sizes=sizes, meanings=[11, 12])                 # testing_voxellist.cu:284-341
    ids, coords, bits = L.read()
    assert len(ids) == 46
    assert np.all(bits == set_bits(11, 12), axis=1).sum() == 8
    prob = gvx.Map(O.PROB_MAP, DIMS, 1.0)
    prob.insert_meta(pts, sizes, None, 1)
    bitm = gvx.Map(O.BITVECTOR_MAP, DIMS, 1.0)
    bitm.insert_meta(pts, sizes, None, 1)
    thr = float(np.float32(199.0 / 254.0))
    assert L.collide_map(prob, thr, mode=0)[0] == 46                     # testing_voxellist.cu:111-148
    for m in (prob, bitm):                                               # testing_voxellist.cu:151-238
        L2 = gvx.VoxelList(DIMS, 1.0)
        L2.insert_points(pts, sizes=sizes, meanings=[11, 12])
        n, mask = L2.collide_map(m, 1.0, mode=1)
        assert n == 46 and bits_of(mask) == [11, 12]
    box = create_box_of_points((1.1, 1.1, 1.1), (3.9, 3.9, 3.9), 0.1)   # testing_voxellist.cu:241-281
    L3 = gvx.VoxelList(DIMS, 1.0)
    L3.insert_points(box, meaning=1)
    p3 = gvx.Map(O.PROB_MAP, DIMS, 1.0)
    p3.insert_points(box, 1)
    for shift, expect in ((0, 27), (1, 18), (2, 9), (3, 0)):
        assert L3.collide_map(p3, thr, offset=(shift, 0, 0), mode=0)[0] == expect


def test_known_answer_bitmasked(gvx_lib):
    # testing_voxellist.cu:859-985
    gvx = gvx_lib
    side = 0.025
    L = gvx.VoxelList(DIMS, side)
    prob = gvx.Map(O.PROB_MAP, DIMS, side)
    prob.insert_points(create_box_of_points((0.111, 0.111, 0.111), (0.799, 0.399, 0.399), 0.005), 1)
    for lo, m in (((0.111, 0.311, 0.111), 34), ((0.311, 0.311, 0.111), 63), ((0.511, 0.311, 0.111), 102)):
        L.insert_points(create_box_of_points(lo, (lo[0] + 0.288, 0.599, 0.399), 0.005), meaning=m)
    expect = {(0, 0, 0): [0, 0, 9, 15, 21], (0, -4, 0): [0, 0, 18, 30, 42], (0, -4, -4): [0, 0, 12, 20, 28]}
    for off, counts in expect.items():
        tm = np.zeros(32, np.uint8)
        got = [L.collide_map(prob, 1.0, offset=off, mode=2, typemask=tm)[0]]
        for b in (33, 34, 63, 102):
            tm |= set_bits(b)
            got.append(L.collide_map(prob, 1.0, offset=off, mode=2, typemask=tm)[0])
        assert got == [c * 64 for c in counts], off


def test_unique_fallback_paths(gvx_lib):
    """Duplicates are merged through a dense bitmap when the map is small next to the
    list; the radix sort still runs for a sparse 1024^3 list (key width from the map
    size) and after a merge with an offset (ids may leave the map: key width from the
    largest key).  Every path against the oracle."""
    gvx = gvx_lib
    rng = np.random.default_rng(7)
    dims, side = (1024, 1024, 1024), 0.01
    p = rng.uniform(-0.5, 10.74, size=(5000, 3)).astype(np.float32)
    p[:1000] = p[:1000] // np.float32(0.3) * np.float32(0.3)
    g, o = gvx.VoxelList(dims, side), O.List(dims, side)
    g.insert_points(p, meaning=5)
    o.insert_points(p, meaning=5)
    _same(g, o)
    dims, side = (64, 48, 40), 0.05
    a, b = gvx.VoxelList(dims, side), gvx.VoxelList(dims, side)
    oa, ob = O.List(dims, side), O.List(dims, side)
    pa, pb = _pts(rng, 6000, dims, side), _pts(rng, 6000, dims, side)
    a.insert_points(pa, meaning=1); oa.insert_points(pa, meaning=1)
    b.insert_points(pb, meaning=2); ob.insert_points(pb, meaning=2)
    a.merge_list(b, offset=(-7, 5, 3), meaning=-1)
    oa.merge_list(ob, offset=(-7, 5, 3), meaning=None)
    _same(a, oa)
    pc = _pts(rng, 4000, dims, side)
    a.insert_points(pc, meaning=3); oa.insert_points(pc, meaning=3)     # sorted by the radix path now
    _same(a, oa)
    a.clear(); oa = O.List(dims, side)
    a.insert_points(pc, meaning=4); oa.insert_points(pc, meaning=4)     # cleared: bitmap path again
    _same(a, oa)
