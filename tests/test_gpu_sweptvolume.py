"""Swept-volume operations on the GPU against the oracle and the reference's known answers:
windowed SV collision, shiftLeftSweptVolumeIDs, collideWithBitcheck, collideCountingPerMeaning."""
import numpy as np
import pytest

from oracle import oracle as O
from refgen import DIM_X, DIM_Y, DIM_Z, create_box_of_points

pytestmark = pytest.mark.gpu
DIMS = (DIM_X, DIM_Y, DIM_Z)


def _sv_map(rng, n, density):
    m = np.zeros((n, 32), np.uint8)
    occ = rng.random(n) < density
    k = int(occ.sum())
    bits = rng.random((k, 256)) < 0.03
    m[occ] = np.packbits(bits, axis=1, bitorder="little")
    return m


@pytest.mark.parametrize("window,sv_offset", [(0, 0), (1, 0), (2, 0), (3, 0), (1, 3), (2, 9), (0, 17), (4, 5)])
def test_map_collide_sv_parity(gvx_lib, window, sv_offset):
    gvx = gvx_lib
    rng = np.random.default_rng(window * 31 + sv_offset)
    dims = (37, 29, 23)
    n = dims[0] * dims[1] * dims[2]
    a, b = _sv_map(rng, n, 0.3), _sv_map(rng, n, 0.3)
    ga, gb = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, 0.1), gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, 0.1)
    ga.write(a)
    gb.write(b)
    c, mask = ga.collide_sv(gb, 1.0, window, sv_offset)
    oc, omask = O.map_collide_sv(O.BITVECTOR_MAP, a, O.BITVECTOR_MAP, b, dims, 1.0, window, sv_offset)
    assert c == oc and c > 0
    assert np.array_equal(mask, omask)
    assert np.array_equal(ga.read(), a)          # eBVM_COLLISION marks on self, as the oracle
    assert np.array_equal(gb.read(), b)


@pytest.mark.parametrize("shift", [0, 1, 3, 7, 8, 13, 31, 50, 56])
def test_shift_sv_parity(gvx_lib, shift):
    gvx = gvx_lib
    rng = np.random.default_rng(shift)
    dims = (20, 20, 20)
    a = rng.integers(0, 256, (8000, 32), dtype=np.uint8)
    g = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, 0.1)
    g.write(a)
    g.shift_sv(shift)
    assert O.shift_sv(a, shift)
    assert np.array_equal(g.read(), a)
    with pytest.raises(gvx.GvxError):
        g.shift_sv(57)


def test_map_bitshift_known_answer(gvx_lib):
    # testing_voxelmap.cu:244-299: collisionCheckBitvector with SVCollider(1.0, window 1)
    gvx = gvx_lib
    pts = np.array([[0.3, 0.3, 0.3], [1.3, 1.3, 1.3], [2.3, 2.3, 2.3], [3.3, 3.3, 3.3]], np.float32)
    m1, m2 = gvx.Map(gvx.BITVECTOR_VOXELMAP, DIMS, 1.0), gvx.Map(gvx.BITVECTOR_VOXELMAP, DIMS, 1.0)
    for shift in range(0, 49):
        m1.clear()
        m2.clear()
        m2.insert_points(pts, 54)
        m1.insert_points(pts, 54 - shift)
        m2.shift_sv(shift)
        assert m1.collide_sv(m2, 1.0, 1)[0] == 4


def _list(gvx, dims, side, pts, meaning=None, sizes=None, meanings=None):
    g = gvx.VoxelList(dims, side)
    o = O.List(dims, side)
    if sizes is None:
        g.insert_points(pts, meaning=meaning)
        o.insert_points(pts, meaning=meaning)
    else:
        g.insert_points(pts, sizes=sizes, meanings=meanings)
        o.insert_points(pts, sizes=sizes, meanings=meanings)
    return g, o


def test_list_bitcheck_and_per_meaning_parity(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(5)
    dims, side = (60, 50, 40), 0.05
    p1 = rng.uniform(0, 2.0, (30000, 3)).astype(np.float32)
    p2 = np.concatenate([p1[:12000], rng.uniform(0, 2.0, (20000, 3)).astype(np.float32)])
    s1 = np.array([10000, 10000, 10000], np.uint32)
    s2 = np.array([16000, 16000], np.uint32)
    g1, o1 = _list(gvx, dims, side, p1, sizes=s1, meanings=np.array([4, 30, 31], np.uint32))
    g2, o2 = _list(gvx, dims, side, p2, sizes=s2, meanings=np.array([5, 32], np.uint32))
    for margin in (0, 1, 2, 5):
        assert g1.collide_bitcheck(g2, margin) == o1.collide_bitcheck(o2, margin)
    t, c = g1.collide_per_meaning(g2)
    ot, oc = o1.collide_per_meaning(o2)
    assert t == ot and np.array_equal(c, oc) and t > 0
    for shift in (1, 26):
        g2.shift_sv(shift)
        o2.shift_sv(shift)
        assert np.array_equal(g2.read()[2], o2.bits)
        for margin in (0, 1, 3):
            assert g1.collide_bitcheck(g2, margin) == o1.collide_bitcheck(o2, margin)


def test_list_known_answers(gvx_lib):
    gvx = gvx_lib
    # testing_voxellist.cu:502-546 (bitvoxellist_bitshift_collision)
    pts = np.array([[0.3, 0.3, 0.3]], np.float32)
    for shift in range(0, 49):
        l1, l2 = gvx.VoxelList(DIMS, 1.0), gvx.VoxelList(DIMS, 1.0)
        l2.insert_points(pts, meaning=54)
        l1.insert_points(pts, meaning=54 - shift)
        l2.shift_sv(shift)
        assert l1.collide_bitcheck(l2, 1) == 1
    # testing_voxellist.cu:46-101 (collideCountingPerMeaning, swept volume vs obstacle)
    obstacle = gvx.VoxelList(DIMS, 1.0)
    obstacle.insert_points(create_box_of_points((8.0, 8.0, 8.0), (11.0, 11.0, 11.0), 0.1), meaning=1)
    clouds = []
    for i in range(50):
        p = np.float32(np.float32(np.float32(30) * np.float32(0.02)) * np.float32(i))
        clouds.append(create_box_of_points((p, p, p), (p + np.float32(2.5),) * 3, 0.1))
    swept = gvx.VoxelList(DIMS, 1.0)
    swept.insert_points(np.concatenate(clouds), sizes=np.array([len(c) for c in clouds], np.uint32),
                        meanings=np.arange(4, 54, dtype=np.uint32))
    total, counts = swept.collide_per_meaning(obstacle)
    assert {i: int(c) for i, c in enumerate(counts) if c} == {14: 1, 15: 8, 16: 8, 17: 27, 18: 27, 19: 27, 20: 8,
                                                              21: 1, 22: 1}
    assert total == 108
