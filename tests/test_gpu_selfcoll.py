"""insertMetaPointCloudWithSelfCollisionCheck and insertRobotIntoMapSelfCollAware on the GPU:
the reference's known answers and parity with the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from refgen import DIM_X, DIM_Y, DIM_Z, bits_of, create_box_of_points, set_bits

pytestmark = pytest.mark.gpu
DIMS = (DIM_X, DIM_Y, DIM_Z)


@pytest.mark.parametrize("j", [2, 10])
def test_selfcollision_known_answers_gpu(gvx_lib, j):
    # test/testing_voxelmap.cu:366-479 (selfcollision_metapointcloud)
    gvx = gvx_lib
    d = 1.0 / j
    boxes_c = [((2.1,) * 3, (5.1,) * 3), ((4.1,) * 3, (7.1,) * 3), ((6.1,) * 3, (9.1,) * 3), ((10.1,) * 3, (12.1,) * 3)]
    boxes_n = [((30.1,) * 3, (32.1,) * 3), ((40.1,) * 3, (43.1,) * 3), ((60.1,) * 3, (64.1,) * 3),
               ((65.1,) * 3, (68.1,) * 3)]
    cc = [create_box_of_points(a, b, d) for a, b in boxes_c]
    cn = [create_box_of_points(a, b, d) for a, b in boxes_n]
    pc, sc = np.concatenate(cc), [len(c) for c in cc]
    pn, sn = np.concatenate(cn), [len(c) for c in cn]
    m = gvx.Map(gvx.BITVECTOR_VOXELMAP, DIMS, 1.0)
    assert m.insert_meta_selfcoll(pc, sc)[0] is True
    assert m.insert_meta_selfcoll(pn, sn)[0] is False
    m.clear()
    meanings = np.array([51, 52, 53, 54], np.uint32)
    full = set_bits(*range(256)) & ~set_bits(2)
    masks = np.stack([full & ~set_bits(52, 51), full & ~set_bits(51, 52), full & ~set_bits(53), full & ~set_bits(54)])
    coll, col, _ = m.insert_meta_selfcoll(pc, sc, meanings, masks)
    assert coll is True and bits_of(col) == [52, 53]
    masks = np.stack([full & ~set_bits(k) for k in (51, 52, 53, 54)])
    coll, col, _ = m.insert_meta_selfcoll(pn, sn, meanings, masks)
    assert coll is False and not col.any()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_selfcollision_parity(gvx_lib, seed):
    gvx = gvx_lib
    rng = np.random.default_rng(70 + seed)
    dims, side = (57, 41, 29), 0.05
    n_vox = int(np.prod(dims))
    start = O.new_map(O.BITVECTOR_MAP, dims)
    sel = rng.random(n_vox) < 0.02                      # prior content, some FREE-only voxels
    start[sel, rng.integers(0, 32, sel.sum())] = rng.integers(1, 256, sel.sum()).astype(np.uint8)
    start[rng.random(n_vox) < 0.02, 0] |= 1
    for variant in range(3):
        ncl = int(rng.integers(1, 9))
        sizes = rng.integers(0, 3000, ncl).astype(np.uint32)
        pts = rng.uniform(-0.1, 3.0, size=(int(sizes.sum()), 3)).astype(np.float32)
        pts[rng.random(len(pts)) < 0.01] = np.nan
        meanings = masks = None
        if variant >= 1:
            meanings = rng.integers(0, 256, ncl).astype(np.uint32)
        if variant == 2:                                 # arbitrary masks, own meaning and COLLISION kept
            masks = rng.integers(0, 256, (ncl, 32)).astype(np.uint8)
        g = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        g.write(start)
        o = start.copy()
        got = g.insert_meta_selfcoll(pts, sizes, meanings, masks)
        want = O.map_insert_meta_selfcoll(o, dims, side, pts, sizes, meanings, masks)
        assert got[0] == want[0] and np.array_equal(got[1], want[1]) and got[2] == want[2], variant
        assert np.array_equal(g.read(), o), variant
    with pytest.raises(gvx.GvxError):
        gvx.Map(gvx.PROBAB_VOXELMAP, dims, side).insert_meta_selfcoll(pts, sizes)


def test_robot_selfcoll_and_insert_parity(gvx_lib):
    gvx = gvx_lib
    from gvx import scenes
    c = scenes.cfg2()
    r = c["arm"]
    dims, side = (256, 256, 256), 0.02
    rng = np.random.default_rng(5)
    off = np.concatenate([[0], np.cumsum(r["link_sizes"])]).astype(np.int64)
    for k in range(3):
        q = rng.uniform(-np.pi, np.pi, 7).astype(np.float32)
        base = scenes.translation(2.56, 2.56, 0.5)
        T = O.fk_chain(r["dh"], base, q)
        posed = np.concatenate([O.transform_points(T[l], r["link_points"][off[l]:off[l + 1]]) for l in range(7)])
        g = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        got = gvx.map_insert_robot_selfcoll(g, r["link_points"], r["link_sizes"], r["dh"], q, base)
        o = O.new_map(O.BITVECTOR_MAP, dims)
        want = O.map_insert_meta_selfcoll(o, dims, side, posed, r["link_sizes"])
        assert got[0] == want[0] and np.array_equal(got[1], want[1]) and got[2] == want[2]
        assert np.array_equal(g.read(), o)
        # plain insertRobotIntoMap with per-link meanings against the oracle's meta insert
        g2 = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        mn = np.arange(4, 11, dtype=np.uint32)
        out = gvx.map_insert_robot(g2, r["link_points"], r["link_sizes"], r["dh"], q, base, mn)
        o2 = O.new_map(O.BITVECTOR_MAP, dims)
        assert out == O.map_insert_meta(O.BITVECTOR_MAP, o2, dims, side, posed, r["link_sizes"], mn)
        assert np.array_equal(g2.read(), o2)


def test_robot_transform_prob_and_list_insert(gvx_lib):
    """The other users of the FK kernel: robot_transform (every posed point equal to the
    oracle's, as floats), insertRobotIntoMap into a ProbVoxelMap (one meaning) and
    into a VoxelList with per-link meanings, and the fused BitVector path with part of
    the arm outside the map."""
    gvx = gvx_lib
    from gvx import scenes
    r = scenes.cfg2()["arm"]
    dims, side = (128, 128, 128), 0.02
    rng = np.random.default_rng(12)
    off = np.concatenate([[0], np.cumsum(r["link_sizes"])]).astype(np.int64)
    mn = np.arange(4, 11, dtype=np.uint32)
    for k in range(2):
        q = rng.uniform(-np.pi, np.pi, 7).astype(np.float32)
        base = scenes.translation(1.28, 1.28, 0.3 if k == 0 else 0.05)
        T = O.fk_chain(r["dh"], base, q)
        posed = np.concatenate([O.transform_points(T[l], r["link_points"][off[l]:off[l + 1]]) for l in range(7)])
        got = gvx.robot_transform(r["link_points"], r["link_sizes"], r["dh"], q, base)
        assert np.array_equal(np.asarray(got).reshape(-1, 3), posed)
        gp = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
        outp = gvx.map_insert_robot(gp, r["link_points"], r["link_sizes"], r["dh"], q, base, None, gvx.eBVM_OCCUPIED)
        op = O.new_map(O.PROB_MAP, dims)
        assert outp == O.map_insert_points(O.PROB_MAP, op, dims, side, posed, gvx.eBVM_OCCUPIED)
        assert np.array_equal(gp.read(), op)
        gb = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
        outb = gvx.map_insert_robot(gb, r["link_points"], r["link_sizes"], r["dh"], q, base, mn)
        ob = O.new_map(O.BITVECTOR_MAP, dims)
        assert outb == O.map_insert_meta(O.BITVECTOR_MAP, ob, dims, side, posed, r["link_sizes"], mn)
        assert np.array_equal(gb.read(), ob)
        gl = gvx.VoxelList(dims, side)
        gvx.list_insert_robot(gl, r["link_points"], r["link_sizes"], r["dh"], q, base, mn)
        ol = O.List(dims, side)
        ol.insert_points(posed, sizes=r["link_sizes"], meanings=mn)
        ids, coords, bits = gl.read()
        assert np.array_equal(ids, ol.ids) and np.array_equal(coords, ol.coords) and np.array_equal(bits, ol.bits)
