"""GPU parity of dense maps: voxelization and map x map collision, HIP library vs oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from refgen import (DIM_X, DIM_Y, DIM_Z, NUM_POINTS, bits_of, create_box_of_points, create_checkerboard,
                    create_equidistant_points_in_box)

pytestmark = pytest.mark.gpu
DIMS = (DIM_X, DIM_Y, DIM_Z)
TYPES = (O.BITVECTOR_MAP, O.PROB_MAP, O.DIST_MAP)


def _pts(rng, n, dims, side, spread=1.1):
    hi = np.array(dims) * side * spread
    p = rng.uniform(-0.05 * hi, hi, size=(n, 3)).astype(np.float32)
    p[rng.random(n) < 0.01] = np.nan
    return p


def _map(gvx, mtype, dims, side):
    return gvx.Map(mtype, dims, side)


@pytest.mark.parametrize("mtype", TYPES)
def test_insert_points_parity(gvx_lib, mtype):
    gvx = gvx_lib
    rng = np.random.default_rng(10 + mtype)
    dims, side = (37, 29, 41), 0.05
    g = _map(gvx, mtype, dims, side)
    o = O.new_map(mtype, dims)
    # several calls with different meanings: serial semantics per call
    meanings = (1, 7, 200) if mtype == O.BITVECTOR_MAP else ((1, 0, 1, 4, 254, 129) if mtype == O.PROB_MAP else (1,))
    for m in meanings:
        p = _pts(rng, 20000, dims, side)
        out_g = g.insert_points(p, m)
        out_o = O.map_insert_points(mtype, o, dims, side, p, m)
        assert out_g == out_o
    assert np.array_equal(g.read(), o)


@pytest.mark.parametrize("mtype", [O.BITVECTOR_MAP, O.PROB_MAP])
def test_insert_runs_of_one_voxel(gvx_lib, mtype):
    # points in runs on one voxel (as neighbouring pixels of a depth image), runs
    # broken by other voxels, by a change of meaning and by points outside the map,
    # the same voxel again out of order, and a count that is not a multiple of 64.
    # insert_points and the single-meaning insert_meta go through the run-merging
    # kernel; insert_meta with per-cloud meanings through the per-cloud kernel.  Prob
    # runs stay below 26 points: 127 - 10 k saturates at -127 beyond that, and longer
    # runs would no longer show in the map.
    gvx = gvx_lib
    rng = np.random.default_rng(90 + mtype)
    dims, side = (16, 12, 10), 0.1
    centres = (rng.integers(0, 10, size=(300, 3)) + 0.5).astype(np.float32) * np.float32(side)
    lengths = rng.integers(1, 70 if mtype == O.BITVECTOR_MAP else 26, size=300)
    p = np.repeat(centres, lengths, axis=0)
    p[rng.random(len(p)) < 0.03] = np.float32(50.0)          # outside, inside a run
    p = p[: (len(p) // 64 - 1) * 64 + 37]
    sizes = np.array([len(p) // 3, 0, len(p) - len(p) // 3], np.uint32)
    mn = np.array([5, 9, 5], np.uint32) if mtype == O.BITVECTOR_MAP else np.array([1, 0, 1], np.uint32)
    g = _map(gvx, mtype, dims, side)
    o = O.new_map(mtype, dims)
    # Prob: the free pass after the occupied one leaves 127 - 10 x (points in the
    # voxel), so every run length shows in the map
    for m in ((7,) if mtype == O.BITVECTOR_MAP else (1, 0)):
        assert g.insert_points(p, m) == O.map_insert_points(mtype, o, dims, side, p, m)
        assert np.array_equal(g.read(), o)
    assert g.insert_meta(p, sizes, mn) == O.map_insert_meta(mtype, o, dims, side, p, sizes, mn)
    assert np.array_equal(g.read(), o)
    g.clear()
    o = O.new_map(mtype, dims)
    for m in ((7,) if mtype == O.BITVECTOR_MAP else (1, 0)):
        assert g.insert_meta(p, sizes, meaning=m) == O.map_insert_meta(mtype, o, dims, side, p, sizes, meaning=m)
        assert np.array_equal(g.read(), o)


def test_insert_points_heavy_contention_prob(gvx_lib):
    gvx = gvx_lib
    dims, side = (8, 8, 8), 0.1
    g = _map(gvx, O.PROB_MAP, dims, side)
    o = O.new_map(O.PROB_MAP, dims)
    p = np.tile(np.float32([[0.35, 0.45, 0.55]]), (50000, 1))
    for m in (0, 0, 1):
        g.insert_points(p, m)
        O.map_insert_points(O.PROB_MAP, o, dims, side, p, m)
        assert np.array_equal(g.read(), o)


@pytest.mark.parametrize("mtype", TYPES)
def test_insert_coords_and_meta_parity(gvx_lib, mtype):
    gvx = gvx_lib
    rng = np.random.default_rng(20 + mtype)
    dims, side = (33, 17, 25), 0.02
    g = _map(gvx, mtype, dims, side)
    o = O.new_map(mtype, dims)
    c = rng.integers(0, 40, size=(5000, 3)).astype(np.uint32)
    assert g.insert_coords(c, 1) == O.map_insert_coords(mtype, o, dims, c, 1)
    sizes = np.array([300, 0, 4000, 1234, 17], np.uint32)
    p = _pts(rng, int(sizes.sum()), dims, side)
    mn = np.array([1, 1, 1, 1, 1], np.uint32) if mtype == O.DIST_MAP else np.array([5, 9, 1, 130, 77], np.uint32)
    assert g.insert_meta(p, sizes, mn) == O.map_insert_meta(mtype, o, dims, side, p, sizes, mn)
    assert g.insert_meta(p, sizes, None, 1) == O.map_insert_meta(mtype, o, dims, side, p, sizes, None, 1)
    assert np.array_equal(g.read(), o)


def _random_map(gvx, rng, mtype, dims, dens):
    o = O.new_map(mtype, dims)
    n = int(np.prod(dims))
    if mtype == O.BITVECTOR_MAP:
        sel = rng.random(n) < dens
        o[sel, rng.integers(0, 32, sel.sum())] = rng.integers(1, 256, sel.sum()).astype(np.uint8)
        o[rng.random(n) < dens, 0] = 1  # FREE bit only: not occupied
    else:
        sel = rng.random(n) < dens
        o[sel] = rng.integers(-127, 128, sel.sum()).astype(np.int8)
    g = gvx.Map(mtype, dims, 1.0)
    g.write(o)
    return g, o


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 4), (4, 0), (4, 4)])
@pytest.mark.parametrize("offset", [None, (0, 0, 0), (1, -2, 3), (-3, 0, -1), (0, 0, 40)])
def test_collide_parity(gvx_lib, ta, tb, offset):
    gvx = gvx_lib
    rng = np.random.default_rng(30 + ta + tb)
    dims = (31, 23, 19)
    ga, oa = _random_map(gvx, rng, ta, dims, 0.3)
    gb, ob = _random_map(gvx, rng, tb, dims, 0.3)
    for thr in (0.0, 0.1, 0.5, 0.7, 1.0):
        n = ga.collide(gb, thr, offset)
        assert n == O.map_collide(ta, oa, tb, ob, dims, thr, (0, 0, 0) if offset is None else offset)
        assert np.array_equal(ga.read(), oa)  # eBVM_COLLISION side effect on self


@pytest.mark.parametrize("ta,tb", [(0, 4), (4, 0), (4, 4), (0, 0)])
def test_collide_large_sparse(gvx_lib, ta, tb):
    # more voxels than one grid-stride pass covers (16 Prob voxels per lane), sparse
    # content as in a robot-vs-environment check, odd voxel count, signed offsets
    gvx = gvx_lib
    rng = np.random.default_rng(60 + ta + tb)
    dims = (257, 250, 201)
    ga, oa = _random_map(gvx, rng, ta, dims, 0.002)
    gb, ob = _random_map(gvx, rng, tb, dims, 0.01)
    for thr, offset in ((1.0, None), (0.3, (5, -7, 2)), (0.55, (-1, 0, -3))):
        n = ga.collide(gb, thr, offset)
        assert n == O.map_collide(ta, oa, tb, ob, dims, thr, (0, 0, 0) if offset is None else offset)
        assert n > 0 or thr == 1.0
        assert np.array_equal(ga.read(), oa)
    if ta == 0:
        for thr in (0.2, 1.0):
            n, mask = ga.collide_types(gb, thr)
            n2, mask2 = O.map_collide_types(0, oa, tb, ob, dims, thr)
            assert n == n2 and np.array_equal(mask, mask2)
            assert np.array_equal(ga.read(), oa)


@pytest.mark.parametrize("ta,tb", [(0, 4), (4, 0), (4, 4)])
def test_collide_prob_every_voxel_passes(gvx_lib, ta, tb):
    # a fully known Prob map: at threshold 0.0 every voxel of every lane passes, so a
    # wave's list of passing voxels is full (64 lanes x 16)
    gvx = gvx_lib
    rng = np.random.default_rng(70 + ta + tb)
    dims = (64, 33, 17)
    ga, oa = _random_map(gvx, rng, ta, dims, 1.0 if ta == 4 else 0.5)
    gb, ob = _random_map(gvx, rng, tb, dims, 1.0 if tb == 4 else 0.5)
    for thr, offset in ((0.0, None), (0.0, (3, -1, 2)), (0.6, None)):
        n = ga.collide(gb, thr, offset)
        assert n == O.map_collide(ta, oa, tb, ob, dims, thr, (0, 0, 0) if offset is None else offset)
        assert np.array_equal(ga.read(), oa)
    if ta == 0:
        n, mask = ga.collide_types(gb, 0.0)
        n2, mask2 = O.map_collide_types(0, oa, tb, ob, dims, 0.0)
        assert n == n2 and np.array_equal(mask, mask2)
        assert np.array_equal(ga.read(), oa)


@pytest.mark.parametrize("tb", [0, 4])
def test_collide_types_parity(gvx_lib, tb):
    gvx = gvx_lib
    rng = np.random.default_rng(40 + tb)
    dims = (40, 30, 20)
    ga, oa = _random_map(gvx, rng, 0, dims, 0.2)
    gb, ob = _random_map(gvx, rng, tb, dims, 0.4)
    for thr in (0.0, 0.2, 0.5, 1.0):
        n, mask = ga.collide_types(gb, thr)
        n2, mask2 = O.map_collide_types(0, oa, tb, ob, dims, thr)
        assert n == n2 and np.array_equal(mask, mask2)
        assert np.array_equal(ga.read(), oa)


def test_clear_and_clear_bits(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(50)
    g, o = _random_map(gvx, rng, 0, (16, 16, 16), 0.5)
    bits = rng.integers(0, 256, 32).astype(np.uint8)
    g.clear_bits(bits)
    O.map_clear_bits(o, (16, 16, 16), bits)
    assert np.array_equal(g.read(), o)
    for t in TYPES:
        m = gvx.Map(t, (7, 5, 3), 1.0)
        m.insert_coords(np.array([[1, 1, 1]], np.uint32), 1)
        m.clear()
        assert np.array_equal(m.read(), O.new_map(t, (7, 5, 3)))


# ------------------------------------------------------------ reference known answers on the GPU
def _prob(gvx, pts):
    m = gvx.Map(O.PROB_MAP, DIMS, 1.0)
    m.insert_points(pts, 1)
    return m


def test_known_answers_collision(gvx_lib):
    gvx = gvx_lib
    pts = create_equidistant_points_in_box(NUM_POINTS, DIMS, 1.0)
    assert _prob(gvx, pts).collide(_prob(gvx, pts), 0.1) == NUM_POINTS          # testing_voxelmap.cu:96-115
    p1 = create_box_of_points((2.1, 2.1, 2.1), (4.1, 4.1, 4.1), 0.5)
    p2 = create_box_of_points((3.1, 3.1, 3.1), (5.1, 5.1, 5.1), 0.5)
    assert _prob(gvx, p1).collide(_prob(gvx, p2), 0.1) == 8                     # testing_voxelmap.cu:119-142
    assert _prob(gvx, p1).collide(_prob(gvx, p2), 0.1, (-1, 0, -1)) == 18
    assert _prob(gvx, p2).collide(_prob(gvx, p1), 0.1, (1, 0, 1)) == 18
    black, white = create_checkerboard(NUM_POINTS, DIMS, 1.0)
    assert _prob(gvx, black).collide(_prob(gvx, white), 0.1) == 0              # testing_voxelmap.cu:144-168


def test_known_answer_types_probab(gvx_lib):
    gvx = gvx_lib
    pts = np.array([[2, 2, 2], [2, 2, 3], [2, 3, 2], [2, 3, 3], [3, 2, 2], [3, 2, 3], [3, 3, 2], [3, 3, 3]],
                   np.float32)
    a = gvx.Map(O.BITVECTOR_MAP, DIMS, 1.0)
    for m in (14, 46, 103):
        a.insert_points(pts, m)
    n, mask = a.collide_types(_prob(gvx, pts), float(np.float32(199.0 / 254.0)))   # testing_voxelmap.cu:313-360
    assert n == 8 and bits_of(mask) == [14, 46, 103]


# ------------------------------------------------------------ cfg1 and cfg2 workloads
def test_cfg1_scene(gvx_lib):
    gvx = gvx_lib
    from gvx import scenes
    c = scenes.cfg1()
    d, s = c["dims"], c["side"]
    for meaning, expect_zero in ((0, True), (1, False)):
        env = gvx.Map(O.BITVECTOR_MAP, d, s)
        env.insert_points(c["env_points"], meaning)
        oenv = O.new_map(O.BITVECTOR_MAP, d)
        O.map_insert_points(O.BITVECTOR_MAP, oenv, d, s, c["env_points"], meaning)
        assert np.array_equal(env.read(), oenv)
        L = gvx.VoxelList(d, s)
        L.insert_points(c["list_points"], sizes=c["list_sizes"], meanings=c["list_meanings"])
        OL = O.List(d, s)
        OL.insert_points(c["list_points"], sizes=c["list_sizes"], meanings=c["list_meanings"])
        n, mask = L.collide_map(env, 1.0, mode=1)
        n2, mask2 = OL.collide_map(O.BITVECTOR_MAP, oenv, 1.0, mode=1)
        assert n == n2 and np.array_equal(mask, mask2)
        assert (n == 0) == expect_zero   # meaning bit 0 alone never counts as occupied


def test_cfg2_scene(gvx_lib):
    gvx = gvx_lib
    from gvx import scenes
    c = scenes.cfg2()
    d, s = c["dims"], c["side"]
    env = gvx.Map(O.PROB_MAP, d, s)
    assert not env.insert_points(c["env_points"], 1)
    oenv = O.new_map(O.PROB_MAP, d)
    O.map_insert_points(O.PROB_MAP, oenv, d, s, c["env_points"], 1)
    assert np.array_equal(env.read(), oenv)
    r = c["arm"]
    T = O.fk_chain(r["dh"], c["base"], c["q"])
    off = np.concatenate([[0], np.cumsum(r["link_sizes"])]).astype(np.int64)
    posed = np.concatenate([O.transform_points(T[l], r["link_points"][off[l]:off[l + 1]]) for l in range(7)])
    robot = gvx.Map(O.BITVECTOR_MAP, d, s)
    meanings = np.arange(4, 11, dtype=np.uint32)
    robot.insert_meta(posed, r["link_sizes"], meanings)
    orob = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_meta(O.BITVECTOR_MAP, orob, d, s, posed, r["link_sizes"], meanings)
    n, mask = robot.collide_types(env, 1.0)
    n2, mask2 = O.map_collide_types(O.BITVECTOR_MAP, orob, O.PROB_MAP, oenv, d, 1.0)
    assert n == n2 and np.array_equal(mask, mask2)
    assert robot.collide(env, 1.0) == O.map_collide(O.BITVECTOR_MAP, orob, O.PROB_MAP, oenv, d, 1.0)
