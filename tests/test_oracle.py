"""CPU tests that pin the oracle: reference known answers, closed forms, brute force."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from refgen import (DIM_X, DIM_Y, DIM_Z, NUM_POINTS, bits_of, bitvector_stream, create_box_of_points,
                    create_checkerboard, create_equidistant_points_in_box, set_bits)

DIMS = (DIM_X, DIM_Y, DIM_Z)


# ------------------------------------------------------------------ scalar rules
def test_sincos_accuracy():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-8, 8, 20000), np.arange(-16, 17) * np.pi / 4]).astype(np.float32)
    for x in xs:
        s, c = O.sincosf(x)
        assert abs(float(s) - math.sin(float(x))) <= 2.5e-7
        assert abs(float(c) - math.cos(float(x))) <= 2.5e-7


def test_quantize_closed_form():
    rng = np.random.default_rng(1)
    p = rng.uniform(-1, 6, 20000).astype(np.float32)
    side = np.float32(0.01)
    ref = np.floor(p / side)  # numpy float32 division is IEEE
    for v, r in zip(p, ref):
        expect = 0 if r < 0 else int(r)
        assert O.quantize(v, side) == expect
    assert O.quantize(float("nan"), 0.01) == 0
    assert O.quantize(-1e-30, 0.01) == 0
    assert O.quantize(1e30, 0.01) == 0xFFFFFFFF
    # IEEE division, not multiplication by the reciprocal: 0.3 / 0.1 rounds below 3 in float32
    assert O.quantize(np.float32(0.3), np.float32(0.1)) == int(np.floor(np.float32(0.3) / np.float32(0.1)))


def test_bitvector_stream_ground_truth():
    # test/testing_voxelmap.cu:482-516 (iostream_bitvoxel)
    m = O.new_map(O.BITVECTOR_MAP, (1, 1, 1))
    for meaning in (0, 2, 3, 4, 111, 254, 255):
        O.map_insert_coords(O.BITVECTOR_MAP, m, (1, 1, 1), [[0, 0, 0]], meaning)
    truth = ("10111" + "0" * 106 + "1" + "0" * 142 + "11")
    assert len(truth) == 256
    assert bitvector_stream(m[0]) == truth


def test_prob_rules():
    U = -128
    # ProbabilisticVoxel::insert: FREE -10, OCCUPIED/COLLISION +72, SV m -> m - 129, UNKNOWN -> -128
    assert O.lib().gvo_prob_insert(U, 1) == 72
    assert O.lib().gvo_prob_insert(U, 0) == -10
    assert O.lib().gvo_prob_insert(U, 2) == 72
    assert O.lib().gvo_prob_insert(U, 4) == 4 - 129
    assert O.lib().gvo_prob_insert(U, 129) == 0
    assert O.lib().gvo_prob_insert(U, 254) == 125
    assert O.lib().gvo_prob_insert(50, 3) == U
    assert O.lib().gvo_prob_insert(50, 255) == U
    assert O.lib().gvo_prob_insert(100, 1) == 127          # saturates
    assert O.lib().gvo_prob_insert(-120, 0) == -127        # saturates at MIN, never reaches UNKNOWN
    # k equal updates commute: closed form clamp(v + k*d)
    for v in (-127, -5, 0, 90):
        cur = v
        for _ in range(5):
            cur = O.lib().gvo_prob_insert(cur, 1)
        assert cur == max(-127, min(127, v + 5 * 72))


def test_float_to_probability():
    assert O.float_to_probability(1.0) == 127
    assert O.float_to_probability(0.0) == -127
    assert O.float_to_probability(2.0) == 127
    assert O.float_to_probability(0.1) == -101   # 25.4 - 127 = -101.6, truncated toward zero
    assert O.float_to_probability(0.5) == 0


def test_bitmargin_margin0_is_masked_and():
    rng = np.random.default_rng(2)
    for _ in range(300):
        a = rng.integers(0, 256, 32, dtype=np.uint8)
        b = rng.integers(0, 256, 32, dtype=np.uint8)
        hit, out = O.bitmargin(a, b, 0, 0)
        expect = a & b
        expect[0] &= 0xF0
        assert np.array_equal(out, expect)
        assert hit == bool(expect.any())


def test_bitmargin_window_brute_force():
    # with margin m, SV bit i of v1 collides when v2 has an SV bit in [i-m, i+m]
    rng = np.random.default_rng(3)
    for m in (1, 2, 3):
        for _ in range(100):
            a = (rng.random(256) < 0.05)
            b = (rng.random(256) < 0.05)
            hit, out = O.bitmargin(np.packbits(a, bitorder="little"), np.packbits(b, bitorder="little"), m, 0)
            ob = np.unpackbits(out, bitorder="little").astype(bool)
            for i in range(4, 254):
                if i >= 248 and m > 0:
                    continue  # the sliding buffer stops feeding bytes near the end
                lo, hi = max(4, i - m), min(255, i + m)
                expect = a[i] and b[lo:hi + 1].any()
                assert ob[i] == expect, (m, i)


# ------------------------------------------------------------------ map known answers
def _prob_map_with(points):
    m = O.new_map(O.PROB_MAP, DIMS)
    O.map_insert_points(O.PROB_MAP, m, DIMS, 1.0, points, 1)
    return m


def test_collision_known_answer():
    # test/testing_voxelmap.cu:96-115
    pts = create_equidistant_points_in_box(NUM_POINTS, DIMS, 1.0)
    a, b = _prob_map_with(pts), _prob_map_with(pts)
    assert O.map_collide(O.PROB_MAP, a, O.PROB_MAP, b, DIMS, 0.1) == NUM_POINTS


def test_collision_with_offset_known_answer():
    # test/testing_voxelmap.cu:119-142
    p1 = create_box_of_points((2.1, 2.1, 2.1), (4.1, 4.1, 4.1), 0.5)
    p2 = create_box_of_points((3.1, 3.1, 3.1), (5.1, 5.1, 5.1), 0.5)
    a, b = _prob_map_with(p1), _prob_map_with(p2)
    assert O.map_collide(O.PROB_MAP, a, O.PROB_MAP, b, DIMS, 0.1) == 8
    a, b = _prob_map_with(p1), _prob_map_with(p2)
    assert O.map_collide(O.PROB_MAP, a, O.PROB_MAP, b, DIMS, 0.1, (-1, 0, -1)) == 18
    a, b = _prob_map_with(p1), _prob_map_with(p2)
    assert O.map_collide(O.PROB_MAP, b, O.PROB_MAP, a, DIMS, 0.1, (1, 0, 1)) == 18


def test_no_collision_checkerboard():
    # test/testing_voxelmap.cu:144-168
    black, white = create_checkerboard(NUM_POINTS, DIMS, 1.0)
    a, b = _prob_map_with(black), _prob_map_with(white)
    assert O.map_collide(O.PROB_MAP, a, O.PROB_MAP, b, DIMS, 0.1) == 0


def test_collision_with_types_probab():
    # test/testing_voxelmap.cu:313-360
    pts = np.array([[2, 2, 2], [2, 2, 3], [2, 3, 2], [2, 3, 3], [3, 2, 2], [3, 2, 3], [3, 3, 2], [3, 3, 3]],
                   np.float32)
    a = O.new_map(O.BITVECTOR_MAP, DIMS)
    for m in (14, 46, 103):
        O.map_insert_points(O.BITVECTOR_MAP, a, DIMS, 1.0, pts, m)
    b = _prob_map_with(pts)
    thr = np.float32((72.0 + 127.0) / 254.0)  # probabilityToFloat(cSENSOR_MODEL_OCCUPIED)
    n, mask = O.map_collide_types(O.BITVECTOR_MAP, a, O.PROB_MAP, b, DIMS, float(thr))
    assert n == 8
    assert bits_of(mask) == [14, 46, 103]


def test_collision_marks_self():
    pts = create_box_of_points((2.1, 2.1, 2.1), (4.1, 4.1, 4.1), 0.5)
    a = O.new_map(O.BITVECTOR_MAP, DIMS)
    O.map_insert_points(O.BITVECTOR_MAP, a, DIMS, 1.0, pts, 7)
    b = _prob_map_with(pts)
    assert O.map_collide(O.BITVECTOR_MAP, a, O.PROB_MAP, b, DIMS, 0.5) == 27
    occ = ~O._none_but_empty(a)
    assert np.all(a[occ][:, 0] & 4)  # eBVM_COLLISION set on every colliding voxel


# ------------------------------------------------------------------ list known answers
def _two_boxes():
    c1 = create_box_of_points((1.1, 1.1, 1.1), (3.9, 3.9, 3.9), 0.1)
    c2 = create_box_of_points((2.1, 2.1, 2.1), (4.9, 4.9, 4.9), 0.1)
    return np.concatenate([c1, c2]), [len(c1), len(c2)]


def test_list_insert_meta_46_and_8():
    # test/testing_voxellist.cu:284-341
    pts, sizes = _two_boxes()
    L = O.List(DIMS, 1.0)
    L.insert_points(pts, sizes=sizes, meanings=[11, 12])
    assert len(L) == 46
    both = set_bits(11, 12)
    sel = np.all(L.bits == both, axis=1)
    assert sel.sum() == 8
    assert np.all((L.coords[sel] >= 2) & (L.coords[sel] <= 3))
    assert np.all(np.diff(L.ids.astype(np.int64)) > 0)


def test_list_collide_prob_and_types():
    # test/testing_voxellist.cu:111-238
    pts, sizes = _two_boxes()
    prob = O.new_map(O.PROB_MAP, DIMS)
    O.map_insert_meta(O.PROB_MAP, prob, DIMS, 1.0, pts, sizes, None, 1)
    bitm = O.new_map(O.BITVECTOR_MAP, DIMS)
    O.map_insert_meta(O.BITVECTOR_MAP, bitm, DIMS, 1.0, pts, sizes, None, 1)
    thr = float(np.float32(199.0 / 254.0))
    L = O.List(DIMS, 1.0)
    L.insert_points(pts, sizes=sizes, meanings=[11, 12])
    assert L.collide_map(O.PROB_MAP, prob, thr, mode=0)[0] == 46
    for mt, m in ((O.PROB_MAP, prob), (O.BITVECTOR_MAP, bitm)):
        L2 = O.List(DIMS, 1.0)
        L2.insert_points(pts, sizes=sizes, meanings=[11, 12])
        n, mask = L2.collide_map(mt, m, 1.0, mode=1)
        assert n == 46 and bits_of(mask) == [11, 12]
        n, mask = L2.collide_map(mt, m, 1.0, mode=1)   # second call: eBVM_COLLISION now in the entries
        assert n == 46 and bits_of(mask) == [2, 11, 12]


def test_list_shifting():
    # test/testing_voxellist.cu:241-281
    box = create_box_of_points((1.1, 1.1, 1.1), (3.9, 3.9, 3.9), 0.1)
    L = O.List(DIMS, 1.0)
    L.insert_points(box, meaning=1)
    prob = O.new_map(O.PROB_MAP, DIMS)
    O.map_insert_points(O.PROB_MAP, prob, DIMS, 1.0, box, 1)
    thr = float(np.float32(199.0 / 254.0))
    for shift, expect in ((0.0, 27), (0.5, 27), (1.0, 18), (1.5, 18), (2.0, 9), (2.5, 9), (3.0, 0), (3.5, 0)):
        assert L.collide_map(O.PROB_MAP, prob, thr, offset=(int(shift), 0, 0), mode=0)[0] == expect


def test_list_bitmasked_collision():
    # test/testing_voxellist.cu:859-985
    side = 0.025
    L = O.List(DIMS, side)
    prob = O.new_map(O.PROB_MAP, DIMS)
    O.map_insert_points(O.PROB_MAP, prob, DIMS, side, create_box_of_points((0.111, 0.111, 0.111),
                                                                         (0.799, 0.399, 0.399), 0.005), 1)
    for lo, m in (((0.111, 0.311, 0.111), 34), ((0.311, 0.311, 0.111), 63), ((0.511, 0.311, 0.111), 102)):
        hi = (lo[0] + 0.288, 0.599, 0.399)
        L.insert_points(create_box_of_points(lo, hi, 0.005), meaning=m)
    expect = {(0, 0, 0): [0, 0, 9, 15, 21], (0, -4, 0): [0, 0, 18, 30, 42], (0, -4, -4): [0, 0, 12, 20, 28]}
    for off, counts in expect.items():
        tm = np.zeros(32, np.uint8)
        got = [L.collide_map(O.PROB_MAP, prob, 1.0, offset=off, mode=2, typemask=tm)[0]]
        for b in (33, 34, 63, 102):
            tm |= set_bits(b)
            got.append(L.collide_map(O.PROB_MAP, prob, 1.0, offset=off, mode=2, typemask=tm)[0])
        assert got == [c * 64 for c in counts], off


def test_signed_offset_known_answers():
    # test/testing_voxelmap.cu:59-93: getVoxelIndexSigned on dims (15, 16, 17), observed through
    # the id shift that a list merge with an offset applies
    d = (15, 16, 17)
    for off, truth in (((0, -2, 0), -30), ((1, 1, 1), 256), ((1, -2, -3), -749), ((3, -2, -1), -267),
                       ((-1, -2, 3), 689)):
        src = O.List(d, 1.0)
        src.insert_coords([[7, 8, 8]], 5)
        dst = O.List(d, 1.0)
        dst.merge_list(src, offset=off)
        assert int(dst.ids[0]) - int(src.ids[0]) == truth


def test_list_collide_list():
    a = O.List(DIMS, 1.0)
    b = O.List(DIMS, 1.0)
    a.insert_points(create_box_of_points((1.1, 1.1, 1.1), (3.9, 3.9, 3.9), 0.5), meaning=20)
    b.insert_points(create_box_of_points((2.1, 2.1, 2.1), (4.9, 4.9, 4.9), 0.5), meaning=30)
    n, mask = a.collide_list(b)
    assert n == 8 and bits_of(mask) == [20, 30]


def test_list_merge_map_roundtrip():
    rng = np.random.default_rng(5)
    d = (20, 17, 9)
    m = O.new_map(O.BITVECTOR_MAP, d)
    c = rng.integers(0, (20, 17, 9), size=(300, 3)).astype(np.uint32)
    O.map_insert_coords(O.BITVECTOR_MAP, m, d, c, 9)
    O.map_insert_coords(O.BITVECTOR_MAP, m, d, c[:40], 0)   # bit 0 alone is not "occupied"
    L = O.List(d, 1.0)
    L.merge_map(m)
    ids = np.unique(c[:, 2].astype(np.int64) * 340 + c[:, 1] * 20 + c[:, 0])
    assert np.array_equal(L.ids, ids.astype(np.uint32))
    assert np.array_equal(L.bits, m[ids])


# ------------------------------------------------------------------ EDT
def _random_obstacles(rng, d, dens):
    n = int(np.prod(d))
    u = O.dist_uninit(d)
    m = np.full(n, u, np.uint32)
    idx = np.nonzero(rng.random(n) < dens)[0]
    x, y, z = idx % d[0], (idx // d[0]) % d[1], idx // (d[0] * d[1])
    m[idx] = x | (y << 10) | (z << 20)
    return m


@pytest.mark.parametrize("seed", range(12))
def test_pba_equals_bruteforce(seed):
    rng = np.random.default_rng(100 + seed)
    d = tuple(int(v) for v in rng.integers(1, 14, 3))
    for dens in (0.0, 0.003, 0.05, 0.3, 1.0):
        m = _random_obstacles(rng, d, dens)
        bf = O.edt_bruteforce(m, d)
        p = m.copy()
        O.pba(p, d)
        assert np.array_equal(p, bf), (d, dens)


@pytest.mark.parametrize("dens", [0.0, 0.002, 0.05, 0.5])
def test_pba_threaded_equals_serial(dens):
    """The OpenMP oracle (lines of each pass over threads) checks the 1024^3 GPU grids; it must give
    the serial oracle's grid, which the brute-force tests pin."""
    rng = np.random.default_rng(int(dens * 1000) + 5)
    for d in ((37, 29, 41), (64, 64, 64), (5, 70, 3)):
        m = _random_obstacles(rng, d, dens)
        a, b = m.copy(), m.copy()
        O.pba(a, d)
        O.pba(b, d, threads=4)
        assert np.array_equal(a, b), (d, dens)


def test_pba_single_obstacle_closed_form():
    d = (17, 9, 23)
    m = _random_obstacles(np.random.default_rng(0), d, 0.0)
    ox, oy, oz = 5, 7, 11
    m[oz * 17 * 9 + oy * 17 + ox] = ox | (oy << 10) | (oz << 20)
    O.pba(m, d)
    sq = O.dist_squared(m, d)
    z, y, x = np.meshgrid(np.arange(23), np.arange(9), np.arange(17), indexing="ij")
    assert np.array_equal(sq, ((x - ox) ** 2 + (y - oy) ** 2 + (z - oz) ** 2).ravel())


def test_pba_matches_exact_distances3d():
    # test/testing_distance.cu:39-127, scaled to 32 x 32 x 35 so the O(N M) scan stays short
    d = (32, 32, 35)
    obst = create_equidistant_points_in_box(2000, d, 1.0)
    m = O.new_map(O.DIST_MAP, d)
    O.map_insert_points(O.DIST_MAP, m, d, 1.0, obst, 1)
    ex = m.copy()
    O.exact_distances(ex, d, 1.0, obst)
    p = m.copy()
    O.pba(p, d)
    assert np.array_equal(O.dist_squared(p, d), O.dist_squared(ex, d))


@pytest.mark.parametrize("mtype", [O.BITVECTOR_MAP, O.PROB_MAP, O.DIST_MAP])
def test_threaded_insert_and_collide_equal_serial(mtype):
    # the OpenMP variants used as multi-core CPU baselines give the serial results: points
    # repeated on one voxel (Prob saturation depends on their order), points outside
    rng = np.random.default_rng(11 + mtype)
    d, side = (40, 30, 20), 0.1
    p = (rng.integers(-2, 42, size=(3000, 3)) + 0.5).astype(np.float32) * np.float32(side)
    p = np.repeat(p, rng.integers(1, 40, size=3000), axis=0)
    a, b = O.new_map(mtype, d), O.new_map(mtype, d)
    oa = O.map_insert_points(mtype, a, d, side, p, 1)
    ob = O.map_insert_points(mtype, b, d, side, p, 1, threads=7)
    assert oa == ob and oa
    assert np.array_equal(a, b)
    if mtype == O.DIST_MAP:
        return
    other = O.new_map(O.PROB_MAP, d)
    O.map_insert_points(O.PROB_MAP, other, d, side, p[::3], 1)
    for off in ((0, 0, 0), (2, -1, 3)):
        a2, b2 = a.copy(), b.copy()
        ca = O.map_collide(mtype, a2, O.PROB_MAP, other, d, 0.6, off)
        cb = O.map_collide(mtype, b2, O.PROB_MAP, other, d, 0.6, off, threads=5)
        assert ca == cb and ca > 0
        assert np.array_equal(a2, b2)


def test_merge_occupied_closed_form():
    # DistanceVoxelMap::mergeOccupied (DistanceVoxelMap.hpp:109-136, DistanceVoxelMap.h:92-134): a voxel
    # whose log-odds is strictly above floatToProbability(threshold) becomes an obstacle at its own
    # coordinates plus the offset; every other voxel keeps what it held
    d = (5, 4, 3)
    n = 5 * 4 * 3
    rng = np.random.default_rng(7)
    for thr, off in ((0.5, (0, 0, 0)), (0.8, (3, 1, 2)), (0.2, (0, 7, 0))):
        t = int(O.float_to_probability(thr))
        prob = rng.choice(np.array([-128, -127, t - 1, t, t + 1, 127], np.int8), n).astype(np.int8)
        old = rng.integers(0, 1 << 30, n, dtype=np.uint32)
        dist = old.copy()
        O.dist_merge_occupied(dist, prob, d, off, thr)
        i = np.arange(n)
        x, y, z = i % 5, (i // 5) % 4, i // 20
        want = np.where(prob.astype(np.int32) > t,
                        ((x + off[0]) | ((y + off[1]) << 10) | ((z + off[2]) << 20)).astype(np.uint32), old)
        assert np.array_equal(dist, want), thr
        assert (prob == t).any()  # the strict comparison is exercised


def test_dist_uninit_extension():
    assert O.dist_uninit((512, 512, 512)) == 0x3FFFFFFF
    assert O.dist_uninit((1024, 1024, 1024)) == 0xFFFFFFFF


def test_pba_1024_wide_line():
    # a 1024-long x axis: coordinate 1023 is valid, the sentinel is 0xFFFFFFFF
    d = (1024, 1, 3)
    m = np.full(1024 * 3, 0xFFFFFFFF, np.uint32)
    m[1 * 1024 + 1023] = 1023 | (0 << 10) | (1 << 20)
    O.pba(m, d)
    sq = O.dist_squared(m, d)
    x = np.arange(1024)
    for z in range(3):
        assert np.array_equal(sq[z * 1024:(z + 1) * 1024], (x - 1023) ** 2 + (z - 1) ** 2)


# ------------------------------------------------------------------ robot, batch, cost
def test_dh_fk_closed_form():
    rng = np.random.default_rng(7)
    dh = np.stack([np.zeros(7), rng.uniform(0.05, 0.4, 7), rng.uniform(0.05, 0.4, 7),
                   rng.choice([-np.pi / 2, 0, np.pi / 2], 7), np.zeros(7)], axis=1).astype(np.float32)
    base = np.array([1, 0, 0, 0.5, 0, 1, 0, 0.2, 0, 0, 1, 0.1], np.float32)
    for _ in range(20):
        q = rng.uniform(-np.pi, np.pi, 7).astype(np.float32)
        T = O.fk_chain(dh, base, q)
        t = np.eye(4)
        t[:3, :] = base.reshape(3, 4)
        for l in range(7):
            assert np.allclose(T[l].reshape(3, 4), t[:3, :], atol=2e-6)
            th, dd, a, al = float(q[l]), float(dh[l, 1]), float(dh[l, 2]), float(dh[l, 3])
            A = np.array([[np.cos(th), -np.sin(th) * np.cos(al), np.sin(th) * np.sin(al), a * np.cos(th)],
                          [np.sin(th), np.cos(th) * np.cos(al), -np.cos(th) * np.sin(al), a * np.sin(th)],
                          [0, np.sin(al), np.cos(al), dd], [0, 0, 0, 1]])
            t = t @ A


def test_batch_oracle_equals_primitive_pipeline():
    """The batched oracle equals clearMap + FK + insertMetaPointCloud + collideWith composed
    from the oracle's primitive functions (count), and the list-form collideWithTypes (mask)."""
    rng = np.random.default_rng(8)
    d = (64, 64, 64)
    side = 0.02
    env = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_points(O.BITVECTOR_MAP, env, d, side, rng.uniform(0, 1.28, (6000, 3)).astype(np.float32), 1)
    env_bits = O.occupancy_bits(env)
    nl = 4
    dh = np.stack([np.zeros(nl), rng.uniform(0.05, 0.3, nl), rng.uniform(0.05, 0.3, nl),
                   rng.choice([-np.pi / 2, 0, np.pi / 2], nl), np.zeros(nl)], axis=1).astype(np.float32)
    sizes = np.array([300, 500, 200, 400], np.uint32)
    pts = rng.uniform(-0.1, 0.3, (sizes.sum(), 3)).astype(np.float32)
    base = np.array([1, 0, 0, 0.64, 0, 1, 0, 0.64, 0, 0, 1, 0.3], np.float32)
    q = rng.uniform(-np.pi, np.pi, (6, nl)).astype(np.float32)
    counts, masks = O.batch_collide(env_bits, d, side, pts, sizes, dh, base, q)
    for p in range(q.shape[0]):
        T = O.fk_chain(dh, base, q[p])
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        posed = np.concatenate([O.transform_points(T[l], pts[off[l]:off[l + 1]]) for l in range(nl)])
        robot = O.new_map(O.BITVECTOR_MAP, d)
        O.map_insert_meta(O.BITVECTOR_MAP, robot, d, side, posed, sizes, np.arange(4, 4 + nl, dtype=np.uint32))
        assert O.map_collide(O.BITVECTOR_MAP, robot, O.BITVECTOR_MAP, env, d, 1.0) == counts[p]
        L = O.List(d, side)
        L.insert_points(posed, sizes=sizes, meanings=np.arange(4, 4 + nl))
        n, mask = L.collide_map(O.BITVECTOR_MAP, env, 1.0, mode=1)
        assert n == counts[p] and np.array_equal(mask, masks[p])
    # the OpenMP variant (all-cores CPU baseline) gives the same results
    q2 = rng.uniform(-np.pi, np.pi, (37, nl)).astype(np.float32)
    c1, m1 = O.batch_collide(env_bits, d, side, pts, sizes, dh, base, q2)
    c4, m4 = O.batch_collide(env_bits, d, side, pts, sizes, dh, base, q2, threads=4)
    assert np.array_equal(c1, c4) and np.array_equal(m1, m4) and c1.sum() > 0


def test_cost_oracle_gradient_fd():
    rng = np.random.default_rng(9)
    d = (24, 24, 24)
    side = 0.05
    grid = rng.uniform(0.0, 0.3, int(np.prod(d))).astype(np.float32)
    nl, spl = 3, 5
    dh = np.stack([np.zeros(nl), rng.uniform(0.05, 0.2, nl), rng.uniform(0.05, 0.2, nl),
                   rng.choice([-np.pi / 2, 0, np.pi / 2], nl), np.zeros(nl)], axis=1)
    base = np.array([1, 0, 0, 0.6, 0, 1, 0, 0.6, 0, 0, 1, 0.5], np.float64)
    sph = np.concatenate([rng.uniform(-0.05, 0.15, (nl, spl, 3)), rng.uniform(0.1, 0.3, (nl, spl, 1))], axis=2)
    q = rng.uniform(-np.pi, np.pi, (2, 3, nl))
    c, dq, _ = O.cost_d(grid, d, side, dh, base, sph, q, 0.05, 1.0)
    h = 1e-6
    checked = 0
    for t in range(2):
        for w in range(3):
            for j in range(nl):
                qp, qm = q.copy(), q.copy()
                qp[t, w, j] += h
                qm[t, w, j] -= h
                fd = (O.cost_d(grid, d, side, dh, base, sph, qp, 0.05, 1.0, False)[0][t] -
                      O.cost_d(grid, d, side, dh, base, sph, qm, 0.05, 1.0, False)[0][t]) / (2 * h)
                if abs(fd) > 1e-3:
                    assert abs(fd - dq[t, w, j]) <= 1e-4 * abs(fd) + 1e-6
                    checked += 1
    assert checked >= 3


# ---------------------------------------------------------------- sensor data (insertSensorData)
def test_ray_cells_properties():
    # RayCaster::rayCast: starts at the sensor voxel, ends at the point voxel, one axis step per
    # cell (6-connected), 1 + |dx| + |dy| + |dz| cells, and never far from the straight segment
    rng = np.random.default_rng(21)
    for _ in range(3000):
        f = rng.integers(0, 60, 3)
        t = rng.integers(0, 60, 3)
        c = O.ray_cells(f, t).astype(np.int64)
        assert len(c) == 1 + int(np.abs(t - f).sum())
        assert np.array_equal(c[0], f) and np.array_equal(c[-1], t)
        assert np.all(np.abs(np.diff(c, axis=0)).sum(axis=1) == 1)
        seg = (t - f).astype(np.float64)
        L2 = float(seg @ seg)
        if L2 > 0:
            u = np.clip(((c - f) @ seg) / L2, 0, 1)
            dist = np.linalg.norm(c - (f + u[:, None] * seg), axis=1)
            assert dist.max() <= 1.5
    # axis-aligned and diagonal closed forms
    assert np.array_equal(O.ray_cells([2, 5, 7], [6, 5, 7]), [[x, 5, 7] for x in range(2, 7)])
    assert np.array_equal(O.ray_cells([3, 3, 3], [3, 3, 0]), [[3, 3, z] for z in (3, 2, 1, 0)])


# Walks traced by hand through RayCaster::rayCast (voxelmap/kernels/VoxelMapOperations.h:302-335):
# error_xy = dx - dy, error_xz = dx - dz, error_yz = dy - dz; per cell, step x while
# error_xy > 0 and error_xz > 0, else y while error_yz > 0, else z.  The lists below were worked
# out step by step on paper (errors written beside each cell), not produced by the oracle; every
# one is non-axis-aligned and the tie cases (an error exactly 0) fix which axis moves first.
RAY_WALKS = [
    # (from, to, cells); the comment gives (error_xy, error_xz, error_yz) and the axis stepped
    # after each cell but the last
    # (1,2,1) x; (-1,2,1) y; (3,2,1) x
    ((0, 0, 0), (2, 1, 0), [(0, 0, 0), (1, 0, 0), (1, 1, 0), (2, 1, 0)]),
    # (0,1,1) y (error_xy == 0: y before x); (2,1,1) x
    ((0, 0, 0), (1, 1, 0), [(0, 0, 0), (0, 1, 0), (1, 1, 0)]),
    # (-1,-1,0) z (error_yz == 0: z before y); (-1,-1,2) y
    ((0, 0, 0), (0, 1, 1), [(0, 0, 0), (0, 0, 1), (0, 1, 1)]),
    # (0,0,0) z; (0,2,2) y; (2,2,0) x
    ((0, 0, 0), (1, 1, 1), [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1)]),
    # negative increments: (1,-1,-2) z; (1,3,0) x; (-1,-3,0) z; (-1,1,2) y; (3,1,-4) x; (1,-5,-4) z
    ((5, 5, 5), (3, 4, 2), [(5, 5, 5), (5, 5, 4), (4, 5, 4), (4, 5, 3), (4, 4, 3), (3, 4, 3), (3, 4, 2)]),
    # (1,3,2) x; (-5,1,2) y; (3,1,0) x; (-3,-1,0) z; (-3,7,6) y; (5,7,4) x; (-1,5,4) y; (7,5,2) x
    ((2, 0, 1), (6, 3, 2), [(2, 0, 1), (3, 0, 1), (3, 1, 1), (4, 1, 1), (4, 1, 2), (4, 2, 2), (5, 2, 2),
                            (5, 3, 2), (6, 3, 2)]),
]


@pytest.mark.parametrize("frm,to,cells", RAY_WALKS)
def test_ray_cells_hand_traced(frm, to, cells):
    assert O.ray_cells(frm, to).tolist() == [list(c) for c in cells]


# ---------------------------------------------------------------- trilinear field (clearance cost)
def _multilinear_grid(d):
    """g[i, j, k] = 1 + 2i - j + 3k + ij - 2jk + ik/2 + ijk/4 (exact in float32 for small grids):
    trilinear interpolation reproduces a function in span{1, x, y, z, xy, yz, xz, xyz} exactly."""
    k, j, i = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
    g = 1 + 2 * i - j + 3 * k + i * j - 2 * j * k + 0.5 * i * k + 0.25 * i * j * k
    return g.astype(np.float32).ravel()


def _multilinear(u):
    x, y, z = u
    v = 1 + 2 * x - y + 3 * z + x * y - 2 * y * z + 0.5 * x * z + 0.25 * x * y * z
    g = np.array([2 + y + 0.5 * z + 0.25 * y * z, -1 + x - 2 * z + 0.25 * x * z,
                  3 - 2 * y + 0.5 * x + 0.25 * x * y])
    return v, g


def test_trilinear_voxel_centres_exact():
    # samples sit at voxel centres (i + 1/2) * side: there the interpolant is the sample itself
    rng = np.random.default_rng(41)
    d, side = (7, 5, 6), 0.25
    grid = rng.uniform(-3, 3, int(np.prod(d))).astype(np.float32)
    for _ in range(200):
        i = [int(rng.integers(0, d[a])) for a in range(3)]
        v, _ = O.trilinear_d(grid, d, side, [(i[a] + 0.5) * side for a in range(3)])
        assert v == float(grid[(i[2] * d[1] + i[1]) * d[0] + i[0]])


def test_trilinear_multilinear_field_exact():
    rng = np.random.default_rng(42)
    d, side = (9, 8, 7), 0.25
    grid = _multilinear_grid(d)
    for _ in range(500):
        # interior points (u = x / side - 1/2 in [0, dim - 1])
        u = np.array([rng.uniform(0, d[a] - 1) for a in range(3)])
        x = (u + 0.5) * side
        v, g = O.trilinear_d(grid, d, side, x)
        ev, eg = _multilinear(u)
        assert abs(v - ev) <= 1e-12 * max(1.0, abs(ev))
        assert np.allclose(g, eg / side, rtol=1e-12, atol=1e-12)


def test_trilinear_clamps_outside():
    # outside the centre lattice the coordinate clamps to the border sample and the gradient along
    # that axis is 0; inside along the other axes the field stays multilinear
    d, side = (6, 5, 4), 0.5
    grid = _multilinear_grid(d)
    x = np.array([-1.0, 1.3, 0.9])        # u_x < 0 -> 0
    v, g = O.trilinear_d(grid, d, side, x)
    u = np.array([0.0, 1.3 / side - 0.5, 0.9 / side - 0.5])
    ev, eg = _multilinear(u)
    assert abs(v - ev) <= 1e-12 and g[0] == 0.0
    assert np.allclose(g[1:], eg[1:] / side, rtol=1e-12)
    x = np.array([1.1, 10.0, 0.9])        # u_y > dim - 1 -> dim - 1
    v, g = O.trilinear_d(grid, d, side, x)
    u = np.array([1.1 / side - 0.5, d[1] - 1.0, 0.9 / side - 0.5])
    ev, eg = _multilinear(u)
    assert abs(v - ev) <= 1e-12 and g[1] == 0.0
    assert np.allclose(g[[0, 2]], eg[[0, 2]] / side, rtol=1e-12)


def test_sensor_insert_closed_forms():
    d = (16, 8, 8)
    m = O.new_map(O.PROB_MAP, d)
    # one ray along +x from voxel (1, 2, 3) to voxel (6, 2, 3): six cells FREE (-10 from unknown),
    # then the end voxel OCCUPIED: -10 + 72 = 62
    O.prob_insert_sensor(m, d, 1.0, [1.5, 2.5, 3.5], np.array([[6.2, 2.9, 3.1]], np.float32))
    v = m.reshape(8, 8, 16)
    assert list(v[3, 2, 1:6]) == [-10] * 5 and v[3, 2, 6] == 62
    assert int((m != -128).sum()) == 6
    # without ray casting only the measured voxel changes; NaN and outside points are skipped
    m2 = O.new_map(O.PROB_MAP, d)
    O.prob_insert_sensor(m2, d, 1.0, [1.5, 2.5, 3.5],
                         np.array([[6.2, 2.9, 3.1], [np.nan, 1, 1], [20, 1, 1]], np.float32), raycast=False)
    assert int((m2 != -128).sum()) == 1 and m2.reshape(8, 8, 16)[3, 2, 6] == 72
    # the sensor outside the map: nothing happens
    m3 = O.new_map(O.PROB_MAP, d)
    O.prob_insert_sensor(m3, d, 1.0, [1.5, 20.0, 3.5], np.array([[6.2, 2.9, 3.1]], np.float32))
    assert int((m3 != -128).sum()) == 0
    # saturation: 20 rays through the same cells clamp them at -127; all FREE updates of a
    # scan come before its OCCUPIED ones, so the 20 hits then take the end voxel to 127
    m4 = O.new_map(O.PROB_MAP, d)
    O.prob_insert_sensor(m4, d, 1.0, [0.5, 0.5, 0.5], np.repeat(np.array([[3.5, 0.5, 0.5]], np.float32), 20, axis=0))
    assert list(m4[:4]) == [-127, -127, -127, 127]
    # a robot voxel carrying the robot bit removes the point (and its ray)
    robot = np.zeros((16 * 8 * 8, 32), np.uint8)
    robot[(3 * 8 + 2) * 16 + 6, 1] = 1 << 2      # bit 10
    m5 = O.new_map(O.PROB_MAP, d)
    O.prob_insert_sensor(m5, d, 1.0, [1.5, 2.5, 3.5], np.array([[6.2, 2.9, 3.1]], np.float32), robot=robot, robot_bit=10)
    assert int((m5 != -128).sum()) == 0
    O.prob_insert_sensor(m5, d, 1.0, [1.5, 2.5, 3.5], np.array([[6.2, 2.9, 3.1]], np.float32), robot=robot, robot_bit=11)
    assert np.array_equal(m5, m)


def test_sensor_insert_point_limit_and_update_rule():
    # the reference's loop stops at min(points, voxels); updateOccupancy starts unknown voxels at 0
    d = (2, 2, 2)
    m = O.new_map(O.PROB_MAP, d)
    pts = np.array([[0.5, 0.5, 0.5]] * 8 + [[1.5, 1.5, 1.5]] * 5, np.float32)
    O.prob_insert_sensor(m, d, 1.0, [0.5, 0.5, 0.5], pts, raycast=False)
    assert m[0] == 127 and m[7] == -128          # 8 hits saturate; the 5 later points are beyond the limit
    upd = O.lib().gvo_prob_update
    assert upd(-128, -10) == -10 and upd(-120, -10) == -127 and upd(100, 72) == 127


# ---------------------------------------------------------------- swept volumes (reference known answers)
def _bv(*bits):
    v = np.zeros(32, np.uint8)
    for b in bits:
        v[b >> 3] |= 1 << (b & 7)
    return v


def _bit(v, b):
    return bool((v[b >> 3] >> (b & 7)) & 1)


def test_perform_left_shift_known_answers():
    # testing_bitvector.cu:144-196 (bitvector_bitshift)
    for shift in (3, 13):
        v = O.perform_left_shift(_bv(0, 1, 2, 3, 4, 14, 24, 254, 255), shift)
        for b in (0, 1, 2, 3, 4, 14, 24, 254, 255):
            assert not _bit(v, b)
        assert _bit(v, 24 - shift) == (24 - shift > 3)
        assert _bit(v, 254 - shift) and _bit(v, 255 - shift)
    # closed form on random vectors: bits 4.. of the result = bits shifted down, bits 0..3 cleared;
    # the top `shift` bits are zero
    rng = np.random.default_rng(4)
    for _ in range(200):
        v = rng.integers(0, 256, 32, dtype=np.uint8)
        s = int(rng.integers(0, 57))
        bits = np.unpackbits(v, bitorder="little")
        want = np.zeros(256, np.uint8)
        want[:256 - s] = bits[s:]
        want[:4] = 0
        assert np.array_equal(np.unpackbits(O.perform_left_shift(v, s), bitorder="little"), want)
    assert not O.shift_sv(np.zeros((2, 32), np.uint8), 57)


def test_map_bitshift_collision_known_answer():
    # testing_voxelmap.cu:244-299 (bitvoxelmap_bitshift): SVCollider(1.0, window 1)
    pts = np.array([[0.3, 0.3, 0.3], [1.3, 1.3, 1.3], [2.3, 2.3, 2.3], [3.3, 3.3, 3.3]], np.float32)
    type_int = 4 + 50
    for shift in range(0, 50 + 4 - 5):
        m1 = O.new_map(O.BITVECTOR_MAP, DIMS)
        m2 = O.new_map(O.BITVECTOR_MAP, DIMS)
        O.map_insert_points(O.BITVECTOR_MAP, m2, DIMS, 1.0, pts, type_int)
        O.map_insert_points(O.BITVECTOR_MAP, m1, DIMS, 1.0, pts, type_int - shift)
        O.shift_sv(m2, shift)
        n, _ = O.map_collide_sv(O.BITVECTOR_MAP, m1, O.BITVECTOR_MAP, m2, DIMS, 1.0, window=1)
        assert n == (len(pts) if shift <= 50 else 0)


def test_list_bitshift_collision_known_answer():
    # testing_voxellist.cu:502-546 (bitvoxellist_bitshift_collision): collideWithBitcheck, window 1
    pts = np.array([[0.3, 0.3, 0.3]], np.float32)
    for shift in range(0, 49):
        l1, l2 = O.List(DIMS, 1.0), O.List(DIMS, 1.0)
        l2.insert_points(pts, 54)
        l1.insert_points(pts, 54 - shift)
        l2.shift_sv(shift)
        assert l1.collide_bitcheck(l2, 1) == 1
    # margin 0 compares the full vectors (bits 0..3 included)
    l1, l2 = O.List(DIMS, 1.0), O.List(DIMS, 1.0)
    l1.insert_points(pts, 1)
    l2.insert_points(pts, 1)
    assert l1.collide_bitcheck(l2, 0) == 1 and l1.collide_bitcheck(l2, 1) == 0


def test_counting_per_meaning_swept_volume_known_answer():
    # testing_voxellist.cu:46-101: 50 swept-volume boxes vs one obstacle box
    obstacle, swept = O.List(DIMS, 1.0), O.List(DIMS, 1.0)
    obstacle.insert_points(create_box_of_points((8.0, 8.0, 8.0), (11.0, 11.0, 11.0), 0.1), meaning=1)
    clouds, sizes = [], []
    for i in range(50):
        # (max - min) * swept_ratio_delta * i + min with a float ratio: float arithmetic throughout
        p = np.float32(np.float32(np.float32(30) * np.float32(0.02)) * np.float32(i)) + np.float32(0)
        c = create_box_of_points((p, p, p), (p + np.float32(2.5), p + np.float32(2.5), p + np.float32(2.5)), 0.1)
        clouds.append(c)
        sizes.append(len(c))
    swept.insert_points(np.concatenate(clouds), sizes=np.array(sizes, np.uint32),
                        meanings=np.arange(4, 54, dtype=np.uint32))
    total, counts = swept.collide_per_meaning(obstacle)
    got = {i: int(c) for i, c in enumerate(counts) if c}
    assert got == {14: 1, 15: 8, 16: 8, 17: 27, 18: 27, 19: 27, 20: 8, 21: 1, 22: 1}
    assert total == 108


def test_counting_per_meaning_find_matching_known_answer():
    # testing_voxellist.cu:343-414 (bitvoxellist_findMatchingVoxels)
    d, side = (500, 500, 200), 0.01
    l1, l2 = O.List(d, side), O.List(d, side)
    cp = [np.array([[np.float32(j / n * 5.0), y, 1.0] for j in range(n)], np.float32) for n, y in ((40, 1.0), (30, 2.0), (20, 3.0))]
    l1.insert_points(create_box_of_points((2.1, 0.1, 5.1), (4.9, 4.9, 9.9), 0.02), meaning=10)
    l2.insert_points(create_box_of_points((5.1, 5.1, 10.1), (9.9, 9.9, 14.9), 0.3), meaning=20)
    for k in range(3):
        l1.insert_points(cp[k], meaning=11 + k)
        l2.insert_points(cp[k], meaning=21 + k)
    t1, c1 = l1.collide_per_meaning(l2)
    assert t1 == 90 and (c1[11], c1[12], c1[13]) == (40, 30, 20) and c1[21] == c1[22] == c1[23] == 0
    assert c1[10] == 0 and c1[20] == 0
    t2, c2 = l2.collide_per_meaning(l1)
    assert t2 == 90 and (c2[21], c2[22], c2[23]) == (40, 30, 20) and c2[11] == c2[12] == c2[13] == 0


# ------------------------------------------------------------------ self-collision-checking insert
def _bv_map(dims):
    return O.new_map(O.BITVECTOR_MAP, dims)


@pytest.mark.parametrize("j", [2, 10])
def test_selfcollision_known_answers(j):
    # test/testing_voxelmap.cu:366-479 (selfcollision_metapointcloud), both parts
    dims, d = DIMS, 1.0 / j
    boxes_c = [((2.1,) * 3, (5.1,) * 3), ((4.1,) * 3, (7.1,) * 3), ((6.1,) * 3, (9.1,) * 3), ((10.1,) * 3, (12.1,) * 3)]
    boxes_n = [((30.1,) * 3, (32.1,) * 3), ((40.1,) * 3, (43.1,) * 3), ((60.1,) * 3, (64.1,) * 3),
               ((65.1,) * 3, (68.1,) * 3)]
    clouds_c = [create_box_of_points(a, b, d) for a, b in boxes_c]
    clouds_n = [create_box_of_points(a, b, d) for a, b in boxes_n]
    pc, sc = np.concatenate(clouds_c), [len(c) for c in clouds_c]
    pn, sn = np.concatenate(clouds_n), [len(c) for c in clouds_n]
    m = _bv_map(dims)
    coll, _, _ = O.map_insert_meta_selfcoll(m, dims, 1.0, pc, sc)
    assert coll is True
    coll, _, _ = O.map_insert_meta_selfcoll(m, dims, 1.0, pn, sn)
    assert coll is False
    # part 2: meanings 51..54; cloud 0 ignores cloud 1, cloud 1 ignores cloud 0, each ignores itself
    m = _bv_map(dims)
    meanings = np.array([51, 52, 53, 54], np.uint32)
    full = set_bits(*range(256)) & ~set_bits(2)
    masks = np.stack([full] * 4)
    masks[0] &= ~set_bits(52, 51)
    masks[1] &= ~set_bits(51, 52)
    masks[2] &= ~set_bits(53)
    masks[3] &= ~set_bits(54)
    coll, col, _ = O.map_insert_meta_selfcoll(m, dims, 1.0, pc, sc, meanings, masks)
    assert coll is True and bits_of(col) == [52, 53]
    masks = np.stack([full & ~set_bits(k) for k in (51, 52, 53, 54)])
    coll, col, _ = O.map_insert_meta_selfcoll(m, dims, 1.0, pn, sn, meanings, masks)
    assert coll is False and not col.any()


def test_selfcollision_against_voxel_sets():
    # independent model: with the default masks a sub-cloud collides exactly when one of its
    # voxels already holds an earlier sub-cloud; the colliding meanings are those of the earlier
    # sub-clouds it overlaps plus its own, and only the overlapped voxels get eBVM_COLLISION
    rng = np.random.default_rng(12)
    dims, side = (23, 17, 11), 0.1
    for trial in range(6):
        ncl = int(rng.integers(1, 7))
        sizes = rng.integers(0, 300, ncl)
        pts = rng.uniform(-0.2, 2.4, size=(int(sizes.sum()), 3)).astype(np.float32)
        m = _bv_map(dims)
        coll, col, outside = O.map_insert_meta_selfcoll(m, dims, side, pts, sizes)
        off, occupied, want_col, coll_vox = 0, {}, set(), set()
        q = np.maximum(np.floor(pts / np.float32(side)), 0)  # the saturating float -> unsigned cast
        inside = np.all(q < np.array(dims), axis=1)
        vid = (q[:, 2] * dims[1] + q[:, 1]) * dims[0] + q[:, 0]
        for c, n in enumerate(sizes):
            vs = set(int(v) for v in vid[off:off + n][inside[off:off + n]])
            hit = vs & set(occupied)
            if hit:
                for v_ in hit:
                    want_col |= occupied[v_]
                want_col.add(4 + c)
                coll_vox |= hit
            for v_ in vs:
                occupied.setdefault(v_, set()).add(4 + c)
            off += n
        assert coll == bool(want_col) and bits_of(col) == sorted(want_col)
        assert outside == bool((~inside).any())
        want = _bv_map(dims)
        for v_, ms in occupied.items():
            want[v_] = set_bits(*sorted(ms | ({2} if v_ in coll_vox else set())))
        assert np.array_equal(m, want)
