"""GPU parity of the DistanceVoxelMap EDT (PBA semantics), read-outs and exactDistances3D."""
import numpy as np
import pytest

from oracle import oracle as O
from refgen import create_equidistant_points_in_box

pytestmark = pytest.mark.gpu


def _obstacle_map(d, idx):
    u = O.dist_uninit(d)
    m = np.full(int(np.prod(np.array(d, dtype=np.int64))), u, np.uint32)
    idx = np.asarray(idx, np.int64)
    x, y, z = idx % d[0], (idx // d[0]) % d[1], idx // (d[0] * d[1])
    m[idx] = (x | (y << 10) | (z << 20)).astype(np.uint32)
    return m


def _gpu_pba(gvx, d, m):
    g = gvx.Map(O.DIST_MAP, d, 0.01)
    g.write(m)
    g.pba()
    return g


@pytest.mark.parametrize("seed", range(6))
def test_pba_small_grids_vs_bruteforce(gvx_lib, seed):
    rng = np.random.default_rng(seed)
    for _ in range(4):
        d = tuple(int(v) for v in rng.integers(1, 40, 3))
        n = int(np.prod(d))
        for dens in (0.0, 0.001, 0.02, 0.2, 1.0):
            idx = np.nonzero(rng.random(n) < dens)[0]
            m = _obstacle_map(d, idx)
            g = _gpu_pba(gvx_lib, d, m)
            assert np.array_equal(g.read(), O.edt_bruteforce(m, d)), (d, dens)


@pytest.mark.parametrize("dens,seed", [(0.001, 31), (0.01, 32), (0.1, 33)])
def test_pba_64_crops_vs_bruteforce(gvx_lib, dens, seed):
    d = (64, 64, 64) if dens < 0.05 else (40, 40, 40)
    idx = np.nonzero(np.random.default_rng(seed).random(int(np.prod(d))) < dens)[0]
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    assert np.array_equal(g.read(), O.edt_bruteforce(m, d))


def test_distance_correctness_reference_setup(gvx_lib):
    # test/testing_distance.cu:39-127: 128 x 128 x 131, equidistant obstacles, PBA vs exactDistances3D
    gvx = gvx_lib
    d = (128, 128, 131)
    obst = create_equidistant_points_in_box(10000, d, 1.0)
    p = gvx.Map(O.DIST_MAP, d, 1.0)
    p.insert_points(obst, 1)
    e = gvx.Map(O.DIST_MAP, d, 1.0)
    e.insert_points(obst, 1)
    p.pba()
    e.exact_distances(obst)
    sp, se = p.squared_distances(), e.squared_distances()
    assert np.array_equal(sp, se)
    om = O.new_map(O.DIST_MAP, d)
    O.map_insert_points(O.DIST_MAP, om, d, 1.0, obst, 1)
    O.pba(om, d)
    assert np.array_equal(p.read(), om)


def test_exact_distances_parity(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(7)
    d = (30, 20, 25)
    obst = rng.uniform(0, 1, (300, 3)).astype(np.float32) * np.array(d, np.float32)
    g = gvx.Map(O.DIST_MAP, d, 1.0)
    g.insert_points(obst[:100], 1)
    om = O.new_map(O.DIST_MAP, d)
    O.map_insert_points(O.DIST_MAP, om, d, 1.0, obst[:100], 1)
    g.exact_distances(obst)
    O.exact_distances(om, d, 1.0, obst)
    assert np.array_equal(g.read(), om)


def test_merge_occupied_and_readouts(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(8)
    d = (48, 40, 36)
    prob = O.new_map(O.PROB_MAP, d)
    prob[:] = rng.integers(-128, 128, prob.shape[0]).astype(np.int8)
    gp = gvx.Map(O.PROB_MAP, d, 0.02)
    gp.write(prob)
    gd = gvx.Map(O.DIST_MAP, d, 0.02)
    od = O.new_map(O.DIST_MAP, d)
    for thr in (0.9, 0.5):
        gd.merge_occupied(gp, (0, 0, 0), thr)
        O.dist_merge_occupied(od, prob, d, (0, 0, 0), thr)
    assert np.array_equal(gd.read(), od)
    gd.pba()
    O.pba(od, d)
    assert np.array_equal(gd.read(), od)
    idx = rng.integers(0, od.shape[0], 5000).astype(np.uint32)
    assert np.array_equal(gd.squared_distances(idx), O.dist_squared(od, d, idx))
    full = O.dist_squared(od, d)
    assert np.array_equal(gd.squared_distances(), full)
    met = gd.metric()
    assert np.array_equal(met, (np.sqrt(full.astype(np.float32)) * np.float32(0.02)).astype(np.float32))


def _cfg3a_ids(d):
    from gvx import scenes
    c = scenes.cfg1()
    return scenes.points_to_voxel_ids(c["env_points"], d, 0.01)


@pytest.mark.parametrize("which", ["cfg1", 0.001, 0.01, 0.1])
def test_pba_512_full_grid_vs_oracle(gvx_lib, which):
    d = (512, 512, 512)
    if which == "cfg1":
        idx = _cfg3a_ids(d)
    else:
        from gvx import scenes
        idx = scenes.cfg3_random(d, which, {0.001: 31, 0.01: 32, 0.1: 33}[which])
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    O.pba(m, d)
    assert np.array_equal(g.read(), m)


@pytest.mark.parametrize("d,dens,seed", [((200, 130, 70), 0.01, 41), ((201, 130, 70), 0.1, 42),
                                          ((70, 333, 129), 0.03, 43), ((1000, 67, 33), 0.3, 44),
                                          ((96, 1024, 40), 0.02, 45)])
def test_pba_ragged_dense_vs_oracle(gvx_lib, d, dens, seed):
    """Dense fields on dims that are not multiples of 64: partial x ranges and row blocks in the
    windowed y and x passes, halos clipped at both ends, a 1024-long y axis."""
    n = int(np.prod(np.array(d, dtype=np.int64)))
    idx = np.nonzero(np.random.default_rng(seed).random(n) < dens)[0]
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    O.pba(m, d)
    assert np.array_equal(g.read(), m), d


def _ypass_best(occ, zc, reach=40):
    """Per (y, x) of slice zc, the y pass's minimum over rows y' of (y - y')^2 + dz(y')^2 (dz: the
    nearest obstacle of column (x, y') in z), candidates up to `reach` rows away."""
    big = 1 << 20
    zz = np.arange(occ.shape[0])[:, None, None]
    dz = np.where(occ, np.abs(zz - zc), big).min(axis=0).astype(np.int64)
    c = np.minimum(dz * dz, big)
    best = c.copy()
    for k in range(1, reach + 1):
        best[k:] = np.minimum(best[k:], c[:-k] + k * k)
        best[:-k] = np.minimum(best[:-k], c[k:] + k * k)
    return best


def test_pba_cavity_mixed_window_and_stack_vs_oracle(gvx_lib):
    """A 2 % field with an empty box (x, y in [64, 192), z in [30, 110)): in the slices through
    the box's middle some voxel's nearest site in its (y, z) plane lies beyond the largest window
    radius, so those slices go to the stack kernel (packed sites) while the slices outside the box
    are resolved by the window (relative sites); the x pass reads both formats in one call, and
    the box's rows leave their x groups to the x stack kernels.  Every voxel against the oracle."""
    d = (256, 256, 160)
    rng = np.random.default_rng(47)
    occ = rng.random((d[2], d[1], d[0])) < 0.02
    occ[30:110, 64:192, 64:192] = False
    # the construction does what it says: a middle slice has voxels no window of radius 32
    # resolves (minimum at or above 33^2), an outer slice has none
    assert (_ypass_best(occ, 70) >= 33 * 33).any()
    assert (_ypass_best(occ, 5) < 33 * 33).all()
    idx = np.nonzero(occ.reshape(-1))[0]
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    O.pba(m, d)
    assert np.array_equal(g.read(), m)


@pytest.mark.parametrize("d,x_empty", [((256, 128, 64), (128, 192)), ((250, 100, 48), (64, 128))])
def test_pba_empty_x_range_vs_oracle(gvx_lib, d, x_empty):
    """A 5 % field with one 64-column x range holding no obstacle in any slice: the y window
    resolves every other range (relative sites) and skips the empty one, so no x tile mask of
    any slice is complete and every row goes to the x stack kernels, which read the relative
    sites of the windowed ranges.  Every voxel against the oracle."""
    rng = np.random.default_rng(48)
    occ = rng.random((d[2], d[1], d[0])) < 0.05
    occ[:, :, x_empty[0]:x_empty[1]] = False
    idx = np.nonzero(occ.reshape(-1))[0]
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    O.pba(m, d)
    assert np.array_equal(g.read(), m), d


@pytest.mark.parametrize("d", [(256, 128, 96), (200, 100, 70)])
def test_pba_mixed_density_x_ranges_vs_oracle(gvx_lib, d):
    """x below 128 at 3 % density, above at 0.05 % (below the window's 1/256 limit): the y
    window resolves the dense 64-column ranges (relative sites) and leaves the sparse ranges'
    groups to the y stack kernel (packed sites), so every slice holds both formats and its rows
    go to the x stack kernels, which read each tile in its own format.  Every voxel against
    the oracle."""
    rng = np.random.default_rng(49)
    u = rng.random((d[2], d[1], d[0]))
    occ = np.where(np.arange(d[0])[None, None, :] < 128, u < 0.03, u < 0.0005)
    idx = np.nonzero(occ.reshape(-1))[0]
    m = _obstacle_map(d, idx)
    g = _gpu_pba(gvx_lib, d, m)
    O.pba(m, d)
    assert np.array_equal(g.read(), m), d


@pytest.mark.parametrize("which", ["cfg1", 0.001, 0.01, 0.1])
def test_pba_1024_full_grid_vs_oracle(gvx_lib, which):
    """cfg3 at 1024^3 (the 0xFFFFFFFF sentinel): every voxel of the GPU grid against the oracle's
    PBA on the same field, run over this host's cores."""
    import os
    from gvx import scenes
    d = (1024, 1024, 1024)
    if which == "cfg1":
        idx = _cfg3a_ids(d)
    else:
        idx = scenes.cfg3_random(d, which, {0.001: 31, 0.01: 32, 0.1: 33}[which])
    m = _obstacle_map(d, idx)
    assert O.dist_uninit(d) == 0xFFFFFFFF
    g = _gpu_pba(gvx_lib, d, m)
    got = g.read()
    del g
    O.pba(m, d, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    assert np.array_equal(got, m), which
    # obstacles keep themselves; a probe of the read-out path on the same grid
    assert np.array_equal(got[idx], m[idx])


def test_pba_1024_large_plane_costs(gvx_lib):
    """x-pass costs above 2^20 (in-plane distances beyond 1024): planes x = 0 mod 16 hold one
    obstacle at (x, 0, 0), planes x = 8 mod 16 one at (x, 23, 46).  For the row (y, z) = (1023, 46)
    the first cost 1023^2 + 46^2 = 1048645 exceeds 2^20 and the second (1000^2) does not, so a
    window search that let the first wrap would pick it; every voxel is compared with the oracle."""
    import os
    d = (1024, 1024, 1024)
    xs = np.arange(1024)
    a, b = xs[xs % 16 == 0], xs[xs % 16 == 8]
    ids = np.concatenate([a, b + 23 * 1024 + 46 * 1024 * 1024]).astype(np.int64)
    m = _obstacle_map(d, ids)
    g = _gpu_pba(gvx_lib, d, m)
    got = g.read()
    del g
    O.pba(m, d, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    row = (46 * 1024 + 1023) * 1024
    assert np.array_equal(got[row:row + 1024], m[row:row + 1024])
    assert np.array_equal(got, m)


def test_pba_sparse_after_dense_reuses_scratch(gvx_lib):
    """A dense field first leaves valid-looking z-pass values in every column of the
    shared scratch; a sparse field afterwards stores only its obstacle segments and
    must ignore the stale ones (widths that are and are not multiples of 64)."""
    rng = np.random.default_rng(17)
    for d in ((256, 192, 160), (200, 96, 70)):
        n = d[0] * d[1] * d[2]
        dense = np.unique(rng.integers(0, n, n // 50)).astype(np.int64)
        m = _obstacle_map(d, dense)
        g = _gpu_pba(gvx_lib, d, m)
        O.pba(m, d)
        assert np.array_equal(g.read(), m), d
        # sparse: a few obstacles in one corner and one far away
        pts = np.array([[3, 4, 5], [10, 2, 7], [d[0] - 2, d[1] - 3, d[2] - 1], [70, 40, 30]], np.int64)
        ids = np.unique(pts[:, 2] * d[0] * d[1] + pts[:, 1] * d[0] + pts[:, 0])
        m2 = _obstacle_map(d, ids)
        g2 = _gpu_pba(gvx_lib, d, m2)
        O.pba(m2, d)
        assert np.array_equal(g2.read(), m2), d
