"""ProbVoxelMap::insertSensorData on the GPU (ray casting, robot cut) against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _scan(rng, n, dims, side, sensor):
    hi = np.array(dims) * side
    p = rng.uniform(-0.05 * hi, 1.05 * hi, size=(n, 3)).astype(np.float32)
    p[rng.random(n) < 0.01] = np.nan
    # some points on top of each other and some on the sensor voxel itself
    p[: n // 20] = p[n // 20: 2 * (n // 20)]
    p[-3:] = np.asarray(sensor, np.float32)
    return p


@pytest.mark.parametrize("seed,raycast", [(0, True), (1, True), (2, False), (3, True)])
def test_sensor_scan_parity(gvx_lib, seed, raycast):
    gvx = gvx_lib
    rng = np.random.default_rng(seed)
    dims, side = (61, 47, 33), 0.02
    sensor = rng.uniform(0.1, 0.6, 3).astype(np.float32)
    g = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
    o = O.new_map(O.PROB_MAP, dims)
    start = rng.integers(-128, 128, o.size).astype(np.int8)   # any prior occupancy
    start[rng.random(o.size) < 0.5] = -128
    g.write(start)
    o[:] = start
    for k in range(3):                                        # successive scans accumulate
        pts = _scan(rng, 4000 + 3000 * k, dims, side, sensor)
        g.insert_sensor_data(pts, sensor, raycast=raycast)
        O.prob_insert_sensor(o, dims, side, sensor, pts, raycast=raycast)
        got = g.read()
        assert np.array_equal(got, o), f"scan {k}: {int((got != o).sum())} voxels differ"


def test_sensor_robot_cut_and_edges(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(9)
    dims, side = (40, 40, 40), 0.05
    sensor = np.array([1.0, 1.0, 1.0], np.float32)
    pts = _scan(rng, 20000, dims, side, sensor)
    robot_bits = np.zeros((40 ** 3, 32), np.uint8)
    hit = rng.random(40 ** 3) < 0.3
    robot_bits[hit, 1] |= 1 << 3                             # meaning 11
    robot = gvx.Map(gvx.BITVECTOR_VOXELMAP, dims, side)
    robot.write(robot_bits)
    for bit in (11, 12):
        g = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
        o = O.new_map(O.PROB_MAP, dims)
        g.insert_sensor_data(pts, sensor, robot=robot, robot_meaning=bit)
        O.prob_insert_sensor(o, dims, side, sensor, pts, robot=robot_bits, robot_bit=bit)
        assert np.array_equal(g.read(), o)
    # sensor beyond the map: no voxel changes; a negative coordinate saturates to voxel 0
    # (the float -> unsigned conversion of mapToVoxels), so that sensor is inside
    g = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
    g.insert_sensor_data(pts, [3.0, 1.0, 1.0])
    assert np.all(g.read() == -128)
    g.insert_sensor_data(pts, [-1.0, 1.0, 1.0])
    o = O.new_map(O.PROB_MAP, dims)
    O.prob_insert_sensor(o, dims, side, [-1.0, 1.0, 1.0], pts)
    assert np.array_equal(g.read(), o) and int((o != -128).sum()) > 0
    # more points than voxels: only the first (voxel count) points are used
    d2 = (2, 2, 2)
    g2 = gvx.Map(gvx.PROBAB_VOXELMAP, d2, 1.0)
    o2 = O.new_map(O.PROB_MAP, d2)
    many = np.array([[0.5, 0.5, 0.5]] * 8 + [[1.5, 1.5, 1.5]] * 5, np.float32)
    g2.insert_sensor_data(many, [0.5, 0.5, 0.5])
    O.prob_insert_sensor(o2, d2, 1.0, [0.5, 0.5, 0.5], many)
    assert np.array_equal(g2.read(), o2) and o2[7] == -128
    with pytest.raises(gvx.GvxError):
        gvx.Map(gvx.BITVECTOR_VOXELMAP, d2, 1.0).insert_sensor_data(many, [0.5, 0.5, 0.5])
    # the robot map is read 32 bytes per voxel: a map of another type or size is refused on the host
    for bad in (gvx.Map(gvx.PROBAB_VOXELMAP, dims, side), gvx.Map(gvx.BITVECTOR_VOXELMAP, (20, 40, 40), side)):
        with pytest.raises(ValueError):
            gvx.Map(gvx.PROBAB_VOXELMAP, dims, side).insert_sensor_data(pts, sensor, robot=bad, robot_meaning=11)


def test_sensor_depth_image_512(gvx_lib):
    # cfg2's 640x480 depth image (~304k points) from its camera into a 512^3 ProbVoxelMap
    gvx = gvx_lib
    from gvx import scenes
    c2 = scenes.cfg2()
    dims, side = c2["dims"], c2["side"]
    pts = c2["env_points"]
    cam = scenes.CAM_POS.astype(np.float32)
    g = gvx.Map(gvx.PROBAB_VOXELMAP, dims, side)
    g.insert_sensor_data(pts, cam)
    o = O.new_map(O.PROB_MAP, dims)
    O.prob_insert_sensor(o, dims, side, cam, pts)
    got = g.read()
    assert np.array_equal(got, o), f"{int((got != o).sum())} voxels differ"
    assert int((o == 127).sum()) > 1000 and int((o == -127).sum()) > 100000
