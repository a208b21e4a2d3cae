"""GPU parity of batched multi-pose collision vs the oracle and vs the per-pose GPU path."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _robot(rng, nl, sizes):
    dh = np.stack([np.zeros(nl), rng.uniform(0.05, 0.3, nl), rng.uniform(0.05, 0.3, nl),
                   rng.choice([-np.pi / 2, 0, np.pi / 2], nl), np.zeros(nl)], axis=1).astype(np.float32)
    pts = rng.uniform(-0.08, 0.35, (int(np.sum(sizes)), 3)).astype(np.float32)
    return dh, pts


@pytest.mark.parametrize("env_type", [O.BITVECTOR_MAP, O.PROB_MAP])
def test_batch_small_vs_oracle(gvx_lib, env_type):
    gvx = gvx_lib
    rng = np.random.default_rng(1 + env_type)
    d, side = (96, 80, 64), 0.02
    env_pts = rng.uniform(0, 1.3, (40000, 3)).astype(np.float32)
    oenv = O.new_map(env_type, d)
    O.map_insert_points(env_type, oenv, d, side, env_pts, 1)
    env = gvx.Map(env_type, d, side)
    env.write(oenv)
    thr = 1.0 if env_type == O.BITVECTOR_MAP else 0.7
    if env_type == O.BITVECTOR_MAP:
        env_bits = O.occupancy_bits(oenv)
    else:
        env_bits = np.packbits(oenv >= O.float_to_probability(thr), bitorder="little")
    sizes = np.array([700, 1300, 90, 2100, 513], np.uint32)
    dh, pts = _robot(rng, 5, sizes)
    base = np.array([1, 0, 0, 0.9, 0, 1, 0, 0.8, 0, 0, 1, 0.4], np.float32)
    q = rng.uniform(-np.pi, np.pi, (77, 5)).astype(np.float32)
    b = gvx.Batch(env, pts, sizes, dh, base, max_poses=128, threshold=thr)
    oc, om = O.batch_collide(env_bits, d, side, pts, sizes, dh, base, q)
    assert oc.sum() > 0
    for s in range(16):          # 0-7 direct launches, 8-15 graph replays
        b.set_schedule(s)
        for _ in range(2):
            c, m = b.collide(q)
            assert np.array_equal(c, oc), s
            assert np.array_equal(m, om), s
    robot = gvx.Map(O.BITVECTOR_MAP, d, side)
    assert np.array_equal(b.collide_per_pose(robot, q[:12]), oc[:12])


def test_batch_env_refresh(gvx_lib):
    gvx = gvx_lib
    rng = np.random.default_rng(5)
    d, side = (64, 64, 64), 0.02
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    sizes = np.array([500, 500], np.uint32)
    dh, pts = _robot(rng, 2, sizes)
    base = np.array([1, 0, 0, 0.6, 0, 1, 0, 0.6, 0, 0, 1, 0.6], np.float32)
    q = rng.uniform(-np.pi, np.pi, (40, 2)).astype(np.float32)
    b = gvx.Batch(env, pts, sizes, dh, base, max_poses=64)
    c, m = b.collide(q)
    assert c.sum() == 0 and not m.any()
    env_pts = rng.uniform(0, 1.28, (30000, 3)).astype(np.float32)
    env.insert_points(env_pts, 1)
    b.refresh_env()
    oenv = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_points(O.BITVECTOR_MAP, oenv, d, side, env_pts, 1)
    oc, om = O.batch_collide(O.occupancy_bits(oenv), d, side, pts, sizes, dh, base, q)
    c, m = b.collide(q)
    assert np.array_equal(c, oc) and np.array_equal(m, om)


def test_cfg4_full_size(gvx_lib):
    """1024 poses of the 7-link arm against the 512^3 cfg2 environment."""
    gvx = gvx_lib
    from gvx import scenes
    c2 = scenes.cfg2()
    c4 = scenes.cfg4()
    d, side = c2["dims"], c2["side"]
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.insert_points(c2["env_points"], 1)
    # occupancy bits for the oracle straight from the point cloud (same quantization)
    ids = scenes.points_to_voxel_ids(c2["env_points"], d, side)
    env_bits = np.zeros(512 ** 3 // 8, np.uint8)
    np.bitwise_or.at(env_bits, ids // 8, (1 << (ids % 8)).astype(np.uint8))
    arm = c4["arm"]
    b = gvx.Batch(env, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"], max_poses=1024)
    c, m = b.collide(c4["q"])
    sel = np.arange(0, 1024, 4)
    oc, om = O.batch_collide(env_bits, d, side, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"],
                             c4["q"][sel])
    assert np.array_equal(c[sel], oc)
    assert np.array_equal(m[sel], om)
    assert (c > 0).sum() > 10
    robot = gvx.Map(O.BITVECTOR_MAP, d, side)
    assert np.array_equal(b.collide_per_pose(robot, c4["q"][:8]), c[:8])


def test_batch_culling_at_map_faces(gvx_lib):
    """Links reaching past every face of the map: points below 0 quantize to voxel 0 and
    the ones past the far faces are skipped, so the per-chunk broad phase has to keep
    every pose whose box touches a face brick."""
    gvx = gvx_lib
    rng = np.random.default_rng(11)
    d, side = (64, 48, 40), 0.02
    ext = np.array(d, np.float32) * side
    face = []
    for a in range(3):
        for v in (0.5 * side, ext[a] - 0.5 * side):
            p = rng.uniform(0, 1, (1500, 3)).astype(np.float32) * ext
            p[:, a] = v
            face.append(p)
    env_pts = np.concatenate(face + [rng.uniform(0, 1, (3000, 3)).astype(np.float32) * ext])
    oenv = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_points(O.BITVECTOR_MAP, oenv, d, side, env_pts, 1)
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.write(oenv)
    sizes = np.array([900, 2500, 1700, 333], np.uint32)
    dh, pts = _robot(rng, 4, sizes)
    q = rng.uniform(-np.pi, np.pi, (300, 4)).astype(np.float32)
    for base_xyz in ((0.05, 0.05, 0.05), (1.2, 0.9, 0.75), (0.64, 0.48, 0.4)):
        base = np.array([1, 0, 0, base_xyz[0], 0, 1, 0, base_xyz[1], 0, 0, 1, base_xyz[2]], np.float32)
        b = gvx.Batch(env, pts, sizes, dh, base, max_poses=512)
        oc, om = O.batch_collide(O.occupancy_bits(oenv), d, side, pts, sizes, dh, base, q)
        assert (oc > 0).sum() > 20
        for s in (0, 3, 4, 6, 8, 12):
            b.set_schedule(s)
            c, m = b.collide(q)
            assert np.array_equal(c, oc), (base_xyz, s)
            assert np.array_equal(m, om), (base_xyz, s)


def test_batch_device_buffers_and_async(gvx_lib):
    """Device-resident q / counts / masks (direct launch) and the asynchronous entry
    point agree with the host path; the broad-phase statistics are consistent."""
    import torch
    gvx = gvx_lib
    rng = np.random.default_rng(21)
    d, side = (80, 64, 48), 0.02
    env_pts = rng.uniform(0, 1.0, (20000, 3)).astype(np.float32)
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.insert_points(env_pts, 1)
    sizes = np.array([640, 1000, 300], np.uint32)
    dh, pts = _robot(rng, 3, sizes)
    base = np.array([1, 0, 0, 0.7, 0, 1, 0, 0.6, 0, 0, 1, 0.45], np.float32)
    q = rng.uniform(-np.pi, np.pi, (200, 3)).astype(np.float32)
    b = gvx.Batch(env, pts, sizes, dh, base, max_poses=256)
    hc, hm = b.collide(q)
    assert (hc > 0).sum() > 10
    dev = torch.device("cuda:0")
    qd = torch.from_numpy(q).to(dev)
    cd = torch.full((200,), 7, dtype=torch.int32, device=dev)
    md = torch.full((200, 32), 7, dtype=torch.uint8, device=dev)
    b.collide(qd, cd, md)
    assert np.array_equal(cd.cpu().numpy().astype(np.uint32), hc)
    assert np.array_equal(md.cpu().numpy(), hm)
    tests, live, live_pp = b.last_stats()
    assert tests > 0 and 0 < live <= tests and live <= live_pp <= live * 256
    for sched in (0, 5, 8, 13):            # direct launches and graph replays on the caller's buffers
        b.set_schedule(sched)
        ca = torch.full((200,), 9, dtype=torch.int32, device=dev)
        ma = torch.full((200, 32), 9, dtype=torch.uint8, device=dev)
        for _ in range(3):                 # back-to-back calls on the library stream
            b.collide_async(qd, ca, ma)
        gvx.sync()
        assert np.array_equal(ca.cpu().numpy().astype(np.uint32), hc), sched
        assert np.array_equal(ma.cpu().numpy(), hm), sched
        cd.fill_(7)
        md.fill_(7)
        b.collide(qd, cd, md)
        assert np.array_equal(cd.cpu().numpy().astype(np.uint32), hc), sched
        assert np.array_equal(md.cpu().numpy(), hm), sched


def test_batch_many_links_and_tiny_links(gvx_lib):
    """20 links (two 16-link segments of the forward-kinematics prologue), empty and
    one-point links, a pose count that is not a multiple of 32."""
    gvx = gvx_lib
    rng = np.random.default_rng(31)
    d, side = (72, 72, 72), 0.02
    env_pts = rng.uniform(0, 1.44, (30000, 3)).astype(np.float32)
    oenv = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_points(O.BITVECTOR_MAP, oenv, d, side, env_pts, 1)
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.write(oenv)
    nl = 20
    sizes = rng.integers(0, 700, nl).astype(np.uint32)
    sizes[3], sizes[7], sizes[19] = 0, 1, 1
    dh = np.stack([np.zeros(nl), rng.uniform(0.0, 0.08, nl), rng.uniform(0.0, 0.08, nl),
                   rng.choice([-np.pi / 2, 0, np.pi / 2], nl), np.zeros(nl)], axis=1).astype(np.float32)
    pts = rng.uniform(-0.05, 0.1, (int(sizes.sum()), 3)).astype(np.float32)
    base = np.array([1, 0, 0, 0.7, 0, 1, 0, 0.7, 0, 0, 1, 0.5], np.float32)
    q = rng.uniform(-np.pi, np.pi, (97, nl)).astype(np.float32)
    b = gvx.Batch(env, pts, sizes, dh, base, max_poses=100)
    oc, om = O.batch_collide(O.occupancy_bits(oenv), d, side, pts, sizes, dh, base, q)
    assert (oc > 0).sum() > 5
    for s in range(16):
        b.set_schedule(s)
        c, m = b.collide(q)
        assert np.array_equal(c, oc), s
        assert np.array_equal(m, om), s


def test_batch_extreme_translations(gvx_lib):
    """Base translations of +-1e20 (every point outside the map on that axis; below zero
    the quantization saturates to voxel 0) and +-1e31 (row magnitude above the box
    test's 1e30 guard, so the pose is never culled): counts equal the oracle's."""
    gvx = gvx_lib
    rng = np.random.default_rng(41)
    d, side = (64, 64, 64), 0.02
    env_pts = np.concatenate([rng.uniform(0, 1.28, (20000, 3)),
                              np.array([[0.005, 0.005, 0.005], [0.005, 0.6, 0.6], [0.6, 0.005, 0.6]])]).astype(np.float32)
    oenv = O.new_map(O.BITVECTOR_MAP, d)
    O.map_insert_points(O.BITVECTOR_MAP, oenv, d, side, env_pts, 1)
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.write(oenv)
    sizes = np.array([300, 500], np.uint32)
    dh, pts = _robot(rng, 2, sizes)
    q = rng.uniform(-np.pi, np.pi, (64, 2)).astype(np.float32)
    bits = O.occupancy_bits(oenv)
    for t in ((-1e20, -1e20, -1e20), (-1e31, -1e31, -1e31), (1e20, 0.6, 0.6), (1e31, 0.6, 0.6),
              (-1e20, 0.6, 0.6), (0.6, -1e31, 0.6)):
        base = np.array([1, 0, 0, t[0], 0, 1, 0, t[1], 0, 0, 1, t[2]], np.float32)
        b = gvx.Batch(env, pts, sizes, dh, base, max_poses=64)
        oc, om = O.batch_collide(bits, d, side, pts, sizes, dh, base, q)
        c, m = b.collide(q)
        assert np.array_equal(c, oc), t
        assert np.array_equal(m, om), t
        b.close()


def test_batch_dense_environment(gvx_lib):
    """The cfg4 arm against a 512^3 BitVector environment with 10% of its voxels occupied (i.i.d.,
    seed 33: 12.8 M occupied voxels, a 1.6 GB hit matrix): calls of different sizes in a row, each
    against the oracle, so the words one call touched are cleared for the next."""
    import os
    gvx = gvx_lib
    from gvx import scenes
    c4 = scenes.cfg4(n_poses=96)
    d, side = (512, 512, 512), 0.01
    ids = scenes.cfg3_random(d, 0.1, 33)
    coords = np.stack([ids % 512, (ids // 512) % 512, ids // (512 * 512)], axis=1).astype(np.uint32)
    env = gvx.Map(O.BITVECTOR_MAP, d, side)
    env.insert_coords(coords, 1)
    bits = np.zeros(512 ** 3 // 8, np.uint8)
    np.bitwise_or.at(bits, ids // 8, (1 << (ids % 8)).astype(np.uint8))
    arm = c4["arm"]
    b = gvx.Batch(env, arm["link_points"], arm["link_sizes"], arm["dh"], c4["base"], max_poses=96)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    ref = {}
    for sched in (0, 13):          # a direct launch and a graph replay
        b.set_schedule(sched)
        for sl in (slice(0, 96), slice(10, 40), slice(0, 96)):
            q = c4["q"][sl]
            if (sl.start, sl.stop) not in ref:
                ref[sl.start, sl.stop] = O.batch_collide(bits, d, side, arm["link_points"], arm["link_sizes"],
                                                         arm["dh"], c4["base"], q, threads=threads)
            oc, om = ref[sl.start, sl.stop]
            assert oc.sum() > 0
            c, m = b.collide(q)
            assert np.array_equal(c, oc), (sched, sl)
            assert np.array_equal(m, om), (sched, sl)
