"""Seeded synthetic inputs for the five benchmark workloads (numpy PCG64 streams).

cfg1  20 boxes in a 1.28 m cube, 100k noisy surface points -> 128^3 BitVector map;
      50k points on 5 spheres -> VoxelList with meanings 1..5.
cfg2  640x480 depth image of a 60-box room (pinhole fx=fy=525, cx=320, cy=240,
      depth 0.3..5 m, 1 % NaN dropouts) back-projected to ~300k points for a
      512^3 map at 1 cm; 7-link DH arm with capsule link clouds.
cfg3  obstacle fields for the EDT: the cfg1 occupancy embedded at the origin,
      or i.i.d. voxels at a given density.
cfg4  1024 poses of the 7-link arm (link clouds of 2k..40k points) against the
      cfg2 environment.
cfg5  trajectories of sphere-approximated links for the clearance cost.
"""
import numpy as np

SIDE = 0.01


def _boxes(rng, n, lo, hi, smin, smax):
    sides = rng.uniform(smin, smax, size=(n, 3))
    centres = rng.uniform(lo + sides / 2, hi - sides / 2)
    return centres - sides / 2, centres + sides / 2


def sample_box_surfaces(rng, bmin, bmax, n):
    """n points uniform on the union of box surfaces (area-weighted face choice)."""
    ext = bmax - bmin
    # faces: for axis a in 0..2, sides 0/1; area = product of the other two extents
    areas = []
    for a in range(3):
        o = [k for k in range(3) if k != a]
        ar = ext[:, o[0]] * ext[:, o[1]]
        areas += [ar, ar]
    areas = np.stack(areas, axis=1)  # (nbox, 6)
    flat = areas.ravel() / areas.sum()
    pick = rng.choice(flat.size, size=n, p=flat)
    box, face = pick // 6, pick % 6
    axis, side = face // 2, face % 2
    u = rng.random((n, 3))
    pts = bmin[box] + u * ext[box]
    pts[np.arange(n), axis] = np.where(side == 1, bmax[box, axis], bmin[box, axis])
    return pts


def cfg1(seed=1):
    rng = np.random.default_rng(seed)
    bmin, bmax = _boxes(rng, 20, 0.0, 1.28, 0.05, 0.30)
    env = sample_box_surfaces(rng, bmin, bmax, 100_000)
    env = env + rng.normal(0.0, 0.002, size=env.shape)
    keep = np.all((env >= 0.0) & (env < 1.28), axis=1)
    env = env[keep].astype(np.float32)
    radii = rng.uniform(0.05, 0.20, size=5)
    centres = rng.uniform(0.2, 1.08, size=(5, 3))
    per = np.full(5, 10_000)
    sph = []
    for k in range(5):
        v = rng.normal(size=(per[k], 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        sph.append(centres[k] + radii[k] * v)
    sph = np.concatenate(sph).astype(np.float32)
    return {"env_points": env, "list_points": sph, "list_sizes": per.astype(np.uint32),
            "list_meanings": np.arange(1, 6, dtype=np.uint32), "dims": (128, 128, 128), "side": SIDE}


# camera: world x = image right, world y = optical axis, world z = image up
CAM_POS = np.array([2.56, 0.05, 1.60])
CAM_R = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, -1.0, 0.0]])  # columns: cam x, y, z in world


def _room_boxes(rng):
    walls_min = np.array([[0.05, 0.05, 0.00], [0.05, 4.90, 0.00], [0.00, 0.05, 0.00], [5.00, 0.05, 0.00],
                          [0.05, 0.05, 3.60]])
    walls_max = np.array([[5.05, 5.05, 0.08], [5.05, 5.00, 3.70], [0.08, 5.05, 3.70], [5.08, 5.05, 3.70],
                          [5.05, 5.05, 3.68]])
    bmin, bmax = _boxes(rng, 55, np.array([0.3, 0.6, 0.08]), np.array([4.8, 4.8, 2.6]), 0.10, 0.90)
    return np.concatenate([walls_min, bmin]), np.concatenate([walls_max, bmax])


def render_depth(bmin, bmax, rng, w=640, h=480, fx=525.0, fy=525.0, cx=320.0, cy=240.0):
    """Ray-cast axis-aligned boxes; depth along the optical axis, NaN where nothing is hit."""
    u, v = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    dc = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], axis=-1).reshape(-1, 3)
    dw = dc @ CAM_R.T
    inv = 1.0 / np.where(np.abs(dw) < 1e-12, 1e-12, dw)
    best = np.full(dw.shape[0], np.inf)
    for k in range(bmin.shape[0]):
        t0 = (bmin[k] - CAM_POS) * inv
        t1 = (bmax[k] - CAM_POS) * inv
        tn = np.minimum(t0, t1).max(axis=1)
        tf = np.maximum(t0, t1).min(axis=1)
        hit = (tf >= tn) & (tf > 0)
        t = np.where(tn > 0, tn, tf)
        best = np.where(hit & (t < best), t, best)
    depth = best  # depth along optical axis equals t because dc.z == 1
    depth[(depth < 0.3) | (depth > 5.0)] = np.nan
    drop = rng.random(depth.shape[0]) < 0.01
    depth[drop] = np.nan
    return depth.reshape(h, w).astype(np.float32)


def backproject(depth, fx=525.0, fy=525.0, cx=320.0, cy=240.0):
    h, w = depth.shape
    u, v = np.meshgrid(np.arange(w, dtype=np.float32), np.arange(h, dtype=np.float32))
    ok = np.isfinite(depth)
    z = depth[ok]
    pc = np.stack([(u[ok] - cx) / fx * z, (v[ok] - cy) / fy * z, z], axis=1)
    pw = pc @ CAM_R.T.astype(np.float32) + CAM_POS.astype(np.float32)
    return pw.astype(np.float32)


def capsule_points(rng, n, p0, p1, r):
    """n points uniform on a capsule surface (cylinder + two hemispheres, area weighted)."""
    ax = p1 - p0
    L = np.linalg.norm(ax)
    ax_u = ax / L
    tmp = np.array([1.0, 0.0, 0.0]) if abs(ax_u[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = np.cross(ax_u, tmp)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(ax_u, e1)
    a_cyl, a_sph = 2 * np.pi * r * L, 4 * np.pi * r * r
    k_cyl = rng.binomial(n, a_cyl / (a_cyl + a_sph))
    phi = rng.uniform(0, 2 * np.pi, k_cyl)
    t = rng.uniform(0, 1, k_cyl)[:, None]
    cyl = p0 + t * ax + r * (np.cos(phi)[:, None] * e1 + np.sin(phi)[:, None] * e2)
    s = rng.normal(size=(n - k_cyl, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    along = s @ ax_u
    cen = np.where(along[:, None] >= 0, p1, p0)
    sph = cen + r * s
    pts = np.concatenate([cyl, sph])
    rng.shuffle(pts)
    return pts


def arm(seed=4, min_pts=2000, max_pts=40000, n_links=7):
    """7 revolute DH links (a, d ~ U[0.05, 0.40], alpha in {+-pi/2, 0}, theta0 = 0) and
    capsule link clouds in link frames."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.05, 0.40, n_links)
    d = rng.uniform(0.05, 0.40, n_links)
    alpha = rng.choice([-np.pi / 2, 0.0, np.pi / 2], n_links)
    dh = np.stack([np.zeros(n_links), d, a, alpha, np.zeros(n_links)], axis=1).astype(np.float32)
    sizes = rng.integers(min_pts, max_pts + 1, n_links).astype(np.uint32)
    clouds = []
    for l in range(n_links):
        # link body in its own frame: from the origin along z by d then x by a (the next joint)
        p1 = np.array([a[l], 0.0, d[l]])
        r = rng.uniform(0.03, 0.07)
        clouds.append(capsule_points(rng, int(sizes[l]), np.zeros(3), p1, r))
    pts = np.concatenate(clouds).astype(np.float32)
    return {"dh": dh, "link_points": pts, "link_sizes": sizes}


def translation(x, y, z):
    return np.array([1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z], dtype=np.float32)


ARM_BASE = translation(2.56, 2.20, 0.10)


def cfg2(seed=2, robot_seed=4):
    rng = np.random.default_rng(seed)
    bmin, bmax = _room_boxes(rng)
    depth = render_depth(bmin, bmax, rng)
    env = backproject(depth)
    r = arm(robot_seed, 2000, 20000)
    q = rng.uniform(-np.pi, np.pi, r["dh"].shape[0]).astype(np.float32)
    return {"depth": depth, "env_points": env, "dims": (512, 512, 512), "side": SIDE, "arm": r, "q": q,
            "base": ARM_BASE}


def cfg3_random(dims, density, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(np.array(dims, dtype=np.int64)))
    k = rng.binomial(n, density)
    idx = np.unique(rng.integers(0, n, size=k, dtype=np.int64))
    return idx


def cfg4(n_poses=1024, seed=4, pose_seed=44):
    r = arm(seed, 2000, 40000)
    rng = np.random.default_rng(pose_seed)
    q = rng.uniform(-np.pi, np.pi, size=(n_poses, r["dh"].shape[0])).astype(np.float32)
    return {"arm": r, "q": q, "base": ARM_BASE}


def cfg5(n_traj=256, n_wp=32, spheres_per_link=64, seed=5, robot_seed=4):
    r = arm(robot_seed, 2000, 2000)
    dh = r["dh"]
    rng = np.random.default_rng(seed)
    nl = dh.shape[0]
    sph = np.zeros((nl, spheres_per_link, 4), np.float32)
    for l in range(nl):
        p1 = np.array([dh[l, 2], 0.0, dh[l, 1]])
        t = np.linspace(0.0, 1.0, spheres_per_link)[:, None]
        sph[l, :, :3] = t * p1
        sph[l, :, 3] = rng.uniform(0.02, 0.08, spheres_per_link)
    q = rng.uniform(-np.pi, np.pi, size=(n_traj, n_wp, nl)).astype(np.float32)
    base = translation(0.64, 0.64, 0.05)
    return {"dh": dh, "spheres": sph, "q": q, "base": base, "eps": 0.05, "w": 1.0}


def points_to_voxel_ids(pts, dims, side):
    """Linear ids of in-bounds points (float32 division, floor) -- for building obstacle fields."""
    q = np.floor(pts.astype(np.float32) / np.float32(side))
    ok = np.all((q >= 0) & (q < np.array(dims)), axis=1)
    q = q[ok].astype(np.int64)
    return np.unique(q[:, 2] * dims[0] * dims[1] + q[:, 1] * dims[0] + q[:, 0])
