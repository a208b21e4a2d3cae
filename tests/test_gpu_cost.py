"""GPU differentiable clearance cost: forward and hand-written adjoint against closed forms, the FP64
oracle and FP64 central finite differences (h = 1e-4)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIDE = 0.01


@pytest.fixture(scope="module")
def edt_512(gvx_lib):
    """cfg5 grid: the cfg3(a) 512^3 EDT (cfg1 occupancy at the origin) as float metres."""
    import torch
    from gvx import scenes
    gvx = gvx_lib
    d = (512, 512, 512)
    ids = scenes.points_to_voxel_ids(scenes.cfg1()["env_points"], d, SIDE)
    coords = np.stack([ids % 512, (ids // 512) % 512, ids // (512 * 512)], axis=1).astype(np.uint32)
    m = gvx.Map(O.DIST_MAP, d, SIDE)
    m.insert_coords(coords, 1)
    m.pba()
    grid = torch.empty(512 ** 3, device="cuda", dtype=torch.float32)
    m.metric(grid)
    return grid, d


def _centres(dh, base, sph, q):
    """FP64 sphere centres (n_links, spheres, 3) of one waypoint."""
    fk = O.fk_chain_d(dh, base, q.astype(np.float64)).reshape(-1, 3, 4)
    return np.einsum("lrc,lsc->lsr", fk[:, :, :3], sph[:, :, :3].astype(np.float64)) + fk[:, None, :, 3]


def _trilinear_fields(g3, x, side):
    """FP64 trilinear value, Frobenius norm of the (mixed) second derivatives, and the cell index per
    axis, with the clamping of the sampled lattice (samples at voxel centres); g3 is (dz, dy, dx)."""
    dims = np.array(g3.shape[::-1])
    u = np.clip(x / side - 0.5, 0, dims - 1)
    i0 = np.minimum(np.floor(u).astype(np.int64), dims - 2)
    t = u - i0
    c = {}
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c[dx, dy, dz] = g3[i0[..., 2] + dz, i0[..., 1] + dy, i0[..., 0] + dx].astype(np.float64)
    tx, ty, tz = t[..., 0], t[..., 1], t[..., 2]
    val = 0.0
    for (dx, dy, dz), v in c.items():
        val = val + v * (tx if dx else 1 - tx) * (ty if dy else 1 - ty) * (tz if dz else 1 - tz)

    def mixed(a, b):  # d^2 / (da db) of the trilinear interpolant, axes a != b, the third interpolated
        o = 3 - a - b
        s = 0.0
        for k, v in c.items():
            sign = (1 if k[a] else -1) * (1 if k[b] else -1)
            s = s + sign * v * (t[..., o] if k[o] else 1 - t[..., o])
        return s / side ** 2

    hess = np.sqrt(2 * (mixed(0, 1) ** 2 + mixed(0, 2) ** 2 + mixed(1, 2) ** 2))
    return val, hess, i0


def _boundary_dist(x, side):
    u = x / side - 0.5
    return np.abs(u - np.rint(u))


def _multilinear_grid(d, scale):
    """g[i, j, k] = (1 + 2i - j + 3k + ij - 2jk + ik/2 + ijk/4) * scale, exact in float32 for a
    power-of-two scale; trilinear interpolation reproduces such a function exactly."""
    k, j, i = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
    g = (1 + 2 * i - j + 3 * k + i * j - 2 * j * k + 0.5 * i * k + 0.25 * i * j * k) * scale
    return g.astype(np.float32).ravel()


def _multilinear(u, scale):
    x, y, z = u[..., 0], u[..., 1], u[..., 2]
    v = (1 + 2 * x - y + 3 * z + x * y - 2 * y * z + 0.5 * x * z + 0.25 * x * y * z) * scale
    g = np.stack([2 + y + 0.5 * z + 0.25 * y * z, -1 + x - 2 * z + 0.25 * x * z,
                  3 - 2 * y + 0.5 * x + 0.25 * x * y], axis=-1) * scale
    return v, g


def test_cost_forward_backward_closed_form(gvx_lib):
    """A multilinear distance grid, which trilinear interpolation reproduces exactly, and one active
    sphere per trajectory (links 0 and 1 carry spheres of radius -10, never active): c = h^2 with
    h = r + eps - d(x), and dc/dx = -2 h grad d, so d and grad d at the sphere centre follow from the
    GPU's c and dc/dx and are compared with the closed form.  The sphere sits at
    base + Rz(q0) (rho + s, 0, q1) (a revolute joint, then a prismatic one along z)."""
    import torch
    gvx = gvx_lib
    d, side, scale = (16, 16, 16), 0.1, 2.0 ** -12
    grid = torch.from_numpy(_multilinear_grid(d, scale)).cuda()
    dh = np.array([[0, 0, 0, 0, 0],          # revolute: Rz(q0)
                   [0, 0, 0.3, 0, 1],        # prismatic: Tz(q1) Tx(0.3)
                   [0, 0, 0, 0, 0]], np.float32)
    sph = np.array([[[0, 0, 0, -10.0]], [[0, 0, 0, -10.0]], [[0.12, 0, 0, 0.9]]], np.float32)
    base = np.array([1, 0, 0, 0.8, 0, 1, 0, 0.8, 0, 0, 1, 0.2], np.float32)
    rng = np.random.default_rng(51)
    n = 512
    q = np.zeros((n, 1, 3), np.float32)
    q[:, 0, 0] = rng.uniform(-np.pi, np.pi, n)
    q[:, 0, 1] = rng.uniform(0.0, 1.1, n)
    cost = gvx.Cost(grid, d, side, dh, sph, 0.05, 1.0, base)
    qd = torch.from_numpy(q).cuda()
    out = torch.empty(n, device="cuda")
    dq = torch.empty_like(qd)
    dx = torch.empty((n, 1, 3, 1, 3), device="cuda")
    torch.cuda.synchronize()
    cost.forward(qd, out)
    cost.backward(qd, dq, dx)
    c, g = out.cpu().numpy().astype(np.float64), dx.cpu().numpy()[:, 0, 2, 0].astype(np.float64)
    ang, hgt = q[:, 0, 0].astype(np.float64), q[:, 0, 1].astype(np.float64)
    x = np.stack([0.8 + 0.42 * np.cos(ang), 0.8 + 0.42 * np.sin(ang), 0.2 + hgt], axis=1)
    u = x / side - 0.5
    assert np.all((u > 0) & (u < 15))
    ev, eg = _multilinear(u, scale)
    h = 0.95 - ev
    assert np.all(h > 0.5)
    # float32: sqrt(c) to ~1e-7 relative, the centre to ~1e-7 m (|grad d| < 0.1 here)
    assert np.max(np.abs((0.95 - np.sqrt(c)) - ev)) <= 2e-6
    grad = g / (-2.0 * h[:, None])
    assert np.allclose(grad, eg / side, rtol=1e-4, atol=1e-6)


def test_cost_forward_backward_vs_oracle(gvx_lib, edt_512):
    import torch
    from gvx import scenes
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=256)
    cost = gvx.Cost(grid, d, SIDE, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
    q = torch.from_numpy(c5["q"]).cuda()
    out = torch.empty(256, device="cuda")
    torch.cuda.synchronize()
    cost.forward(q, out)
    dq = torch.empty_like(q)
    dx = torch.empty((256, 32, 7, 64, 3), device="cuda")
    cost.backward(q, dq, dx)
    out, dq, dx = out.cpu().numpy(), dq.cpu().numpy(), dx.cpu().numpy()
    g = grid.cpu().numpy()
    g3 = g.reshape(d[2], d[1], d[0])
    sel = np.arange(0, 256, 16)
    base = np.asarray(c5["base"], np.float64)
    oc, odq, odx = O.cost_d(g, d, SIDE, c5["dh"], base, c5["spheres"], c5["q"][sel].astype(np.float64),
                            c5["eps"], c5["w"])
    assert (oc > 0).sum() >= 4
    # forward: float32 FK and interpolation vs FP64
    assert np.allclose(out[sel], oc, rtol=1e-4, atol=1e-7)
    # adjoint vs the oracle's analytic gradient, per waypoint, away from cell boundaries
    checked = 0
    for i, t in enumerate(sel):
        for w in range(32):
            ref = odq[i, w]
            nrm = np.linalg.norm(ref)
            if nrm < 1e-6:
                continue
            err = np.linalg.norm(dq[t, w] - ref) / nrm
            assert err <= 1e-4, (t, w, err)
            checked += 1
    assert checked >= 10
    # dc/dx = -2 w h grad d per active sphere: relative error of the 3-vector <= 1e-4.  Float32 gives the
    # sphere centre x to dx32 = 4 ulp(max |x_i|) (the 7-link FK chain and the transform round a few
    # times), so the hinge h = r + eps - d(x) is known to |grad d| dx32 and grad d to |H| dx32 (H: the
    # interpolant's second derivatives); a sphere is checked where that resolution is within the
    # bound, 2 w (|grad d|^2 + h |H|) dx32 <= 1e-4 |dc/dx| (near contact, h below ~5 mm, float32
    # cannot resolve dc/dx to 1e-4), and at least 1e-3 voxel from a cell boundary (where grad d
    # jumps).  Excluded spheres are counted and must stay few.
    bad = n_act = n_res = 0
    for i, t in enumerate(sel):
        for w in range(32):
            x = _centres(c5["dh"], base, c5["spheres"], c5["q"][t, w])
            ref = odx[i, w]
            nrm = np.linalg.norm(ref, axis=-1)
            act = nrm > 0
            dv, hess, _ = _trilinear_fields(g3, x, SIDE)
            hinge = c5["spheres"][:, :, 3] + c5["eps"] - dv
            unres, act, nrm = _unresolved_spheres(x, dv, hess, ref, hinge, c5["w"])
            ok = act & ~unres & (_boundary_dist(x, SIDE).min(axis=-1) > 1e-3)
            n_res += int(unres.sum())
            err = np.linalg.norm(dx[t, w] - ref, axis=-1)[ok]
            n_act += int(ok.sum())
            bad += int((err > 1e-4 * nrm[ok]).sum())
    print(f"dc/dx: {n_act} spheres checked, {n_res} below float32 resolution excluded")
    assert n_act > 10000 and n_res < 0.1 * n_act
    assert bad == 0, f"{bad} of {n_act} sphere gradients off by more than 1e-4"


def _unresolved_spheres(x, dv, hess, odx, hinge, w):
    """Active spheres whose dc/dx float32 cannot resolve to 1e-4: the float32 FK chain gives a sphere
    centre to dx32 = 4 ulp(max |x_i|), so the hinge h = r + eps - d(x) is known to |grad d| dx32 and
    grad d to |H| dx32 (H: the interpolant's second derivatives); dc/dx = -2 w h grad d is then known
    to 2 w (|grad d|^2 + h |H|) dx32, which must stay within 1e-4 |dc/dx|.  Near contact (h below
    ~5 mm) it does not."""
    nrm = np.linalg.norm(odx, axis=-1)
    act = nrm > 0
    gnorm = np.where(act, nrm / np.where(act, 2 * w * hinge, 1.0), 0.0)
    dx32 = 4 * np.spacing(np.abs(x).max(axis=-1).astype(np.float32)).astype(np.float64)
    resolved = 2 * w * (gnorm ** 2 + hinge * hess) * dx32 <= 1e-4 * nrm
    return act & ~resolved, act, nrm


def test_cost_gradient_vs_fp64_central_differences(gvx_lib, edt_512):
    """dc/dq per waypoint against FP64 central differences (h = 1e-4) of the oracle's cost, relative
    error <= 1e-4.  A waypoint is compared when its difference stencil is smooth: for every joint j
    and both q_j +- h, every sphere that is active anywhere on the stencil stays in the same
    interpolation cell and keeps its hinge sign (the cost is then a polynomial in the sphere centres
    along the stencil).  A waypoint whose dc/dq float32 cannot resolve to 1e-4 (the per-sphere
    resolution of the dc/dx test, carried to dc/dq through the joint lever arms) is excluded and
    counted: near contact, float32 sphere centres cannot resolve the hinge, whatever the adjoint."""
    import torch
    from gvx import scenes
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=64)
    cost = gvx.Cost(grid, d, SIDE, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
    q = torch.from_numpy(c5["q"]).cuda()
    dq = torch.empty_like(q)
    torch.cuda.synchronize()
    cost.backward(q, dq)
    dq = dq.cpu().numpy()
    g = grid.cpu().numpy()
    g3 = g.reshape(d[2], d[1], d[0])
    base = np.asarray(c5["base"], np.float64)
    sph = c5["spheres"]
    nl = sph.shape[0]
    h = 1e-4
    checked = worst = excluded = 0
    for t in range(64):
        for w in range(32):
            q0 = c5["q"][t, w].astype(np.float64)
            qs = [q0]
            for j in range(nl):
                for s in (1.0, -1.0):
                    qq = q0.copy()
                    qq[j] += s * h
                    qs.append(qq)
            X = np.stack([_centres(c5["dh"], base, sph, qq) for qq in qs])
            dv, hess, cell = _trilinear_fields(g3, X, SIDE)
            act = (sph[None, :, :, 3] + c5["eps"] - dv) > 0
            if not act[0].any():
                continue
            rel = act.any(axis=0)
            if not (np.all(np.all(cell == cell[:1], axis=(0, -1))[rel]) and np.all(act == act[:1])):
                continue
            _, _, odx = O.cost_d(g, d, SIDE, c5["dh"], base, sph, q0[None, None], c5["eps"], c5["w"])
            fd = np.zeros(nl)
            for j in range(nl):
                cp = O.cost_d(g, d, SIDE, c5["dh"], base, sph, qs[1 + 2 * j][None, None], c5["eps"], c5["w"], False)
                cm = O.cost_d(g, d, SIDE, c5["dh"], base, sph, qs[2 + 2 * j][None, None], c5["eps"], c5["w"], False)
                fd[j] = (cp[0][0] - cm[0][0]) / (2 * h)
            nrm = np.linalg.norm(fd)
            if nrm == 0.0:
                continue
            # float32 resolution of dc/dq: every active sphere's dc/dx is known to
            # e_s = 2 w (|grad d|^2 + h |H|) dx32 (_unresolved_spheres), which reaches dc/dq_j through
            # the lever |x_s - o_j| (o_j: origin of joint j's frame; links l > j); the spheres' rounding
            # errors are independent, so they add in quadrature
            _, act_s, dnrm = _unresolved_spheres(X[0], dv[0], hess[0], odx[0, 0], sph[:, :, 3] + c5["eps"] - dv[0],
                                                 c5["w"])
            hinge = sph[:, :, 3] + c5["eps"] - dv[0]
            gn = np.where(act_s, dnrm / np.where(act_s, 2 * c5["w"] * hinge, 1.0), 0.0)
            dx32 = 4 * np.spacing(np.abs(X[0]).max(axis=-1).astype(np.float32)).astype(np.float64)
            e_s = np.where(act_s, 2 * c5["w"] * (gn ** 2 + hinge * hess[0]) * dx32, 0.0)
            frames = O.fk_chain_d(c5["dh"], base, q0).reshape(-1, 3, 4)
            est = np.array([np.sqrt(sum(float(np.sum((np.linalg.norm(X[0][l] - frames[j][:, 3], axis=-1) * e_s[l]) ** 2))
                                        for l in range(j + 1, nl))) for j in range(nl)])
            if np.linalg.norm(est) > 1e-4 * nrm:
                excluded += 1
                continue
            err = np.linalg.norm(dq[t, w] - fd) / nrm
            worst = max(worst, err)
            assert err <= 1e-4, (t, w, err)
            checked += 1
    print(f"central differences: {checked} waypoints checked (worst relative error {worst:.2e}), "
          f"{excluded} below float32 resolution excluded")
    assert checked >= 400 and excluded <= checked / 4


def test_autograd_function(gvx_lib, edt_512):
    import torch
    from gvx import scenes
    from gvx.torch_cost import clearance_cost
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=32)
    cost = gvx.Cost(grid, d, SIDE, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
    q = torch.from_numpy(c5["q"]).cuda().requires_grad_(True)
    c = clearance_cost(q, cost)
    wts = torch.linspace(0.5, 2.0, 32, device="cuda")
    (c * wts).sum().backward()
    dq = torch.empty_like(q)
    torch.cuda.synchronize()
    cost.backward(q.detach(), dq, None, wts)
    assert torch.equal(q.grad, dq)
    ref = torch.empty(32, device="cuda")
    cost.forward(q.detach(), ref)
    assert torch.equal(c.detach(), ref)
    # shapes and placement are checked before any launch
    with pytest.raises(ValueError):
        cost.forward(q.detach()[:, :, :6].contiguous(), ref)          # 6 joint values for 7 links
    with pytest.raises(ValueError):
        cost.forward(q.detach().cpu(), ref)                            # host tensor
    with pytest.raises(ValueError):
        cost.backward(q.detach(), torch.empty(32, 32, 7, device="cuda"), torch.empty(32, 32, 7, 3, device="cuda"))
    with pytest.raises(ValueError):
        gvx.Cost(grid[:1000], d, SIDE, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
