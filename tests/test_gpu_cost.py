"""GPU differentiable clearance cost: forward vs FP64 oracle, hand-written adjoint vs the
oracle's analytic gradient and FP64 central finite differences (h = 1e-4)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def edt_512(gvx_lib):
    """cfg5 grid: the cfg3(a) 512^3 EDT (cfg1 occupancy at the origin) as float metres."""
    import torch
    from gvx import scenes
    gvx = gvx_lib
    d = (512, 512, 512)
    ids = scenes.points_to_voxel_ids(scenes.cfg1()["env_points"], d, 0.01)
    coords = np.stack([ids % 512, (ids // 512) % 512, ids // (512 * 512)], axis=1).astype(np.uint32)
    m = gvx.Map(O.DIST_MAP, d, 0.01)
    m.insert_coords(coords, 1)
    m.pba()
    grid = torch.empty(512 ** 3, device="cuda", dtype=torch.float32)
    m.metric(grid)
    return grid, d


def _boundary_dist(x, side):
    u = x / side - 0.5
    return np.abs(u - np.rint(u))


def test_cost_forward_backward_vs_oracle(gvx_lib, edt_512):
    import torch
    from gvx import scenes
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=256)
    cost = gvx.Cost(grid.data_ptr(), d, 0.01, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
    q = torch.from_numpy(c5["q"]).cuda()
    out = torch.empty(256, device="cuda")
    torch.cuda.synchronize()
    cost.forward(q, out)
    dq = torch.empty_like(q)
    dx = torch.empty((256, 32, 7, 64, 3), device="cuda")
    cost.backward(q, dq, dx)
    out, dq, dx = out.cpu().numpy(), dq.cpu().numpy(), dx.cpu().numpy()
    g = grid.cpu().numpy()
    sel = np.arange(0, 256, 16)
    base = np.asarray(c5["base"], np.float64)
    oc, odq, odx = O.cost_d(g, d, 0.01, c5["dh"], base, c5["spheres"], c5["q"][sel].astype(np.float64),
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
    # dc/dx per active sphere, relative error of the 3-vector, spheres away from cell boundaries
    bad = 0
    n_act = 0
    for i, t in enumerate(sel):
        for w in range(32):
            fk = O.fk_chain_d(c5["dh"], base, c5["q"][t, w].astype(np.float64)).reshape(7, 3, 4)
            x = np.einsum("lrc,lsc->lsr", fk[:, :, :3], c5["spheres"][:, :, :3].astype(np.float64)) + fk[:, None, :, 3]
            ref = odx[i, w]
            nrm = np.linalg.norm(ref, axis=-1)
            ok = (nrm > 0) & (_boundary_dist(x, 0.01).min(axis=-1) > 1e-3)
            # float32 resolves the hinge h = r + eps - d to ~1e-7 m, so dc/dx = -2 w h grad d carries an
            # absolute error of ~2 w 2e-7 |grad d| (|grad d| <= ~2): allowed on top of the 1e-4 relative bound
            err = np.linalg.norm(dx[t, w] - ref, axis=-1)[ok]
            n_act += int(ok.sum())
            bad += int((err > 1e-4 * nrm[ok] + 1e-6).sum())
    assert n_act > 100
    assert bad == 0, f"{bad} of {n_act} sphere gradients off by more than 1e-4"


def test_cost_gradient_vs_fp64_central_differences(gvx_lib, edt_512):
    import torch
    from gvx import scenes
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=64)
    cost = gvx.Cost(grid.data_ptr(), d, 0.01, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
    q = torch.from_numpy(c5["q"]).cuda()
    dq = torch.empty_like(q)
    torch.cuda.synchronize()
    cost.backward(q, dq)
    dq = dq.cpu().numpy()
    g = grid.cpu().numpy()
    base = np.asarray(c5["base"], np.float64)
    h = 1e-4
    checked = 0
    for t in range(64):
        q64 = c5["q"][t:t + 1].astype(np.float64)
        c0, _, dx0 = O.cost_d(g, d, 0.01, c5["dh"], base, c5["spheres"], q64, c5["eps"], c5["w"])
        if c0[0] == 0:
            continue
        for w in range(32):
            # skip waypoints whose active spheres sit within the FD stencil of a cell boundary
            fk = O.fk_chain_d(c5["dh"], base, q64[0, w]).reshape(7, 3, 4)
            active = np.abs(dx0[0, w]).sum(axis=-1) > 0
            if not active.any():
                continue
            x = np.einsum("lrc,lsc->lsr", fk[:, :, :3], c5["spheres"][:, :, :3].astype(np.float64))
            x += fk[:, None, :, 3]
            if _boundary_dist(x[active], 0.01).min() < 0.05:
                continue
            fd = np.zeros(7)
            for j in range(7):
                qp, qm = q64.copy(), q64.copy()
                qp[0, w, j] += h
                qm[0, w, j] -= h
                fd[j] = (O.cost_d(g, d, 0.01, c5["dh"], base, c5["spheres"], qp, c5["eps"], c5["w"], False)[0][0] -
                         O.cost_d(g, d, 0.01, c5["dh"], base, c5["spheres"], qm, c5["eps"], c5["w"], False)[0][0]) / (2 * h)
            nrm = np.linalg.norm(fd)
            if nrm < 1e-5:
                continue
            err = np.linalg.norm(dq[t, w] - fd) / nrm
            assert err <= 1e-4, (t, w, err)
            checked += 1
    assert checked >= 5


def test_autograd_function(gvx_lib, edt_512):
    import torch
    from gvx import scenes
    from gvx.torch_cost import clearance_cost
    gvx = gvx_lib
    grid, d = edt_512
    c5 = scenes.cfg5(n_traj=32)
    cost = gvx.Cost(grid.data_ptr(), d, 0.01, c5["dh"], c5["spheres"], c5["eps"], c5["w"], c5["base"])
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
