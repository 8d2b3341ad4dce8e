"""GPU: refit of the final layout (HipBVH::update) and the differentiable
closest-hit adjoint, checked against the oracle and FP64 central differences."""
import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import scenes
from hippie_amd.diff import DifferentiableTracer, backward

pytestmark = pytest.mark.gpu


def test_refit_matches_oracle(oracle, ctx):
    v, t = scenes.uniform_soup(30000, seed=61)
    rng = np.random.default_rng(62)
    v2 = (v + rng.normal(0, 0.005, v.shape)).astype(np.float32)
    scene = hx.Scene(v, t)
    bvh = hx.LBVHBuilder(ctx, morton_bits=60).build(scene)
    nodes0 = bvh.nodes_host()
    scene2 = hx.Scene(v2, t)
    bvh.update(scene2)
    nodes = bvh.nodes_host()
    tri = bvh.tri_indices_host()
    assert np.array_equal(nodes[:, 12:16], nodes0[:, 12:16]), "topology changed"
    ref = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, tri, v2, t)
    # min/max are exact: equal as floats (the binary-union order may differ
    # from the reference's fold only in the sign of a zero)
    assert np.array_equal(nodes.view(np.float32)[:, :12], ref.view(np.float32)[:, :12])
    assert np.array_equal(bvh.woop_host().view(np.uint32), oracle.woopify(v2, t, tri).view(np.uint32))
    assert oracle.validate_bin(nodes, bvh.n_interior, bvh.n_leaves, tri, v2, t) == 0


def test_refit_same_vertices_reproduces_build(oracle, ctx):
    v, t = scenes.config2_scene(n_mixture=20000, grid_cells=50, seed=63)
    scene = hx.Scene(v, t)
    bvh = hx.LBVHBuilder(ctx, morton_bits=60).build(scene)
    a = bvh.nodes_host()
    bvh.update()
    assert np.array_equal(a, bvh.nodes_host())


def _plane_t64(v, t, rays, ids):
    tri = t[ids]
    v0, v1, v2 = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    n = np.cross(v0 - v2, v1 - v2)
    return np.sum((v2 - o) * n, axis=1) / np.sum(d * n, axis=1)


def _fd_check(v, t, rays, ids, mask, t_tgt, grad, sample, h=1e-4):
    """FP64 central differences of L = sum_mask (t(V) - t_tgt)^2 with the hit
    triangles held fixed, on the sampled vertices; returns the scaled errors.

    Steps h and h/2 are combined by Richardson extrapolation, (4 D(h/2) -
    D(h)) / 3: on the config-4 mesh (edges ~0.0055) a plain h = 1e-4 central
    difference carries an O((h/edge)^2) ~ 3e-4 truncation error."""
    hit_rays = np.nonzero(mask)[0]
    tri_of = t[ids[hit_rays]]
    vd = v.astype(np.float64)
    g_fd_all, g_all = [], []

    def central(rows, k, c, step):
        vals = []
        for s in (+1, -1):
            vp = vd.copy()
            vp[k, c] += s * step
            tt = _plane_t64(vp, t, rays[rows], ids[rows])
            vals.append(np.sum((tt - t_tgt[rows]) ** 2))
        return (vals[0] - vals[1]) / (2 * step)

    for k in sample:
        rows = hit_rays[np.any(tri_of == k, axis=1)]
        for c in range(3):
            g_fd_all.append((4.0 * central(rows, k, c, h / 2) - central(rows, k, c, h)) / 3.0)
            g_all.append(grad[k, c])
    g_fd = np.array(g_fd_all)
    g = np.array(g_all, np.float64)
    scale = np.max(np.abs(g_fd))
    err = np.abs(g - g_fd) / (np.abs(g_fd) + 1e-3 * scale)
    return err, g, g_fd


@pytest.mark.parametrize("level,res,nsample", [(4, 96, 256), (8, 512, 1024)])
def test_backward_vs_fp64_finite_differences(ctx, level, res, nsample):
    """Config 4 at level 8: 1,310,720 triangles, 16 x 512^2 rays, 1,024 sampled vertices."""
    v, t = scenes.displaced_icosphere(level, seed=1, amplitude=0.05)
    vt, _ = scenes.displaced_icosphere(level, seed=2, amplitude=0.05)
    rays = scenes.orbit_rays(16, res, 3.0)
    dt = DifferentiableTracer(torch.from_numpy(v).cuda(), torch.from_numpy(t).cuda(), ctx)
    tgt_tr = DifferentiableTracer(torch.from_numpy(vt).cuda(), torch.from_numpy(t).cuda(), ctx)
    rays_t = torch.from_numpy(rays).cuda()
    with torch.no_grad():
        t_tgt, _, _, ids_tgt = tgt_tr(torch.from_numpy(vt).cuda(), rays_t)
    verts = torch.from_numpy(v).cuda().requires_grad_(True)
    tt, uu, vv, ids = dt(verts, rays_t)
    # Rays within ~14.5 degrees of grazing their triangle (|cos| < 0.25) are
    # left out: there t(V) is so curved on the small config-4 triangles that
    # the h = 1e-4 differences themselves disagree (Richardson at h and h/2
    # differ by 1.1e-4 at |cos| = 0.107), while the FP64 adjoint agrees with
    # h = 1e-6 differences to 5e-6.
    ids_c = ids.cpu().numpy()
    tri = t[np.maximum(ids_c, 0)]
    nrm = np.cross(v[tri[:, 0]] - v[tri[:, 2]], v[tri[:, 1]] - v[tri[:, 2]]).astype(np.float64)
    cosang = np.abs(np.sum(nrm * rays[:, 4:7], axis=1)) / np.maximum(np.linalg.norm(nrm, axis=1), 1e-30)
    mask = (ids >= 0) & (ids_tgt >= 0) & torch.from_numpy(cosang >= 0.25).cuda()
    loss = (((tt - t_tgt) ** 2) * mask).sum()
    loss.backward()
    grad = verts.grad.cpu().numpy()
    # direct C-ABI call gives the same gradient up to float atomics order
    res = torch.stack([ids, tt.view(torch.int32), uu.view(torch.int32), vv.view(torch.int32)], 1).contiguous()
    g2 = backward(ctx, verts.detach(), dt.triangles, rays_t, res, dl_dt=(2 * (tt - t_tgt) * mask).detach()).cpu()
    assert torch.allclose(torch.from_numpy(grad), g2, rtol=1e-5, atol=1e-6 * float(g2.abs().max()))
    ids_np = ids.cpu().numpy()
    m = mask.cpu().numpy()
    assert m.mean() > 0.2
    rng = np.random.default_rng(64)
    touched = np.unique(t[ids_np[m]])
    sample = rng.choice(touched, min(nsample, touched.size), replace=False)
    err, g, g_fd = _fd_check(v, t, rays, ids_np, m, t_tgt.cpu().numpy().astype(np.float64), grad, sample)
    # Contract: gradient within 1e-4 relative of the FP64 differences (vector
    # norm).  Per coordinate, the differences themselves resolve ~1e-4 on the
    # level-8 mesh, so single coordinates are held to 99.5% < 1e-4, all < 1e-3.
    normwise = np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd)
    assert normwise <= 1e-4, normwise
    assert np.quantile(err, 0.995) < 1e-4, np.quantile(err, 0.995)
    assert np.max(err) < 1e-3, (np.max(err), np.argmax(err))
    # and the gradient is not trivially zero
    assert np.max(np.abs(g_fd)) > 0


def test_backward_u_v_gradients(ctx):
    """dL/du and dL/dv paths: L = sum(u) + 2 sum(v) against FP64 differences."""
    v, t = scenes.displaced_icosphere(3, seed=5, amplitude=0.05)
    rays = scenes.orbit_rays(4, 64, 3.0)
    dt = DifferentiableTracer(torch.from_numpy(v).cuda(), torch.from_numpy(t).cuda(), ctx)
    verts = torch.from_numpy(v).cuda().requires_grad_(True)
    tt, uu, vv, ids = dt(verts, torch.from_numpy(rays).cuda())
    (uu.sum() + 2 * vv.sum()).backward()
    grad = verts.grad.cpu().numpy().astype(np.float64)
    ids_np = ids.cpu().numpy()
    hit = np.nonzero(ids_np >= 0)[0]

    def loss64(vd, rows):
        tri = t[ids_np[rows]]
        v0, v1, v2 = (vd[tri[:, k]] for k in range(3))
        o = rays[rows, 0:3].astype(np.float64)
        d = rays[rows, 4:7].astype(np.float64)
        e1, e2 = v0 - v2, v1 - v2
        M = np.stack([e1, e2, -d], axis=2)
        x = np.linalg.solve(M, (o - v2)[..., None])[..., 0]
        return np.sum(x[:, 0] + 2 * x[:, 1])

    vd = v.astype(np.float64)
    rng = np.random.default_rng(65)
    tri_of = t[ids_np[hit]]
    for k in rng.choice(np.unique(tri_of), 64, replace=False):
        rows = hit[np.any(tri_of == k, axis=1)]
        for c in range(3):
            vp, vm = vd.copy(), vd.copy()
            vp[k, c] += 1e-5
            vm[k, c] -= 1e-5
            g_fd = (loss64(vp, rows) - loss64(vm, rows)) / 2e-5
            assert abs(grad[k, c] - g_fd) <= 2e-3 * max(abs(g_fd), 1.0), (k, c, grad[k, c], g_fd)
