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


@pytest.mark.parametrize("method", ["lbvh", "ploc"])
def test_repeated_refits_and_rebuild(oracle, ctx, method):
    """The first refit after a build runs the full emit and records where each
    node's box rows come from; later refits rewrite only those rows.  Every
    refit must equal a first refit of a fresh build, and a rebuild (new
    topology) must drop the record."""
    cls = hx.LBVHBuilder if method == "lbvh" else hx.PLOCBuilder
    v, t = scenes.gaussian_mixture_soup(30000, seed=64)
    rng = np.random.default_rng(65)
    vs = [(v + rng.normal(0, 0.004 * (k + 1), v.shape)).astype(np.float32) for k in range(3)]
    builder = cls(ctx, morton_bits=60)
    bvh = builder.build(hx.Scene(v, t))
    nodes0, tri = bvh.nodes_host(), bvh.tri_indices_host()
    for vk in vs + [vs[0]]:
        bvh.update(hx.Scene(vk, t))
        fresh = cls(ctx, morton_bits=60).build(hx.Scene(v, t))
        fresh.update(hx.Scene(vk, t))
        assert np.array_equal(bvh.nodes_host(), fresh.nodes_host())
        ref = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, tri, vk, t)
        assert np.array_equal(bvh.nodes_host().view(np.float32)[:, :12], ref.view(np.float32)[:, :12])
    builder.rebuild(bvh, hx.Scene(vs[2], t))
    bvh.update(hx.Scene(vs[1], t))
    fresh = cls(ctx, morton_bits=60).build(hx.Scene(vs[2], t))
    fresh.update(hx.Scene(vs[1], t))
    assert np.array_equal(bvh.nodes_host(), fresh.nodes_host())


def _plane_t64(v, t, rays, ids):
    tri = t[ids]
    v0, v1, v2 = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    n = np.cross(v0 - v2, v1 - v2)
    return np.sum((v2 - o) * n, axis=1) / np.sum(d * n, axis=1)


def _hit_geometry(v, t, rays, ids, chunk=1 << 20):
    """FP64 plane hit of every ray on its triangle: t, barycentrics (b0, b1, b2)
    of the hit point, n / (d.n) (so dt/dv_i = b_i n / (d.n)), and whether the
    ray keeps its triangle under a move of any one vertex coordinate by +-h.

    A move by h shifts the triangles' points by at most h, so the hit point
    slides by at most h / |cos| in the plane (cos between ray and normal) and
    an edge by at most h: a ray whose hit lies farther than 4 h (1 + 1/|cos|)
    from every edge of its triangle keeps that triangle (the factor 4 covers
    the tilt of the moved planes).  Other rays are the ones whose primitive id
    may change, where the loss is not differentiable."""
    R = ids.shape[0]
    tt = np.zeros(R)
    B = np.zeros((R, 3))
    ndn = np.zeros((R, 3))
    margin_h = np.zeros(R)
    vd = v.astype(np.float64)
    for a in range(0, R, chunk):
        sl = slice(a, min(R, a + chunk))
        tri = t[np.maximum(ids[sl], 0)]
        v0, v1, v2 = (vd[tri[:, k]] for k in range(3))
        o = rays[sl, 0:3].astype(np.float64)
        d = rays[sl, 4:7].astype(np.float64)
        n = np.cross(v0 - v2, v1 - v2)
        dn = np.sum(d * n, axis=1)
        th = np.sum((v2 - o) * n, axis=1) / dn
        p = o + th[:, None] * d
        nn = np.sum(n * n, axis=1)
        b0 = np.sum(np.cross(v1 - p, v2 - p) * n, axis=1) / nn
        b1 = np.sum(np.cross(v2 - p, v0 - p) * n, axis=1) / nn
        b = np.stack([b0, b1, 1.0 - b0 - b1], axis=1)
        nl = np.sqrt(nn)
        alt = np.stack([nl / np.linalg.norm(v1 - v2, axis=1), nl / np.linalg.norm(v2 - v0, axis=1),
                        nl / np.linalg.norm(v0 - v1, axis=1)], axis=1)
        cos = np.abs(dn) / nl
        tt[sl], B[sl], ndn[sl] = th, b, n / dn[:, None]
        # edge distance over the per-h margin: keeps the triangle while > 4 h
        margin_h[sl] = np.min(b * alt, axis=1) / (1.0 + 1.0 / np.maximum(cos, 1e-300))
    return tt, B, ndn, margin_h


def _fd_check(v, t, rays, ids, mask, t_tgt, grad, sample, geo, h=1e-7):
    """FP64 central differences (step h) of L = sum_mask (t(V) - t_tgt)^2 with
    the hit triangles held fixed, on the sampled vertices.  Returns per
    coordinate the GPU gradient, the differences and the bound of what FP32
    adds: each forward t rounded once (|dt| <= 2^-24 |t|, so 2^-23 |t| |dt/dV|
    per ray) and the FP32 conversion and atomic summation of the K per-ray
    terms ((K + 1) 2^-24 sum |term|)."""
    tt, B, ndn, _ = geo
    hit_rays = np.nonzero(mask)[0]
    tri_of = t[ids[hit_rays]]
    vd = v.astype(np.float64)
    g_fd_all, g_all, bound_all = [], [], []
    for k in sample:
        sel = np.any(tri_of == k, axis=1)
        rows = hit_rays[sel]
        which = np.argmax(tri_of[sel] == k, axis=1)
        for c in range(3):
            vals = []
            for s in (+1, -1):
                vp = vd.copy()
                vp[k, c] += s * h
                tp = _plane_t64(vp, t, rays[rows], ids[rows])
                vals.append(np.sum((tp - t_tgt[rows]) ** 2))
            g_fd_all.append((vals[0] - vals[1]) / (2 * h))
            g_all.append(grad[k, c])
            dtdv = np.abs(B[rows, which] * ndn[rows, c])
            terms = np.abs(2.0 * (tt[rows] - t_tgt[rows])) * dtdv
            bound_all.append(np.sum(2.0 ** -23 * np.abs(tt[rows]) * dtdv) + (rows.size + 1) * 2.0 ** -24 * terms.sum())
    return np.array(g_all, np.float64), np.array(g_fd_all), np.array(bound_all)


@pytest.mark.parametrize("level,res,nsample", [(4, 96, 256), (8, 512, 1024)])
def test_backward_vs_fp64_finite_differences(ctx, level, res, nsample):
    """Config 4 at level 8: 1,310,720 triangles, 16 x 512^2 rays, 1,024 sampled
    vertices, FP64 central differences at h = 1e-7 (truncation far below 1e-4
    even on the small config-4 triangles).  Only rays whose primitive id may
    change under +-h are left out of the loss."""
    h = 1e-7
    v, t = scenes.displaced_icosphere(level, seed=1, amplitude=0.05)
    vt, _ = scenes.displaced_icosphere(level, seed=2, amplitude=0.05)
    rays = scenes.orbit_rays(16, res, 3.0)
    dt = DifferentiableTracer(torch.from_numpy(v).cuda(), torch.from_numpy(t).cuda(), ctx)
    tgt_tr = DifferentiableTracer(torch.from_numpy(vt).cuda(), torch.from_numpy(t).cuda(), ctx)
    rays_t = torch.from_numpy(rays).cuda()
    with torch.no_grad():
        t_tgt, _, _, ids_tgt = tgt_tr(torch.from_numpy(vt).cuda(), rays_t)
        _, _, _, ids0 = dt(torch.from_numpy(v).cuda(), rays_t)
    ids0 = ids0.cpu().numpy()
    geo = _hit_geometry(v, t, rays, ids0)
    stable = geo[3] > 4.0 * h
    both = (ids0 >= 0) & (ids_tgt.cpu().numpy() >= 0)
    m = both & stable
    assert (both & ~stable).sum() < 0.01 * both.sum()       # few rays sit that close to an edge
    mask = torch.from_numpy(m).cuda()
    verts = torch.from_numpy(v).cuda().requires_grad_(True)
    tt, uu, vv, ids = dt(verts, rays_t)
    assert torch.equal(ids.cpu(), torch.from_numpy(ids0))
    loss = (((tt - t_tgt) ** 2) * mask).sum()
    loss.backward()
    grad = verts.grad.cpu().numpy()
    # direct C-ABI call gives the same gradient up to float atomics order
    res = torch.stack([ids, tt.view(torch.int32), uu.view(torch.int32), vv.view(torch.int32)], 1).contiguous()
    g2 = backward(ctx, verts.detach(), dt.triangles, rays_t, res, dl_dt=(2 * (tt - t_tgt) * mask).detach()).cpu()
    assert torch.allclose(torch.from_numpy(grad), g2, rtol=1e-5, atol=1e-6 * float(g2.abs().max()))
    # the forward t is the FP64 plane t of the hit triangle rounded once
    hit = np.nonzero(ids0 >= 0)[0]
    t64 = geo[0][hit]
    assert np.all(np.abs(tt.detach().cpu().numpy()[hit] - t64) <= 2.0 ** -24 * np.abs(t64) * (1 + 1e-6))
    ids_np = ids.cpu().numpy()
    assert m.mean() > 0.2
    rng = np.random.default_rng(64)
    touched = np.unique(t[ids_np[m]])
    sample = rng.choice(touched, min(nsample, touched.size), replace=False)
    t_tgt64 = t_tgt.cpu().numpy().astype(np.float64)
    g, g_fd, bound = _fd_check(v, t, rays, ids_np, m, t_tgt64, grad, sample, geo, h)
    # Contract: every sampled coordinate within 1e-4 relative of the FP64
    # differences, beyond what the FP32 forward t and the FP32 scatter carry.
    excess = np.abs(g - g_fd) / (1e-4 * np.abs(g_fd) + bound)
    assert np.max(excess) <= 1.0, (np.max(excess), np.argmax(excess), g[np.argmax(excess)], g_fd[np.argmax(excess)])
    assert np.linalg.norm(g - g_fd) / np.linalg.norm(g_fd) <= 1e-4
    # and the gradient is not trivially zero
    assert np.max(np.abs(g_fd)) > 0


def test_backward_u_v_gradients(ctx):
    """dL/du and dL/dv paths: L = sum(u) + 2 sum(v) against FP64 central
    differences at h = 1e-7, every sampled coordinate within 1e-4 relative
    beyond the FP32 conversion and scatter of the per-ray terms."""
    h = 1e-7
    v, t = scenes.displaced_icosphere(3, seed=5, amplitude=0.05)
    rays = scenes.orbit_rays(4, 64, 3.0)
    dt = DifferentiableTracer(torch.from_numpy(v).cuda(), torch.from_numpy(t).cuda(), ctx)
    verts = torch.from_numpy(v).cuda().requires_grad_(True)
    tt, uu, vv, ids = dt(verts, torch.from_numpy(rays).cuda())
    ids_np = ids.cpu().numpy()
    stable = (ids_np >= 0) & (_hit_geometry(v, t, rays, ids_np)[3] > 4.0 * h)
    m = torch.from_numpy(stable).cuda()
    ((uu * m).sum() + 2 * (vv * m).sum()).backward()
    grad = verts.grad.cpu().numpy().astype(np.float64)
    hit = np.nonzero(stable)[0]

    def terms64(vd, rows):
        tri = t[ids_np[rows]]
        v0, v1, v2 = (vd[tri[:, k]] for k in range(3))
        o = rays[rows, 0:3].astype(np.float64)
        d = rays[rows, 4:7].astype(np.float64)
        e1, e2 = v0 - v2, v1 - v2
        M = np.stack([e1, e2, -d], axis=2)
        x = np.linalg.solve(M, (o - v2)[..., None])[..., 0]
        return x[:, 0] + 2 * x[:, 1]

    vd = v.astype(np.float64)
    rng = np.random.default_rng(65)
    tri_of = t[ids_np[hit]]
    worst = 0.0
    for k in rng.choice(np.unique(tri_of), 64, replace=False):
        rows = hit[np.any(tri_of == k, axis=1)]
        for c in range(3):
            vp, vm = vd.copy(), vd.copy()
            vp[k, c] += h
            vm[k, c] -= h
            per_ray = (terms64(vp, rows) - terms64(vm, rows)) / (2 * h)
            g_fd = per_ray.sum()
            bound = (rows.size + 1) * 2.0 ** -24 * np.abs(per_ray).sum()
            worst = max(worst, abs(grad[k, c] - g_fd) / (1e-4 * abs(g_fd) + bound))
    assert worst <= 1.0, worst
