"""Batched meshes (hx_build_batch and friends) against the looped per-mesh path.

Each mesh of a batch must get exactly the BVH a separate build gives: the
oracle builds every mesh alone, and the batched Bin tree (canonical form), Woop
rows and mesh-local triangle ids must match bit for bit.  The batched trace,
refit and adjoint are checked per mesh against the oracle trace, the oracle
refit and the single-mesh adjoint on the GPU.
"""
import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import scenes
from hippie_amd.diff import backward as single_backward

from adjoint_ref import adjoint_terms_fp64, assert_scatter_within_bound

pytestmark = pytest.mark.gpu


def _meshes(n, seed, lo=2, hi=6000):
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(int)
    return [scenes.batch_mesh(int(c), rng, bool(rng.random() < 0.5)) for c in counts]


def _check_meshes(oracle, hb, meshes, bits):
    for m, (v, t) in enumerate(meshes):
        got = hb.mesh_host(m)
        ref = oracle.build_lbvh(v, t, bits=bits)
        assert (got["n_interior"], got["n_leaves"]) == (ref["n_interior"], ref["n_leaves"]), f"mesh {m} counts"
        assert oracle.validate_bin(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"], v, t) == 0
        cg = oracle.bin_canonical(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"])
        co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
        assert np.array_equal(cg, co), f"mesh {m} canonical tree"
        assert np.array_equal(got["woop"].view(np.uint32), oracle.woopify(v, t, got["tri_idx"]).view(np.uint32))


@pytest.mark.parametrize("bits", [30, 60])
def test_batch_build_matches_each_mesh_alone(oracle, ctx, bits):
    meshes = _meshes(40, seed=bits)
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes(meshes), morton_bits=bits)
    _check_meshes(oracle, hb, meshes, bits)


def test_batch_single_mesh_and_duplicates(oracle, ctx):
    """One mesh through the batch API, and meshes whose triangles coincide."""
    v, t = scenes.uniform_soup(5000, seed=3)
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes([(v, t)]))
    _check_meshes(oracle, hb, [(v, t)], 60)
    tri = np.array([[0, 1, 2]] * 7, np.int32)
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    meshes = [(verts, tri), (v, t), (verts, tri[:2])]
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes(meshes))
    _check_meshes(oracle, hb, meshes, 60)


def test_batch_trace_refit_backward(oracle, ctx):
    meshes = _meshes(24, seed=7, lo=16, hi=4000)
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb)
    rays, owner = scenes.batch_rays(meshes, per_1k=2000, seed=5)
    d_rays = torch.as_tensor(rays, device="cuda")
    d_owner = torch.as_tensor(owner, device="cuda")
    res = hb.trace(d_rays, d_owner).cpu().numpy()
    for m in range(len(meshes)):
        sel = owner == m
        g = hb.mesh_host(m)
        exp, ov = oracle.trace(g["nodes"], g["woop"], g["tri_idx"], rays[sel])
        assert ov == 0
        assert np.array_equal(res[sel, 0], exp[:, 0]) and np.array_equal(res[sel, 1], exp[:, 1]), f"mesh {m}"
    # Refit to perturbed vertices: every mesh as the oracle refit of its nodes.
    before = [hb.mesh_host(m) for m in range(len(meshes))]
    rng = np.random.default_rng(9)
    v_all = mb.scene.vertices.cpu().numpy()
    v2 = (v_all + rng.normal(0.0, 0.001, v_all.shape)).astype(np.float32)
    hb.refit(v2)
    base = 0
    for m, (v, t) in enumerate(meshes):
        vm = v2[base:base + len(v)]
        base += len(v)
        g = hb.mesh_host(m)
        exp = oracle.refit_final(before[m]["nodes"], g["n_interior"], g["n_leaves"], g["tri_idx"], vm, t)
        assert np.array_equal(g["nodes"], exp), f"mesh {m} refit"
    # Later refits rewrite only the box rows recorded by the first: a second
    # and a third refit give the oracle's nodes too.
    for k, sigma in enumerate((0.002, 0.001)):
        vk = (v_all + np.random.default_rng(20 + k).normal(0.0, sigma, v_all.shape)).astype(np.float32)
        hb.refit(vk)
        base = 0
        for m, (v, t) in enumerate(meshes):
            vm = vk[base:base + len(v)]
            base += len(v)
            g = hb.mesh_host(m)
            exp = oracle.refit_final(before[m]["nodes"], g["n_interior"], g["n_leaves"], g["tri_idx"], vm, t)
            assert np.array_equal(g["nodes"], exp), f"mesh {m} refit {k + 2}"
    hb.refit(v2)
    # Adjoint: batched scatter equals the single-mesh adjoint mesh by mesh.
    res_t = hb.trace(d_rays, d_owner)
    dl_dt = torch.rand(len(rays), device="cuda", dtype=torch.float32)
    grad = hb.backward(d_rays, d_owner, res_t, dl_dt=dl_dt)
    g64, S, K = adjoint_terms_fp64(mb.scene.vertices, mb.scene.triangles, d_rays, res_t, dl_dt, owner=d_owner,
                                   tri_offsets=mb.offsets[:-1])
    assert_scatter_within_bound(grad, g64, S, K, "batched adjoint")
    base = 0
    for m, (v, t) in enumerate(meshes):
        sel = torch.as_tensor(owner == m, device="cuda")
        vm = torch.as_tensor(v2[base:base + len(v)], device="cuda")
        tm = torch.as_tensor(t, device="cuda")
        gm = single_backward(ctx, vm, tm, d_rays[sel].contiguous(), res_t[sel].contiguous(),
                             dl_dt=dl_dt[sel].contiguous())
        sl = slice(3 * base, 3 * (base + len(v)))
        assert_scatter_within_bound(gm, g64[sl], S[sl], K[sl], f"mesh {m} adjoint")
        base += len(v)


def test_batch_rejects_tiny_meshes(ctx):
    v, t = scenes.uniform_soup(10, seed=1)
    with pytest.raises(hx._lib.HxError):
        hx.HipBatch(ctx, hx.MeshBatch.from_meshes([(v, t), (v, t[:1])]))


def test_batch_diff_step_graph_replays_match_eager(ctx):
    """refit -> trace -> loss -> adjoint captured in one graph: every replay
    equals the eager step (the capture holds no memset node: all clears are
    kernels)."""
    meshes = _meshes(16, seed=11, lo=32, hi=3000)
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb)
    rays, owner = scenes.batch_rays(meshes, per_1k=1000, seed=6)
    d_rays = torch.as_tensor(rays, device="cuda")
    d_owner = torch.as_tensor(owner, device="cuda")
    res = hb.trace(d_rays, d_owner)
    t_ref = res[:, 1].view(torch.float32).clone() * 0.999
    grad = torch.zeros_like(mb.scene.vertices)
    dl_dt = torch.zeros(len(rays), dtype=torch.float32, device="cuda")

    def step():
        hb.refit()
        hb.trace(d_rays, d_owner, results=res)
        t = res[:, 1].view(torch.float32)
        torch.where(res[:, 0] >= 0, 2.0 * (t - t_ref), torch.zeros_like(t), out=dl_dt)
        grad.zero_()
        hb.backward(d_rays, d_owner, res, dl_dt=dl_dt, out=grad)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ctx.set_stream(side)
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    try:
        with torch.cuda.graph(g):
            ctx.set_stream(torch.cuda.current_stream())
            step()
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    torch.cuda.synchronize()
    rng = np.random.default_rng(2)
    for k in range(3):
        v2 = mb.scene.vertices + torch.as_tensor(rng.normal(0.0, 1e-4, tuple(mb.scene.vertices.shape)),
                                                 dtype=torch.float32, device="cuda")
        mb.scene.vertices.copy_(v2)
        step()                                   # eager
        torch.cuda.synchronize()
        res_e, grad_e = res.clone(), grad.clone()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(res, res_e), f"replay {k}: results"
        # Atomic float32 scatter: both within the summation bound of the FP64 sum.
        g64, S, K = adjoint_terms_fp64(mb.scene.vertices, mb.scene.triangles, d_rays, res, dl_dt, owner=d_owner,
                                       tri_offsets=mb.offsets[:-1])
        assert_scatter_within_bound(grad_e, g64, S, K, f"step {k} eager")
        assert_scatter_within_bound(grad, g64, S, K, f"step {k} replay")


def _check_meshes_presplit(oracle, hb, meshes, bits, beta):
    for m, (v, t) in enumerate(meshes):
        got = hb.mesh_host(m)
        ref = oracle.build_lbvh(v, t, bits=bits, presplit_beta=beta)
        assert got["tri_idx"].shape[0] == ref["tri_idx"].shape[0], f"mesh {m} references"
        assert (got["n_interior"], got["n_leaves"]) == (ref["n_interior"], ref["n_leaves"]), f"mesh {m} counts"
        assert oracle.validate_bin_refs(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"], v, t) == 0
        cg = oracle.bin_canonical(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"])
        co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
        assert np.array_equal(cg, co), f"mesh {m} canonical tree"
        assert np.array_equal(got["woop"].view(np.uint32), oracle.woopify(v, t, got["tri_idx"]).view(np.uint32))


def _config5_meshes(n, seed, lo=64, hi=5000):
    """Uneven meshes, 30% of them long thin slivers (aspect up to 1:200)."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n))).astype(int)
    out = []
    for c in counts:
        if rng.random() < 0.3:
            out.append(scenes.sliver_soup(int(c), rng))
        else:
            out.append(scenes.batch_mesh(int(c), rng, bool(rng.random() < 0.5)))
    return out


@pytest.mark.parametrize("beta", [0.3, 1.0])
def test_batch_presplit_matches_each_mesh_alone(oracle, ctx, beta):
    """Config 5 shape: every mesh presplit alone (own box, own budget beta * n_m)."""
    meshes = _config5_meshes(30, seed=21)
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes(meshes), morton_bits=60, presplitting=1, presplitter_beta=beta)
    _check_meshes_presplit(oracle, hb, meshes, 60, beta)
    hb.rebuild()
    _check_meshes_presplit(oracle, hb, meshes, 60, beta)


def test_batch_presplit_trace_refit_backward(oracle, ctx):
    meshes = _config5_meshes(16, seed=22, lo=32, hi=3000)
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb, presplitting=1, presplitter_beta=0.3)
    rays, owner = scenes.batch_rays(meshes, per_1k=2000, seed=5)
    d_rays = torch.as_tensor(rays, device="cuda")
    d_owner = torch.as_tensor(owner, device="cuda")
    res = hb.trace(d_rays, d_owner).cpu().numpy()
    for m, (v, t) in enumerate(meshes):
        sel = owner == m
        g = hb.mesh_host(m)
        exp, ov = oracle.trace(g["nodes"], g["woop"], g["tri_idx"], rays[sel])
        brute = oracle.trace_brute(v, t, rays[sel])
        assert ov == 0
        assert np.array_equal(res[sel, 0], exp[:, 0]) and np.array_equal(res[sel, 1], exp[:, 1]), f"mesh {m}"
        assert np.array_equal(res[sel, 0], brute[:, 0]) and np.array_equal(res[sel, 1], brute[:, 1]), f"mesh {m}"
    # Refit: leaf boxes grow back to whole triangles, as HipBVH::update does.
    before = [hb.mesh_host(m) for m in range(len(meshes))]
    rng = np.random.default_rng(9)
    v_all = mb.scene.vertices.cpu().numpy()
    v2 = (v_all + rng.normal(0.0, 0.001, v_all.shape)).astype(np.float32)
    hb.refit(v2)
    base = 0
    for m, (v, t) in enumerate(meshes):
        vm = v2[base:base + len(v)]
        base += len(v)
        g = hb.mesh_host(m)
        exp = oracle.refit_final(before[m]["nodes"], g["n_interior"], g["n_leaves"], g["tri_idx"], vm, t)
        assert np.array_equal(g["nodes"], exp), f"mesh {m} refit"
    # Adjoint: batched scatter equals the single-mesh adjoint mesh by mesh.
    res_t = hb.trace(d_rays, d_owner)
    dl_dt = torch.rand(len(rays), device="cuda", dtype=torch.float32)
    grad = hb.backward(d_rays, d_owner, res_t, dl_dt=dl_dt)
    g64, S, K = adjoint_terms_fp64(mb.scene.vertices, mb.scene.triangles, d_rays, res_t, dl_dt, owner=d_owner,
                                   tri_offsets=mb.offsets[:-1])
    assert_scatter_within_bound(grad, g64, S, K, "batched adjoint")
    base = 0
    for m, (v, t) in enumerate(meshes):
        sel = torch.as_tensor(owner == m, device="cuda")
        vm = torch.as_tensor(v2[base:base + len(v)], device="cuda")
        tm = torch.as_tensor(t, device="cuda")
        gm = single_backward(ctx, vm, tm, d_rays[sel].contiguous(), res_t[sel].contiguous(),
                             dl_dt=dl_dt[sel].contiguous())
        sl = slice(3 * base, 3 * (base + len(v)))
        assert_scatter_within_bound(gm, g64[sl], S[sl], K[sl], f"mesh {m} adjoint")
        base += len(v)


def _check_meshes_ploc(oracle, hb, meshes, bits, radius, beta=None):
    for m, (v, t) in enumerate(meshes):
        got = hb.mesh_host(m)
        ref = oracle.build_ploc(v, t, bits=bits, radius=radius, presplit_beta=beta)
        assert (got["n_interior"], got["n_leaves"]) == (ref["n_interior"], ref["n_leaves"]), f"mesh {m} counts"
        assert oracle.validate_bin_refs(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"], v, t) == 0
        cg = oracle.bin_canonical(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"])
        co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
        assert np.array_equal(cg, co), f"mesh {m} canonical tree"
        assert np.array_equal(got["woop"].view(np.uint32), oracle.woopify(v, t, got["tri_idx"]).view(np.uint32))


@pytest.mark.parametrize("radius", [1, 8, 25])
def test_batch_ploc_matches_each_mesh_alone(oracle, ctx, radius):
    """Clusters merge only inside their mesh: every mesh gets its own PLOC tree."""
    meshes = _meshes(30, seed=30 + radius, lo=2, hi=5000)
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes(meshes), method="ploc", ploc_radius=radius)
    _check_meshes_ploc(oracle, hb, meshes, 60, radius)


@pytest.mark.parametrize("method", ["lbvh", "ploc"])
def test_batch_two_triangle_meshes(oracle, ctx, method):
    """The smallest meshes a batch takes (two triangles) first, between others
    and last; a one-triangle mesh is rejected."""
    two = [scenes.batch_mesh(2, np.random.default_rng(s), False) for s in range(4)]
    meshes = [two[0]] + _meshes(6, seed=40, lo=2, hi=300) + [two[1], two[2]] + _meshes(3, seed=41, lo=2, hi=50)
    meshes.append(two[3])
    hb = hx.HipBatch(ctx, hx.MeshBatch.from_meshes(meshes), method=method, ploc_radius=8)
    if method == "lbvh":
        _check_meshes(oracle, hb, meshes, 60)
    else:
        _check_meshes_ploc(oracle, hb, meshes, 60, 8)
    one = scenes.batch_mesh(1, np.random.default_rng(5), False)
    with pytest.raises(hx._lib.HxError, match="fewer than 2"):
        hx.HipBatch(ctx, hx.MeshBatch.from_meshes([two[0], one]), method=method)


def test_batch_ploc_presplit_config5(oracle, ctx):
    """Config 5 shape: uneven meshes with 30% slivers, presplit (beta 1.0 so the
    slivers get a budget), PLOC r=25; rebuild, trace and refit per mesh."""
    meshes = _config5_meshes(24, seed=31, lo=64, hi=4000)
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb, method="ploc", ploc_radius=25, presplitting=1, presplitter_beta=1.0)
    _check_meshes_ploc(oracle, hb, meshes, 60, 25, beta=1.0)
    hb.rebuild()
    _check_meshes_ploc(oracle, hb, meshes, 60, 25, beta=1.0)
    rays, owner = scenes.batch_rays(meshes, per_1k=1500, seed=7)
    d_rays = torch.as_tensor(rays, device="cuda")
    d_owner = torch.as_tensor(owner, device="cuda")
    res = hb.trace(d_rays, d_owner).cpu().numpy()
    for m, (v, t) in enumerate(meshes):
        sel = owner == m
        brute = oracle.trace_brute(v, t, rays[sel])
        assert np.array_equal(res[sel, 0], brute[:, 0]) and np.array_equal(res[sel, 1], brute[:, 1]), f"mesh {m}"
    before = [hb.mesh_host(m) for m in range(len(meshes))]
    rng = np.random.default_rng(3)
    v_all = mb.scene.vertices.cpu().numpy()
    v2 = (v_all + rng.normal(0.0, 0.001, v_all.shape)).astype(np.float32)
    hb.refit(v2)
    base = 0
    for m, (v, t) in enumerate(meshes):
        vm = v2[base:base + len(v)]
        base += len(v)
        g = hb.mesh_host(m)
        exp = oracle.refit_final(before[m]["nodes"], g["n_interior"], g["n_leaves"], g["tri_idx"], vm, t)
        assert np.array_equal(g["nodes"], exp), f"mesh {m} refit"
