"""The benchmark's workloads at full size, checked end to end on the GPU.

Config 1: the 65,536-triangle soup, 30-bit LBVH, 1024x1024 primary rays of the
pinhole at (0.5, 0.5, -1.5): tree and every hit equal to the oracle's.

Config 3: 4096 meshes of log-uniform 64..262,144 triangles (~128M), one
batched LBVH build, trace, refit and adjoint.  Each mesh of the batch must be
byte for byte the BVH a separate build of that mesh gives (nodes, Woop rows,
triangle ids), every separate build must pass the device validation, the
batched hits must equal the separate traces' bit for bit, and a seeded sample
of 64 meshes spanning the size range (the largest included) must equal the
oracle's builds and traces.

Config 5: 512 meshes of log-uniform 1k..200k triangles, 30% long thin
slivers, presplit (beta 0.3) PLOC r = 25, 64M rays; the same checks, plus the
graph-captured refit -> trace -> loss -> adjoint step against the eager one.
"""
import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import scenes, workloads
from hippie_amd.diff import backward as single_backward

from adjoint_ref import adjoint_terms_fp64, assert_scatter_within_bound

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _dev_copy(ctx, ptr, shape, dtype):
    out = torch.empty(shape, dtype=dtype, device="cuda")
    if out.numel():
        ctx.memcpy(out.data_ptr(), ptr, out.numel() * out.element_size())
    return out


class _BatchSnapshot:
    """Device copies of every mesh's nodes, Woop rows and triangle ids."""

    def __init__(self, ctx, hb, n_meshes):
        self.views = [hb.mesh_view(m) for m in range(n_meshes)]
        self.nodes, self.woop, self.tri = [], [], []
        for v in self.views:
            nn = v.n_interior + v.n_leaves
            self.nodes.append(_dev_copy(ctx, v.nodes, (nn, 16), torch.int32))
            self.woop.append(_dev_copy(ctx, v.woop, (v.n_references, 12), torch.float32))
            self.tri.append(_dev_copy(ctx, v.tri_indices, (v.n_references,), torch.int32))

    def host(self, m):
        v = self.views[m]
        return dict(nodes=self.nodes[m].cpu().numpy(), woop=self.woop[m].cpu().numpy(),
                    tri_idx=self.tri[m].cpu().numpy(), n_interior=v.n_interior, n_leaves=v.n_leaves)


def _bvh_dev(ctx, bvh):
    v = bvh.view()
    nn = v.n_interior + v.n_leaves
    return (_dev_copy(ctx, v.nodes, (nn, 16), torch.int32),
            _dev_copy(ctx, v.woop, (v.n_references, 12), torch.float32),
            _dev_copy(ctx, v.tri_indices, (v.n_references,), torch.int32))


def _same_bits(a, b):
    return a.shape == b.shape and torch.equal(a.view(torch.int32), b.view(torch.int32))


def _sample(counts, k, seed):
    """k mesh indices spread over the size range: the largest and smallest,
    then a seeded draw from each of k-2 size quantile bins."""
    order = np.argsort(counts, kind="stable")
    pick = {int(order[-1]), int(order[0])}
    rng = np.random.default_rng(seed)
    for b in np.array_split(order[1:-1], k - 2):
        if len(b):
            pick.add(int(rng.choice(b)))
    return sorted(pick)


def _check_oracle_mesh(oracle, got, ref, v, t, rays_m, res_m, tag):
    assert (got["n_interior"], got["n_leaves"]) == (ref["n_interior"], ref["n_leaves"]), f"{tag} counts"
    cg = oracle.bin_canonical(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"])
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co), f"{tag} canonical tree"
    assert np.array_equal(got["woop"].view(np.uint32), oracle.woopify(v, t, got["tri_idx"]).view(np.uint32)), \
        f"{tag} Woop rows"
    exp, ov = oracle.trace(got["nodes"], got["woop"], got["tri_idx"], rays_m)
    assert ov == 0
    assert np.array_equal(res_m[:, :2], exp[:, :2]), f"{tag} hits vs oracle trace"
    h = exp[:, 0] >= 0
    assert np.array_equal(res_m[h, 2:], exp[h, 2:]), f"{tag} u, v"


def _trace_into(ctx, bvh, rays, out):
    hx.HipTracer(ctx, bvh).trace(hx.RayBuffer.wrap(rays, out))
    return out


def test_config1_full_size(oracle, ctx):
    v, t = scenes.uniform_soup(65536, seed=1)
    rays = scenes.config1_rays(1024, 1024)
    assert rays.shape[0] == 1048576
    bvh = hx.LBVHBuilder(ctx, morton_bits=30).build(hx.Scene(v, t))
    assert bvh.validate() == 0
    ref = oracle.build_lbvh(v, t, bits=30)
    got = dict(nodes=bvh.nodes_host(), woop=bvh.woop_host(), tri_idx=bvh.tri_indices_host(),
               n_interior=bvh.n_interior, n_leaves=bvh.n_leaves)
    rb = hx.RayBuffer(rays)
    res = hx.HipTracer(ctx, bvh).trace(rb).cpu().numpy()
    _check_oracle_mesh(oracle, got, ref, v, t, rays, res, "config 1")
    assert ctx.overflow_count() == 0
    hits = res[:, 0] >= 0
    assert 0.05 < hits.mean() < 0.5
    miss = ~hits
    assert np.all(res[miss, 1].view(np.float32) == rays[miss, 7])
    # A BVH-free brute force on a seeded sample of the rays.
    sel = np.random.default_rng(1).choice(rays.shape[0], 4096, replace=False)
    brute = oracle.trace_brute(v, t, rays[sel])
    assert np.array_equal(res[sel, :2], brute[:, :2])
    for method in ("aila", "twopoint"):
        rs = hx.RayBuffer(rays)
        hx.HipTracer(ctx, bvh).trace_sort(rs, method=method, bits=60)
        assert np.array_equal(rs.results.cpu().numpy()[:, :2], res[:, :2]), method


def test_config3_full_size(oracle, ctx):
    """4096 meshes, ~128M triangles, ~33M rays: batched build, trace, refit and
    adjoint against separate per-mesh builds and the oracle."""
    w = workloads.batched_build_trace(torch.device("cuda", 0))
    M, counts = len(w["counts"]), w["counts"]
    N, R = int(w["offs"][-1]), int(w["rays"].shape[0])
    _say(f"config 3: {M} meshes, {N} triangles ({counts.min()}..{counts.max()}), {R} rays")
    assert M == 4096 and counts.min() >= 64 and counts.max() <= 262144 and N > 100_000_000
    mb = hx.MeshBatch(w["V"], w["T"], w["offs"])
    hb = hx.HipBatch(ctx, mb, morton_bits=60)
    snap_a = _BatchSnapshot(ctx, hb, M)
    rays, owner, roffs, vbase = w["rays"], w["owner"], w["roffs"], w["vbase"]
    res_a = hb.trace(rays, owner)
    g = torch.Generator(device="cuda")
    g.manual_seed(33)
    dl_dt = torch.rand(R, generator=g, device="cuda")
    grad_a = hb.backward(rays, owner, res_a, dl_dt=dl_dt)
    # Atomic float32 scatter: held to the summation bound of the FP64 sum.
    g64, S, K = adjoint_terms_fp64(mb.scene.vertices, mb.scene.triangles, rays, res_a, dl_dt, owner=owner,
                                   tri_offsets=mb.offsets[:-1])
    assert_scatter_within_bound(grad_a, g64, S, K, "batched adjoint")
    V2 = w["V"] + 1e-3 * torch.randn(w["V"].shape, generator=g, device="cuda")
    hb.refit(V2)
    snap_b = _BatchSnapshot(ctx, hb, M)
    res_b = hb.trace(rays, owner)
    torch.cuda.synchronize()
    _say("config 3: batched build, trace, adjoint and refit done; per-mesh builds")

    builder = hx.LBVHBuilder(ctx, morton_bits=60)
    for m in range(M):
        v0, v1 = int(vbase[m]), int(vbase[m + 1])
        r0, r1 = int(roffs[m]), int(roffs[m + 1])
        scene = hx.Scene(w["V"][v0:v1], w["tris_local"][m])
        bvh = builder.build(scene)
        assert bvh.validate() == 0, f"mesh {m} validate"
        nodes, woop, tri = _bvh_dev(ctx, bvh)
        assert _same_bits(nodes, snap_a.nodes[m]), f"mesh {m} nodes"
        assert _same_bits(woop, snap_a.woop[m]), f"mesh {m} Woop rows"
        assert torch.equal(tri, snap_a.tri[m]), f"mesh {m} triangle ids"
        out = torch.zeros((r1 - r0, 4), dtype=torch.int32, device="cuda")
        _trace_into(ctx, bvh, rays[r0:r1], out)
        assert torch.equal(out[:, :2], res_a[r0:r1, :2]), f"mesh {m} hits"
        gm = single_backward(ctx, scene.vertices, scene.triangles, rays[r0:r1], res_a[r0:r1],
                             dl_dt=dl_dt[r0:r1].contiguous())
        sl = slice(3 * v0, 3 * v1)
        assert_scatter_within_bound(gm, g64[sl], S[sl], K[sl], f"mesh {m} adjoint")
        s2 = hx.Scene(V2[v0:v1], w["tris_local"][m])
        bvh.update(s2)
        nodes, woop, _ = _bvh_dev(ctx, bvh)
        assert _same_bits(nodes, snap_b.nodes[m]), f"mesh {m} refit nodes"
        assert _same_bits(woop, snap_b.woop[m]), f"mesh {m} refit Woop rows"
        out.zero_()
        _trace_into(ctx, bvh, rays[r0:r1], out)
        assert torch.equal(out[:, :2], res_b[r0:r1, :2]), f"mesh {m} hits after refit"
        del bvh
        if m % 512 == 511:
            _say(f"config 3: {m + 1} meshes equal their separate builds")

    sample = _sample(counts, 64, seed=3)
    assert int(np.argmax(counts)) in sample
    _say(f"config 3: oracle on {len(sample)} meshes of {counts[sample].min()}..{counts[sample].max()} triangles")
    V_h = w["V"].cpu().numpy()
    rays_h, res_h = rays.cpu().numpy(), res_a.cpu().numpy()
    for m in sample:
        v = V_h[int(vbase[m]):int(vbase[m + 1])]
        t = w["tris_local"][m].cpu().numpy()
        ref = oracle.build_lbvh(v, t, bits=60)
        r0, r1 = int(roffs[m]), int(roffs[m + 1])
        _check_oracle_mesh(oracle, snap_a.host(m), ref, v, t, rays_h[r0:r1], res_h[r0:r1], f"mesh {m}")


def test_config5_full_size(oracle, ctx):
    """512 sliver/presplit meshes with PLOC r = 25, 64M rays; batched rebuild,
    trace, refit against separate builds and the oracle; the graph-captured
    differentiable step replays to the eager step."""
    dev = torch.device("cuda", 0)
    w = workloads.batched_presplit_diff(dev)
    meshes, counts, rays, owner, roffs = w["meshes"], w["counts"], w["rays"], w["owner"], w["roffs"]
    M, R = len(meshes), int(rays.shape[0])
    _say(f"config 5: {M} meshes, {int(counts.sum())} triangles ({counts.min()}..{counts.max()}), {R} rays")
    assert M == 512 and R >= 60_000_000
    params = dict(method="ploc", ploc_radius=25, presplitting=1, presplitter_beta=0.3)
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb, **params)
    hb.rebuild()
    snap = _BatchSnapshot(ctx, hb, M)
    res = hb.trace(rays, owner)
    assert hb.info("references") > int(counts.sum())
    vb = np.concatenate([[0], np.cumsum([len(v) for v, _ in meshes])])
    builder = hx.PLOCBuilder(ctx, morton_bits=60, ploc_radius=25, presplitting=1, presplitter_beta=0.3)
    for m, (v, t) in enumerate(meshes):
        r0, r1 = int(roffs[m]), int(roffs[m + 1])
        bvh = builder.build(hx.Scene(v, t))
        assert bvh.validate() == 0, f"mesh {m} validate"
        nodes, woop, tri = _bvh_dev(ctx, bvh)
        assert _same_bits(nodes, snap.nodes[m]), f"mesh {m} nodes"
        assert _same_bits(woop, snap.woop[m]), f"mesh {m} Woop rows"
        assert torch.equal(tri, snap.tri[m]), f"mesh {m} triangle ids"
        out = torch.zeros((r1 - r0, 4), dtype=torch.int32, device="cuda")
        _trace_into(ctx, bvh, rays[r0:r1], out)
        assert torch.equal(out[:, :2], res[r0:r1, :2]), f"mesh {m} hits"
        del bvh
        if m % 128 == 127:
            _say(f"config 5: {m + 1} meshes equal their separate builds")

    sample = _sample(counts, 64, seed=5)
    _say(f"config 5: oracle on {len(sample)} meshes of {counts[sample].min()}..{counts[sample].max()} triangles")
    rays_h, res_h = rays.cpu().numpy(), res.cpu().numpy()
    for m in sample:
        v, t = meshes[m]
        r0, r1 = int(roffs[m]), int(roffs[m + 1])
        ref = oracle.build_ploc(v, t, bits=60, radius=25, presplit_beta=0.3)
        got = snap.host(m)
        assert oracle.validate_bin_refs(got["nodes"], got["n_interior"], got["n_leaves"], got["tri_idx"], v, t) == 0
        _check_oracle_mesh(oracle, got, ref, v, t, rays_h[r0:r1], res_h[r0:r1], f"mesh {m}")
        k = min(r1 - r0, 2048)
        brute = oracle.trace_brute(v, t, rays_h[r0:r0 + k])
        assert np.array_equal(res_h[r0:r0 + k, :2], brute[:, :2]), f"mesh {m} brute force"

    # The differentiable step, eager and replayed from a graph.
    t_ref = res[:, 1].view(torch.float32).clone() * 0.999
    grad = torch.zeros_like(mb.scene.vertices)
    dl_dt = torch.zeros(R, dtype=torch.float32, device=dev)

    def step():
        hb.refit()
        hb.trace(rays, owner, results=res)
        t = res[:, 1].view(torch.float32)
        torch.where(res[:, 0] >= 0, 2.0 * (t - t_ref), torch.zeros_like(t), out=dl_dt)
        grad.zero_()
        hb.backward(rays, owner, res, dl_dt=dl_dt, out=grad)

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ctx.set_stream(side)
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    try:
        with torch.cuda.graph(graph):
            ctx.set_stream(torch.cuda.current_stream())
            step()
    finally:
        ctx.set_stream(torch.cuda.current_stream())
    torch.cuda.synchronize()
    g = torch.Generator(device="cuda")
    g.manual_seed(55)
    for k in range(3):
        mb.scene.vertices.add_(1e-4 * torch.randn(mb.scene.vertices.shape, generator=g, device="cuda"))
        step()
        torch.cuda.synchronize()
        res_e, grad_e = res.clone(), grad.clone()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(res, res_e), f"replay {k}: results"
        # The scatter adds float32 terms atomically, so eager and replayed
        # gradients differ in summation order only: each must lie within the
        # float32 summation bound of the FP64 sum of the same terms.
        g64, S, K = adjoint_terms_fp64(mb.scene.vertices, mb.scene.triangles, rays, res, dl_dt, owner=owner,
                                       tri_offsets=mb.offsets[:-1])
        assert_scatter_within_bound(grad_e, g64, S, K, f"step {k} eager")
        assert_scatter_within_bound(grad, g64, S, K, f"step {k} replay")
        del g64, S, K
    # After the refits: every mesh equals a separate refit of its own build.
    snap2 = _BatchSnapshot(ctx, hb, M)
    V_now = mb.scene.vertices
    for m in sample[:16]:
        v, t = meshes[m]
        bvh = builder.build(hx.Scene(v, t))
        bvh.update(hx.Scene(V_now[int(vb[m]):int(vb[m + 1])], t))
        nodes, woop, _ = _bvh_dev(ctx, bvh)
        assert _same_bits(nodes, snap2.nodes[m]), f"mesh {m} refit nodes"
        assert _same_bits(woop, snap2.woop[m]), f"mesh {m} refit Woop rows"
