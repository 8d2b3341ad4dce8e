"""GPU parity of the trace kernels against the CPU oracle.

The oracle traverses the same Bin arrays one lane at a time; closest-hit ids
must be equal and t bit-identical.  Independently, hits are checked against a
BVH-free brute force over all triangles and against FP64 Moller-Trumbore
(|t - t64| <= 1e-6 |t64|).
"""
import ctypes

import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _build(ctx, v, t, bits=60):
    scene = hx.Scene(v, t)
    bvh = hx.LBVHBuilder(ctx, morton_bits=bits).build(scene)
    return scene, bvh


def _trace(ctx, bvh, rays, any_hit=False, stats=False, sort=None):
    rb = hx.RayBuffer(rays, closest_hit=not any_hit)
    rb.results.fill_(0x7F7F7F7F)
    tr = hx.HipTracer(ctx, bvh)
    if sort:
        tr.trace_sort(rb, method=sort, bits=60)
        return rb.results.cpu().numpy(), None, rb.indices.cpu().numpy()
    if stats:
        res, st = tr.trace_stats(rb)
        return res.cpu().numpy(), st.cpu().numpy(), None
    return tr.trace(rb).cpu().numpy(), None, None


def _moller_trumbore64(v, t, rays, ids):
    """FP64 ray-plane distance for the given hit triangles."""
    tri = t[ids]
    v0, v1, v2 = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=1)
    q = np.cross(o - v0, e1)
    return np.sum(e2 * q, axis=1) / det


def _check_against_oracle(oracle, ctx, v, t, rays, any_hit=False):
    scene, bvh = _build(ctx, v, t)
    nodes, woop, tri = bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host()
    res, _, _ = _trace(ctx, bvh, rays, any_hit=any_hit)
    ref, ov = oracle.trace(nodes, woop, tri, rays, any_hit=any_hit)
    assert ov == 0 and ctx.overflow_count() == 0
    if any_hit:
        hit_g, hit_o = res[:, 0] >= 0, ref[:, 0] >= 0
        assert np.array_equal(hit_g, hit_o)
        assert np.array_equal(res[:, 0], ref[:, 0])
    else:
        assert np.array_equal(res[:, 0], ref[:, 0]), f"{(res[:, 0] != ref[:, 0]).sum()} id mismatches"
        assert np.array_equal(res[:, 1], ref[:, 1]), "t bits"
        h = res[:, 0] >= 0
        assert np.array_equal(res[h, 2:], ref[h, 2:]), "u, v bits"
    return scene, bvh, res, ref


def test_trace_config1_shape(oracle, ctx):
    """65,536-triangle soup, 256x256 pinhole rays of the config-1 camera."""
    v, t = scenes.uniform_soup(65536, seed=1)
    rays = scenes.config1_rays(256, 256)
    scene, bvh, res, ref = _check_against_oracle(oracle, ctx, v, t, rays)
    hits = res[:, 0] >= 0
    assert hits.mean() > 0.05  # the unit box covers ~11% of the 60-degree view from z = -1.5
    # GPU vs oracle t is bit-identical (checked above).  Against the exact FP64
    # distance, FP32 Woop carries its own rounding: <= 2.2e-6 relative on this
    # scene (plain FP32 Moller-Trumbore: 1.9e-6), so the bound here is 1e-5.
    t64 = _moller_trumbore64(v, t, rays[hits], res[hits, 0])
    tg = res[hits, 1].view(np.float32).astype(np.float64)
    assert np.all(np.abs(tg - t64) <= 1e-5 * np.abs(t64))
    miss = ~hits
    assert np.all(res[miss, 1].view(np.float32) == rays[miss, 7])


def test_trace_vs_brute_force(oracle, ctx):
    v, t = scenes.uniform_soup(3000, seed=21)
    rays = scenes.random_rays(20000, seed=22)
    scene, bvh, res, ref = _check_against_oracle(oracle, ctx, v, t, rays)
    brute = oracle.trace_brute(v, t, rays)
    assert np.array_equal(res[:, 0], brute[:, 0])
    assert np.array_equal(res[:, 1], brute[:, 1])


def test_trace_any_hit(oracle, ctx):
    v, t = scenes.uniform_soup(20000, seed=23)
    rays = scenes.random_rays(30000, seed=24)
    _check_against_oracle(oracle, ctx, v, t, rays, any_hit=True)


def test_trace_shared_edges_grid(oracle, ctx):
    """Rays through grid vertices and edges hit shared edges: ties resolve to the smaller id."""
    gv, gt = scenes.ground_grid(64)
    n = 64
    xs = (np.arange(n + 1, dtype=np.float32) / n)
    xx, zz = np.meshgrid(xs, xs)
    m = xx.size
    rays = np.zeros((m, 8), np.float32)
    rays[:, 0] = xx.ravel()
    rays[:, 1] = 1.0
    rays[:, 2] = zz.ravel()
    rays[:, 5] = -1.0
    rays[:, 7] = 3.0
    scene, bvh, res, ref = _check_against_oracle(oracle, ctx, gv, gt, rays)
    brute = oracle.trace_brute(gv, gt, rays)
    # These rays run exactly inside box face planes with zero x/z direction.
    # The slab test of HipTracerBinKernels.cu (reciprocal of +-2^-80) collapses
    # such a slab to t = 0, so the BVH traversal may cull boxes that brute
    # force still hits (and then reports a larger-id triangle of the same
    # plane).  Every BVH hit is still a true closest hit: same t as brute force.
    h = res[:, 0] >= 0
    assert h.sum() > 0
    assert np.array_equal(res[h, 1], brute[h, 1])


def test_trace_stats_and_sorted(oracle, ctx):
    v, t = scenes.uniform_soup(30000, seed=25)
    rays = scenes.random_rays(40000, seed=26)
    scene, bvh = _build(ctx, v, t)
    plain, _, _ = _trace(ctx, bvh, rays)
    res_s, stats, _ = _trace(ctx, bvh, rays, stats=True)
    assert np.array_equal(plain, res_s)
    _, ost, _ = oracle.trace(bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host(), rays, want_stats=True)
    # A lane may traverse further before its held-back leaves are tested than
    # the one-lane oracle does (which tests each leaf as soon as it is found),
    # so counts differ ray for ray, but one way only: both walk the same
    # near-first order, and when the lane reaches a node it has tested a
    # subset of the leaves the oracle had by then, so its hit distance is no
    # smaller and every box the oracle enters it enters too.  Each ray's
    # node and triangle counts are therefore at least the oracle's.  Until a
    # ray's first triangle test its hit distance is tmax on both sides, so
    # whether a ray tests any triangle at all, and a ray that misses the root
    # box (one node, no triangle), agree ray for ray.
    live = rays[:, 7] > 0
    assert np.all(stats[live, 0] >= 1)
    assert np.all(stats[:, 0] >= ost[:, 0])
    assert np.all(stats[:, 1] >= ost[:, 1])
    ratio = stats[:, 0].sum() / max(ost[:, 0].sum(), 1)
    tri_ratio = stats[:, 1].sum() / max(ost[:, 1].sum(), 1)
    print(f"traceStats vs one-lane oracle: nodes x{ratio:.4f}, triangles x{tri_ratio:.4f}")
    assert 1.0 <= ratio <= 1.5, ratio
    assert 1.0 <= tri_ratio <= 1.5, tri_ratio
    assert np.array_equal(stats[:, 1] > 0, ost[:, 1] > 0)
    assert np.array_equal(stats[:, 0] == 1, ost[:, 0] == 1)
    for method in ("aila", "twopoint"):
        res_o, _, order = _trace(ctx, bvh, rays, sort=method)
        assert np.array_equal(res_o[:, :2], plain[:, :2])
        h = plain[:, 0] >= 0
        assert np.array_equal(res_o[h], plain[h])
        keys = (oracle.ray_keys_aila64(rays, 60) if method == "aila"
                else oracle.ray_keys_twopoint64(rays, 60, 0.25))
        assert np.array_equal(order, np.argsort(keys, kind="stable").astype(np.int32))


@pytest.mark.parametrize("edge", [False, True])
def test_ray_keys(oracle, ctx, edge):
    rays = scenes.ray_key_edge_rays(50000, seed=27) if edge else scenes.random_rays(50000, seed=27)
    rt = torch.from_numpy(rays).cuda()
    for bits in (30, 60, 64):
        ka = hx.ray_keys(ctx, rt, "aila", bits).cpu().numpy().view(np.uint64)
        assert np.array_equal(ka, oracle.ray_keys_aila64(rays, bits))
        for rl in (0.25, 2.0):
            kt = hx.ray_keys(ctx, rt, "twopoint", bits, rl).cpu().numpy().view(np.uint64)
            assert np.array_equal(kt, oracle.ray_keys_twopoint64(rays, bits, rl))


def test_trace_single_triangle_and_dead_rays(oracle, ctx):
    v = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0]], np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    rays = np.zeros((4, 8), np.float32)
    rays[:, 0:2] = [[0.2, 0.2], [0.9, 0.9], [0.1, 0.1], [0.3, 0.3]]
    rays[:, 6] = 1.0
    rays[:, 7] = [3.0, 3.0, -1.0, 0.5]
    scene, bvh, res, ref = _check_against_oracle(oracle, ctx, v, t, rays)
    assert list(res[:, 0]) == [0, -1, -1, -1]


@pytest.mark.parametrize("name,value", [("trace_lds_stack", 0), ("trace_lds_stack", 32), ("trace_lds_stack", 64),
                                        ("trace_fetch", 1), ("trace_fetch", 65), ("trace_blocks_per_cu", 1),
                                        ("trace_blocks_per_cu", 3), ("trace_postpone", 2), ("trace_postpone", 4),
                                        ("trace_woop_ahead", 2)])
def test_trace_tunables_keep_results(oracle, ctx, name, value):
    """Every launch-shape choice the autotuner can make gives the same hits."""
    old = ctx.get_tuning(name)
    ctx.set_tuning(name, value)
    try:
        v, t = scenes.gaussian_mixture_soup(40000, seed=31)
        _check_against_oracle(oracle, ctx, v, t, scenes.random_rays(30000, seed=8))
    finally:
        ctx.set_tuning(name, old)


def test_ray_set_launch_applies_per_call(oracle, ctx):
    """RayBuffer.launch holds for that ray set's traces only; results are the same."""
    v, t = scenes.gaussian_mixture_soup(20000, seed=32)
    bvh = hx.LBVHBuilder(ctx).build(hx.Scene(v, t))
    rays = scenes.random_rays(20000, seed=9)
    plain = hx.RayBuffer(rays)
    tuned = hx.RayBuffer(rays)
    tuned.launch = {"trace_fetch": 8, "trace_lds_stack": 32}
    before = (ctx.get_tuning("trace_fetch"), ctx.get_tuning("trace_lds_stack"))
    tr = hx.HipTracer(ctx, bvh)
    a = tr.trace(plain).cpu().numpy()
    b = tr.trace_sort(tuned, method="aila", bits=60).cpu().numpy()
    assert (ctx.get_tuning("trace_fetch"), ctx.get_tuning("trace_lds_stack")) == before
    assert np.array_equal(a, b)


def test_trace_view_on_external_buffers(oracle, ctx):
    """TRACE_FUNC's argument list over Bin BVH arrays the caller holds (copies
    of a build's buffers in fresh tensors): the same hits as hx_trace, and with
    traceSort's ray indices the same hits and order semantics."""
    v, t = scenes.gaussian_mixture_soup(30000, seed=61)
    rays = scenes.random_rays(25000, seed=62)
    scene, bvh = _build(ctx, v, t)
    plain, _, _ = _trace(ctx, bvh, rays)
    nodes = torch.from_numpy(bvh.nodes_host()).cuda()
    woop = torch.from_numpy(bvh.woop_host()).cuda()
    tri = torch.from_numpy(bvh.tri_indices_host()).cuda()
    del bvh  # the view does not depend on the hx_bvh
    rt = torch.from_numpy(rays).cuda()
    res = hx.trace_view(ctx, nodes, woop, tri, rt).cpu().numpy()
    assert np.array_equal(res[:, :2], plain[:, :2])
    h = plain[:, 0] >= 0
    assert np.array_equal(res[h], plain[h])
    ref, _ = oracle.trace(nodes.cpu().numpy(), woop.cpu().numpy(), tri.cpu().numpy(), rays)
    assert np.array_equal(res[:, :2], ref[:, :2])
    order = np.argsort(oracle.ray_keys_aila64(rays, 60), kind="stable").astype(np.int32)
    stats = torch.zeros((rays.shape[0], 2), dtype=torch.int32, device="cuda")
    res_o = hx.trace_view(ctx, nodes, woop, tri, rt, stats=stats,
                          indices=torch.from_numpy(order).cuda()).cpu().numpy()
    assert np.array_equal(res_o[:, :2], plain[:, :2])
    assert np.all(stats.cpu().numpy()[rays[:, 7] > 0, 0] >= 1)
    any_res = hx.trace_view(ctx, nodes, woop, tri, rt, any_hit=True).cpu().numpy()
    assert np.array_equal(any_res[:, 0] >= 0, plain[:, 0] >= 0)


def test_trace_view_rejects_bad_arguments(ctx):
    from hippie_amd import _lib
    L = _lib.lib()
    x = torch.zeros(64, dtype=torch.int32, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    assert L.hx_trace_view(ctx.handle, None, 2, p, p, 1, p, 1, 0, p, None, None) == _lib.HX_ERR_INVALID_ARG
    assert L.hx_trace_view(ctx.handle, p, 1, p, p, 1, p, 1, 0, p, None, None) == _lib.HX_ERR_INVALID_ARG
    assert L.hx_trace_view(ctx.handle, p, 2, p, p, 1, p, 1 << 30, 0, p, None, None) == _lib.HX_ERR_INVALID_ARG


def test_trace_sort_timed(oracle, ctx):
    """traceSort(rays, sortTime, traceTime): both halves timed, same hits."""
    v, t = scenes.uniform_soup(40000, seed=63)
    rays = scenes.config1_rays(256, 256)
    scene, bvh = _build(ctx, v, t)
    plain, _, _ = _trace(ctx, bvh, rays)
    rb = hx.RayBuffer(rays)
    ts, tt = hx.HipTracer(ctx, bvh).trace_sort_timed(rb, "aila", 60)
    assert ts > 0.0 and tt > 0.0
    assert np.array_equal(rb.results.cpu().numpy()[:, :2], plain[:, :2])
    order = rb.indices.cpu().numpy()
    assert np.array_equal(order, np.argsort(oracle.ray_keys_aila64(rays, 60), kind="stable").astype(np.int32))
