"""Quad and Oct layouts (Bvh.layout quad / oct) on the GPU.

The device collapse of each builder's binary tree must equal the oracle's
restatement of the reference's host collapse (BVH::collapseAdaptiveWide /
collapseWide + BVHConverter's breadth-first records) on the same binary tree,
word for word; the wide traversal must return exactly the Bin traversal's hits
(id, t, u, v) and a BVH-free brute force's; refits keep the hits equal to the
brute force on the moved vertices; the SAH cost equals the oracle's."""
import numpy as np
import pytest
import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu

BUILDERS = {"lbvh": hx.LBVHBuilder, "ploc": hx.PLOCBuilder, "atr": hx.ATRBuilder, "hlbvh": hx.HLBVHBuilder,
            "insertion": hx.InsertionBuilder}


def _binary(bvh):
    d = bvh.debug_host()
    parent = d["parent"]
    n = bvh.view().n_references
    if n == 1:
        return n, 0, np.zeros(1, np.int32), np.zeros(1, np.int32), d["node_min"], d["node_max"]
    roots = np.nonzero(parent[:n - 1] == -1)[0]
    assert len(roots) == 1
    return n, int(roots[0]), d["left"], d["right"], d["node_min"], d["node_max"]


def _rays(v, t, w=96, h=64):
    return scenes.config1_rays(w, h)


def _trace(ctx, bvh, rays, sort=None):
    rb = hx.RayBuffer(rays)
    tr = hx.HipTracer(ctx, bvh)
    if sort:
        tr.trace_sort(rb, method=sort, bits=60)
    else:
        tr.trace(rb)
    return rb.results.cpu().numpy()


@pytest.mark.parametrize("layout,N", [("quad", 4), ("oct", 8)])
@pytest.mark.parametrize("method", ["lbvh", "ploc", "atr", "hlbvh", "insertion"])
def test_wide_collapse_matches_oracle_and_bin_hits(oracle, ctx, layout, N, method):
    v, t = scenes.uniform_soup(6000, seed=11)
    kw = dict(morton_bits=30) if method != "ploc" else dict(morton_bits=60, ploc_radius=8)
    wb = BUILDERS[method](ctx, layout=layout, keep_debug=1, **kw).build(hx.Scene(v, t))
    assert wb.layout == N
    nodes = wb.nodes_host()
    n, root, left, right, nmin, nmax = _binary(wb)
    exp, ni = oracle.collapse_wide(N, n, root, left, right, nmin, nmax, 3.0, 2.0, 0)
    assert wb.n_interior == ni and nodes.shape == exp.shape
    assert np.array_equal(nodes, exp), f"{np.argwhere(nodes != exp)[:5]}"
    rays = _rays(v, t)
    got = _trace(ctx, wb, rays)
    assert ctx.overflow_count() == 0
    bb = BUILDERS[method](ctx, **kw).build(hx.Scene(v, t))
    ref = _trace(ctx, bb, rays)
    assert np.array_equal(got, ref)
    brute = oracle.trace_brute(v, t, rays[:2048])
    assert np.array_equal(got[:2048, :2], brute[:, :2])
    cost = wb.get_cost()
    box = wb.scene_box()
    d = box[3:] - box[:3]
    area = np.float32(2.0) * np.float32(d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
    assert cost == pytest.approx(oracle.cost_wide(N, nodes, float(area), 3.0, 2.0), rel=1e-9)


@pytest.mark.parametrize("layout,N", [("quad", 4), ("oct", 8)])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 70000])
def test_wide_sizes_and_fixed_leaf_size(oracle, ctx, layout, N, n):
    v, t = scenes.uniform_soup(n, seed=n)
    for adaptive, mls in ((1, 8), (0, 3)):
        wb = hx.LBVHBuilder(ctx, layout=layout, keep_debug=1, morton_bits=60, adaptive_leaf_size=adaptive,
                            max_leaf_size=mls).build(hx.Scene(v, t))
        nn, root, left, right, nmin, nmax = _binary(wb)
        exp, ni = oracle.collapse_wide(N, nn, root, left, right, nmin, nmax, 3.0, 2.0, 0 if adaptive else mls)
        assert np.array_equal(wb.nodes_host(), exp), (adaptive, mls)
    rays = scenes.config1_rays(64, 64)
    got = _trace(ctx, wb, rays)
    brute = oracle.trace_brute(v, t, rays)
    assert np.array_equal(got[:, :2], brute[:, :2])


@pytest.mark.parametrize("layout", ["quad", "oct"])
def test_wide_sorted_trace_stats_and_any_hit(oracle, ctx, layout):
    v, t = scenes.merge(scenes.gaussian_mixture_soup(30000, seed=3), scenes.ground_grid(64))
    s = hx.Scene(v, t)
    wb = hx.LBVHBuilder(ctx, layout=layout, morton_bits=60).build(s)
    bb = hx.LBVHBuilder(ctx, morton_bits=60).build(s)
    prim = scenes.pinhole_rays(96, 96, (0.5, 0.6, 0.5), (0.2, -1.0, 0.3))
    pres = _trace(ctx, bb, prim)
    rays = np.concatenate([prim, scenes.secondary_rays(v, t, prim, pres)])
    ref = _trace(ctx, bb, rays)
    for key in ("aila", "twopoint"):
        assert np.array_equal(_trace(ctx, wb, rays, sort=key), ref), key
    # traversal statistics: the wide tree visits fewer nodes than the binary one
    rb = hx.RayBuffer(rays)
    hx.HipTracer(ctx, wb).trace_stats(rb)
    ws = rb.stats.cpu().numpy()
    rb2 = hx.RayBuffer(rays)
    hx.HipTracer(ctx, bb).trace_stats(rb2)
    bs = rb2.stats.cpu().numpy()
    assert np.array_equal(rb.results.cpu().numpy(), ref)
    assert 0 < ws[:, 0].sum() < bs[:, 0].sum()
    # any hit: every ray that hits gets some hit, every miss stays a miss
    ra = hx.RayBuffer(rays, closest_hit=False)
    hx.HipTracer(ctx, wb).trace(ra)
    ah = ra.results.cpu().numpy()
    assert np.array_equal(ah[:, 0] >= 0, ref[:, 0] >= 0)


@pytest.mark.parametrize("layout", ["quad", "oct"])
def test_wide_refit_and_rebuild(oracle, ctx, layout):
    v, t = scenes.uniform_soup(20000, seed=8)
    s = hx.Scene(v, t)
    b = hx.LBVHBuilder(ctx, layout=layout, morton_bits=60)
    wb = b.build(s)
    first = wb.nodes_host()
    b.rebuild(wb, s)
    assert np.array_equal(wb.nodes_host(), first)            # rebuilds are bit-identical
    rng = np.random.default_rng(4)
    v2 = (v + rng.normal(0, 0.002, v.shape)).astype(np.float32)
    wb.update(hx.Scene(v2, t))
    rays = scenes.config1_rays(96, 96)
    got = _trace(ctx, wb, rays)
    brute = oracle.trace_brute(v2, t, rays)
    assert np.array_equal(got[:, :2], brute[:, :2])
    assert np.array_equal(wb.woop_host().view(np.uint32),
                          oracle.woopify(v2, t, wb.tri_indices_host()).view(np.uint32))
