"""GPU parity of the PLOC build (libhx through the C ABI) against the CPU oracle.

The oracle numbers the nodes a round creates in cluster order, as the GPU
does, so the binary trees agree exactly; the collapsed Bin tree is compared in
canonical DFS form (layout-independent), the Woop rows bit for bit, and the
FP64 SAH cost within 1e-9 relative (the required bar is 0.1 %).
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _gpu_ploc(ctx, v, t, radius, adaptive=True, max_leaf=8, ploc_adaptive=False, max_radius=512, bits=60):
    scene = hx.Scene(v, t)
    b = hx.PLOCBuilder(ctx, morton_bits=bits, adaptive_leaf_size=int(adaptive), max_leaf_size=max_leaf,
                       ploc_max_radius=max_radius, ploc_radius=radius, ploc_adaptive=int(ploc_adaptive))
    return scene, b.build(scene)


def _check(oracle, ctx, v, t, radius, adaptive=True, max_leaf=8, ploc_adaptive=False, max_radius=512):
    ref = oracle.build_ploc(v, t, radius=radius, adaptive=adaptive, max_leaf_size=max_leaf,
                            max_radius=max_radius, ploc_adaptive=ploc_adaptive)
    scene, bvh = _gpu_ploc(ctx, v, t, radius, adaptive, max_leaf, ploc_adaptive, max_radius)
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    assert (bvh.n_interior, bvh.n_leaves) == (ref["n_interior"], ref["n_leaves"]), "node counts"
    assert oracle.validate_bin(nodes, bvh.n_interior, bvh.n_leaves, tri, v, t) == 0, "validate"
    cg = oracle.bin_canonical(nodes, bvh.n_interior, bvh.n_leaves, tri)
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co), "canonical Bin tree"
    woop = bvh.woop_host()
    assert np.array_equal(woop.view(np.uint32), oracle.woopify(v, t, tri).view(np.uint32)), "Woop rows"
    sb = bvh.scene_box()
    d = sb[3:6] - sb[0:3]
    area = float(np.float32(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]))
    c_ref = oracle.cost(ref["nodes"], area)
    assert abs(bvh.get_cost() - c_ref) <= 1e-9 * abs(c_ref), "SAH cost"
    return ref, scene, bvh


@pytest.mark.parametrize("radius", [1, 25, 200])
def test_ploc_uniform(oracle, ctx, radius):
    v, t = scenes.uniform_soup(30000, seed=11)
    _check(oracle, ctx, v, t, radius)


def test_ploc_mixture_and_grid(oracle, ctx):
    """Clustered geometry plus a regular grid (many equal union areas: ties)."""
    v, t = scenes.merge(scenes.gaussian_mixture_soup(40000, seed=4), scenes.ground_grid(64))
    _check(oracle, ctx, v, t, 25)


def test_ploc_wide_radius_global_window(oracle, ctx):
    """r > 1024 takes the kernel variant that reads the window from HBM."""
    v, t = scenes.uniform_soup(6000, seed=12)
    _check(oracle, ctx, v, t, 1100, max_radius=2048)


def test_ploc_adaptive_radius(oracle, ctx):
    v, t = scenes.uniform_soup(20000, seed=13)
    _check(oracle, ctx, v, t, 1, ploc_adaptive=True, max_radius=48)


def test_ploc_max_leaf_size(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(20000, seed=14)
    _check(oracle, ctx, v, t, 25, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 17, 64, 65, 513, 1000, 4097])
def test_ploc_small_sizes(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n)
    _check(oracle, ctx, v, t, 3)


def test_ploc_trace_matches_oracle(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(50000, seed=15)
    ref, scene, bvh = _check(oracle, ctx, v, t, 25)
    rays = scenes.random_rays(20000, seed=7)
    res = hx.HipTracer(ctx, bvh).trace(hx.RayBuffer(rays)).cpu().numpy()
    exp, overflow = oracle.trace(bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host(), rays)
    assert overflow == 0
    assert np.array_equal(res[:, 0], exp[:, 0]), "hit ids"
    assert np.array_equal(res[:, 1], exp[:, 1]), "hit t bits"


def test_ploc_refit(oracle, ctx):
    """Refit of a PLOC tree: same topology, boxes and Woop rows from new vertices."""
    v, t = scenes.uniform_soup(20000, seed=16)
    ref, scene, bvh = _check(oracle, ctx, v, t, 25)
    rng = np.random.default_rng(3)
    v2 = (v + rng.normal(0.0, 0.002, v.shape)).astype(np.float32)
    scene2 = hx.Scene(v2, t)
    nodes0 = bvh.nodes_host()
    bvh.update(scene2)
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    exp = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, tri, v2, t)
    assert np.array_equal(nodes, exp), "refit nodes"
    assert np.array_equal(bvh.woop_host().view(np.uint32), oracle.woopify(v2, t, tri).view(np.uint32))


def test_ploc_rebuild_reuses_handle(oracle, ctx):
    v, t = scenes.uniform_soup(10000, seed=17)
    scene, bvh = _gpu_ploc(ctx, v, t, 25)
    v2, t2 = scenes.uniform_soup(10000, seed=18)
    scene2 = hx.Scene(v2, t2)
    hx.PLOCBuilder(ctx, ploc_radius=25).rebuild(bvh, scene2)
    ref = oracle.build_ploc(v2, t2, radius=25)
    cg = oracle.bin_canonical(bvh.nodes_host(), bvh.n_interior, bvh.n_leaves, bvh.tri_indices_host())
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co)
