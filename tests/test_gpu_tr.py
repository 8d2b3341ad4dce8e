"""GPU parity of the TR build (libhx through the C ABI) against the CPU oracle.

The GPU runs the reference's subset DP with its partition order and its
4-lane and 32-lane shuffle-down reductions in 32-lane halves of each wave64,
so the restructured trees agree exactly: the Bin tree is compared in
canonical DFS form, the number of kept treelets must match, the Woop rows bit
for bit, and the FP64 SAH cost within 1e-9 relative.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _check(oracle, ctx, v, t, iters=2, adaptive=True, max_leaf=8, bits=60):
    ref = oracle.build_tr(v, t, bits=bits, iterations=iters, adaptive=adaptive, max_leaf_size=max_leaf)
    scene = hx.Scene(v, t)
    b = hx.TRBuilder(ctx, morton_bits=bits, tr_iterations=iters, adaptive_leaf_size=int(adaptive),
                     max_leaf_size=max_leaf)
    bvh = b.build(scene)
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    if t.shape[0] > 1:
        assert bvh.info("atr_accepted") == ref["accepted"], "treelets kept"
    assert (bvh.n_interior, bvh.n_leaves) == (ref["n_interior"], ref["n_leaves"]), "node counts"
    assert oracle.validate_bin(nodes, bvh.n_interior, bvh.n_leaves, tri, v, t) == 0, "validate"
    cg = oracle.bin_canonical(nodes, bvh.n_interior, bvh.n_leaves, tri)
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co), "canonical Bin tree"
    assert np.array_equal(bvh.woop_host().view(np.uint32), oracle.woopify(v, t, tri).view(np.uint32))
    sb = bvh.scene_box()
    d = sb[3:6] - sb[0:3]
    area = float(np.float32(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]))
    c_ref = oracle.cost(ref["nodes"], area)
    assert abs(bvh.get_cost() - c_ref) <= 1e-9 * abs(c_ref), "SAH cost"
    return ref, scene, bvh


@pytest.mark.parametrize("iters", [0, 1, 2, 4])
def test_tr_iterations(oracle, ctx, iters):
    v, t = scenes.gaussian_mixture_soup(30000, seed=31)
    ref, _, _ = _check(oracle, ctx, v, t, iters=iters)
    assert (ref["accepted"] > 0) == (iters > 0)


def test_tr_uniform_30bit(oracle, ctx):
    v, t = scenes.uniform_soup(40000, seed=32)
    _check(oracle, ctx, v, t, bits=30)


def test_tr_grid_ties(oracle, ctx):
    """A regular grid gives many equal subset costs: the reference's partition
    order and shuffle trees decide."""
    v, t = scenes.merge(scenes.ground_grid(96), scenes.gaussian_mixture_soup(8000, seed=33))
    _check(oracle, ctx, v, t)


def test_tr_max_leaf_size(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(20000, seed=34)
    _check(oracle, ctx, v, t, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 6, 7, 8, 13, 14, 15, 64, 513, 4097])
def test_tr_small_sizes(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n + 5)
    _check(oracle, ctx, v, t, iters=3)


def test_tr_beats_lbvh_traces_and_refits(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(60000, seed=36)
    ref, scene, bvh = _check(oracle, ctx, v, t)
    lb = hx.LBVHBuilder(ctx, morton_bits=60).build(scene)
    assert bvh.get_cost() < lb.get_cost()
    rays = scenes.random_rays(20000, seed=9)
    res = hx.HipTracer(ctx, bvh).trace(hx.RayBuffer(rays)).cpu().numpy()
    exp, overflow = oracle.trace(bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host(), rays)
    assert overflow == 0
    assert np.array_equal(res[:, 0], exp[:, 0]) and np.array_equal(res[:, 1], exp[:, 1])
    brute = oracle.trace_brute(v, t, rays[:4000])
    assert np.array_equal(res[:4000, :2], brute[:, :2])
    rng = np.random.default_rng(5)
    v2 = (v + rng.normal(0.0, 0.002, v.shape)).astype(np.float32)
    nodes0 = bvh.nodes_host()
    bvh.update(hx.Scene(v2, t))
    exp_nodes = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, bvh.tri_indices_host(), v2, t)
    assert np.array_equal(bvh.nodes_host(), exp_nodes), "refit nodes"


def test_tr_rebuild_is_bit_identical_and_env(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(25000, seed=37)
    p = hx.BuildParams.from_env({"Bvh.method": "tr", "Bvh.trIterations": "3", "Bvh.trMortonCodeBits": "48"})
    b = hx.TRBuilder(ctx, params=p)
    assert b.params.c.method == hx._lib.HX_METHOD_TR
    assert b.params.c.tr_iterations == 3 and b.params.c.morton_bits == 48
    s = hx.Scene(v, t)
    bvh = b.build(s)
    first = bvh.nodes_host()
    b.rebuild(bvh, s)
    assert np.array_equal(bvh.nodes_host(), first)
    ref = oracle.build_tr(v, t, bits=48, iterations=3)
    assert np.array_equal(oracle.bin_canonical(first, bvh.n_interior, bvh.n_leaves, bvh.tri_indices_host()),
                          oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"]))


@pytest.mark.parametrize("layout,N", [("quad", 4), ("oct", 8)])
def test_tr_wide_layouts(oracle, ctx, layout, N):
    v, t = scenes.uniform_soup(8000, seed=38)
    wb = hx.TRBuilder(ctx, layout=layout, keep_debug=1).build(hx.Scene(v, t))
    assert wb.layout == N
    d = wb.debug_host()
    n = wb.view().n_references
    roots = np.nonzero(d["parent"][:n - 1] == -1)[0]
    assert len(roots) == 1
    exp, ni = oracle.collapse_wide(N, n, int(roots[0]), d["left"], d["right"], d["node_min"], d["node_max"],
                                   3.0, 2.0, 0)
    assert np.array_equal(wb.nodes_host(), exp)
    rays = scenes.config1_rays(64, 64)
    got = hx.HipTracer(ctx, wb).trace(hx.RayBuffer(rays)).cpu().numpy()
    brute = oracle.trace_brute(v, t, rays)
    assert np.array_equal(got[:, :2], brute[:, :2])
