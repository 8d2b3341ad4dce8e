"""GPU parity of the ATRBVH build (libhx through the C ABI) against the CPU oracle.

The GPU reproduces the reference's warp-32 reductions in 32-lane halves of
each wave64, so the restructured trees agree exactly: the Bin tree is compared
in canonical DFS form, the number of kept treelets must match, the Woop rows
bit for bit, and the FP64 SAH cost within 1e-9 relative (the required bar is
0.1 %).
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _check(oracle, ctx, v, t, ts=9, iters=2, adaptive=True, max_leaf=8, bits=60):
    ref = oracle.build_atr(v, t, bits=bits, treelet_size=ts, iterations=iters, adaptive=adaptive,
                           max_leaf_size=max_leaf)
    scene = hx.Scene(v, t)
    b = hx.ATRBuilder(ctx, morton_bits=bits, atr_treelet_size=ts, atr_iterations=iters,
                      adaptive_leaf_size=int(adaptive), max_leaf_size=max_leaf)
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


@pytest.mark.parametrize("ts", [4, 9, 20, 30])
def test_atr_treelet_sizes(oracle, ctx, ts):
    v, t = scenes.gaussian_mixture_soup(30000, seed=21)
    _check(oracle, ctx, v, t, ts=ts)


def test_atr_uniform_three_iterations(oracle, ctx):
    v, t = scenes.uniform_soup(40000, seed=22)
    _check(oracle, ctx, v, t, ts=9, iters=3)


def test_atr_grid_ties(oracle, ctx):
    """A regular grid gives many equal union areas: the warp-32 tie rules decide."""
    v, t = scenes.merge(scenes.ground_grid(96), scenes.gaussian_mixture_soup(8000, seed=23))
    _check(oracle, ctx, v, t, ts=9)


def test_atr_zero_iterations(oracle, ctx):
    v, t = scenes.uniform_soup(10000, seed=24)
    _check(oracle, ctx, v, t, iters=0)


def test_atr_max_leaf_size(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(20000, seed=25)
    _check(oracle, ctx, v, t, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 10, 64, 513, 4097])
def test_atr_small_sizes(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n + 3)
    _check(oracle, ctx, v, t, ts=4, iters=3)


def test_atr_trace_and_refit(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(50000, seed=26)
    ref, scene, bvh = _check(oracle, ctx, v, t)
    rays = scenes.random_rays(20000, seed=9)
    res = hx.HipTracer(ctx, bvh).trace(hx.RayBuffer(rays)).cpu().numpy()
    exp, overflow = oracle.trace(bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host(), rays)
    assert overflow == 0
    assert np.array_equal(res[:, 0], exp[:, 0]) and np.array_equal(res[:, 1], exp[:, 1])
    rng = np.random.default_rng(5)
    v2 = (v + rng.normal(0.0, 0.002, v.shape)).astype(np.float32)
    nodes0 = bvh.nodes_host()
    bvh.update(hx.Scene(v2, t))
    exp_nodes = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, bvh.tri_indices_host(), v2, t)
    assert np.array_equal(bvh.nodes_host(), exp_nodes), "refit nodes"
