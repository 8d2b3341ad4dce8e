"""GPU parity of the HLBVH build (libhx through the C ABI) against the CPU oracle.

Clusters, the binned-SAH upper levels and the Karras lower levels are all
deterministic on both sides, so the collapsed Bin trees agree in canonical
form; Woop rows bit for bit; FP64 SAH cost within 1e-9 relative.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _check(oracle, ctx, v, t, sah_bits=15, bits=60, adaptive=True, max_leaf=8):
    ref = oracle.build_hlbvh(v, t, bits=bits, sah_bits=sah_bits, adaptive=adaptive, max_leaf_size=max_leaf)
    scene = hx.Scene(v, t)
    bvh = hx.HLBVHBuilder(ctx, morton_bits=bits, hlbvh_sah_bits=sah_bits, adaptive_leaf_size=int(adaptive),
                          max_leaf_size=max_leaf).build(scene)
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    if t.shape[0] > 1:
        assert bvh.info("hlbvh_clusters") == ref["clusters"], "clusters"
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


@pytest.mark.parametrize("sah_bits", [6, 9, 12, 15])
def test_hlbvh_mixture(oracle, ctx, sah_bits):
    v, t = scenes.gaussian_mixture_soup(60000, seed=51)
    _check(oracle, ctx, v, t, sah_bits=sah_bits)


def test_hlbvh_uniform_30bit_keys(oracle, ctx):
    v, t = scenes.uniform_soup(50000, seed=52)
    _check(oracle, ctx, v, t, sah_bits=15, bits=30)


def test_hlbvh_grid_median_splits(oracle, ctx):
    """A grid puts many clusters in one bin per axis: median splits by Morton order."""
    v, t = scenes.merge(scenes.ground_grid(128), scenes.gaussian_mixture_soup(4000, seed=53))
    _check(oracle, ctx, v, t, sah_bits=15)


def test_hlbvh_max_leaf_size(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(30000, seed=54)
    _check(oracle, ctx, v, t, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 513, 4097])
def test_hlbvh_small_sizes(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n + 7)
    _check(oracle, ctx, v, t, sah_bits=6)


def test_hlbvh_trace_and_refit(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(50000, seed=55)
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
