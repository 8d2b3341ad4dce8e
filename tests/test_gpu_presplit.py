"""GPU parity of presplit builds (libhx through the C ABI) against the CPU oracle.

The presplitter is deterministic on both sides (exact fixed-point priority
sum, integer budget sums, each reference's slot fixed by its triangle's split
tree), so the reference count, D, and every builder's tree over the
references agree exactly: canonical Bin tree (child boxes are the clipped
reference boxes), Woop rows bit for bit, SAH cost to 1e-9 relative.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu

BUILDERS = {
    "lbvh": (hx.LBVHBuilder, {}),
    "ploc": (hx.PLOCBuilder, {"ploc_radius": 25}),
    "atr": (hx.ATRBuilder, {"atr_treelet_size": 9, "atr_iterations": 2}),
    "hlbvh": (hx.HLBVHBuilder, {"hlbvh_sah_bits": 12}),
    "insertion": (hx.InsertionBuilder, {"insertion_mod": 1}),
}
ORACLE_KW = {"lbvh": {}, "ploc": {"radius": 25}, "atr": {"treelet_size": 9, "iterations": 2},
             "hlbvh": {"sah_bits": 12}, "insertion": {"mod": 1}}


def _scene(n_soup=20000, grid=8, seed=81):
    return scenes.merge(scenes.gaussian_mixture_soup(n_soup, seed=seed), scenes.ground_grid(grid))


def _check(oracle, ctx, v, t, name, beta=0.3, bits=60, adaptive=True, max_leaf=8):
    cls, kw = BUILDERS[name]
    ref = getattr(oracle, "build_" + name)(v, t, bits=bits, adaptive=adaptive, max_leaf_size=max_leaf,
                                           presplit_beta=beta, **ORACLE_KW[name])
    ps = oracle.presplit(v, t, beta)
    scene = hx.Scene(v, t)
    bvh = cls(ctx, morton_bits=bits, adaptive_leaf_size=int(adaptive), max_leaf_size=max_leaf, presplitting=1,
              presplitter_beta=beta, **kw).build(scene)
    assert bvh.info("references") == ps["ref_tri"].shape[0] == ref["tri_idx"].shape[0], "reference count"
    # D is NaN when the budget is 0 (0 * Smax / 0 in the bisection): compare bits.
    d_bits = np.array([bvh.info("presplit_d_bits")], np.uint32)
    assert d_bits[0] == np.array([ps["D"]], np.float32).view(np.uint32)[0], "D"
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    assert (bvh.n_interior, bvh.n_leaves) == (ref["n_interior"], ref["n_leaves"]), "node counts"
    assert oracle.validate_bin_refs(nodes, bvh.n_interior, bvh.n_leaves, tri, v, t) == 0, "validate"
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


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_presplit_builders(oracle, ctx, name):
    v, t = _scene()
    _check(oracle, ctx, v, t, name)


@pytest.mark.parametrize("beta", [0.0, 0.05, 1.0, 4.0])
def test_presplit_betas(oracle, ctx, beta):
    v, t = _scene(seed=82)
    ref, scene, bvh = _check(oracle, ctx, v, t, "lbvh", beta=beta)
    if beta == 0.0:
        assert bvh.info("references") == t.shape[0]


def test_presplit_30_bits_and_leaf_size(oracle, ctx):
    v, t = _scene(seed=83)
    _check(oracle, ctx, v, t, "lbvh", bits=30, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 17, 300])
def test_presplit_small(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n + 11)
    _check(oracle, ctx, v, t, "lbvh", beta=1.0)


def test_presplit_large(oracle, ctx):
    v, t = _scene(n_soup=300000, grid=32, seed=84)
    _check(oracle, ctx, v, t, "lbvh", beta=0.3)


def test_presplit_trace_refit_and_rebuild(oracle, ctx):
    v, t = _scene(seed=85)
    ref, scene, bvh = _check(oracle, ctx, v, t, "ploc")
    rays = scenes.random_rays(20000, seed=9)
    res = hx.HipTracer(ctx, bvh).trace(hx.RayBuffer(rays)).cpu().numpy()
    exp, overflow = oracle.trace(bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host(), rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    assert np.array_equal(res[:, 0], exp[:, 0]) and np.array_equal(res[:, 1], exp[:, 1])
    assert np.array_equal(res[:, 0], brute[:, 0]) and np.array_equal(res[:, 1], brute[:, 1])
    nodes0 = bvh.nodes_host().copy()
    b = hx.PLOCBuilder(ctx, morton_bits=60, ploc_radius=25, presplitting=1, presplitter_beta=0.3)
    b.rebuild(bvh, scene)
    assert np.array_equal(bvh.nodes_host(), nodes0), "rebuild"
    # HipBVH::update refits leaf boxes from whole triangles, as the reference does.
    rng = np.random.default_rng(5)
    v2 = (v + rng.normal(0.0, 0.002, v.shape)).astype(np.float32)
    bvh.update(hx.Scene(v2, t))
    exp_nodes = oracle.refit_final(nodes0, bvh.n_interior, bvh.n_leaves, bvh.tri_indices_host(), v2, t)
    assert np.array_equal(bvh.nodes_host(), exp_nodes), "refit nodes"


def test_presplit_rejects_negative_beta_and_batches(ctx):
    v, t = _scene(n_soup=1000, seed=86)
    with pytest.raises(Exception):
        hx.LBVHBuilder(ctx, presplitting=1, presplitter_beta=-0.5).build(hx.Scene(v, t))
