"""GPU parity of the insertion build (libhx through the C ABI) against the CPU oracle.

Every step of the GPU loop is deterministic (the locks take a maximum, the
kept moves write disjoint nodes, the cost is an exact fixed-point sum), so the
optimised tree agrees exactly: the Bin tree is compared in canonical DFS form,
the step and reinsertion counts must match, the Woop rows bit for bit, and
the FP64 SAH cost within 1e-9 relative.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _oracle_counts(oracle, v, t, bits, mod):
    n = t.shape[0]
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    keys = oracle.morton_codes(bmin, bmax, sb, bits)
    sk, sv = oracle.sort_pairs(keys, np.arange(n, dtype=np.int32), bits)
    left, right, parent = oracle.construct(sk)
    nmin, nmax = oracle.refit(sv, bmin, bmax, left, right, parent)
    *_, k, steps = oracle.insertion(left, right, parent, nmin, nmax, mod=mod)
    return k, steps


def _check(oracle, ctx, v, t, mod=1, adaptive=True, max_leaf=8, bits=60):
    ref = oracle.build_insertion(v, t, bits=bits, mod=mod, adaptive=adaptive, max_leaf_size=max_leaf)
    scene = hx.Scene(v, t)
    b = hx.InsertionBuilder(ctx, morton_bits=bits, insertion_mod=mod, adaptive_leaf_size=int(adaptive),
                            max_leaf_size=max_leaf)
    bvh = b.build(scene)
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    if t.shape[0] > 1:
        k, steps = _oracle_counts(oracle, v, t, bits, mod)
        assert bvh.info("insertion_steps") == steps == ref["steps"], "steps"
        assert bvh.info("insertion_reinserted") == k, "reinsertions"
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


def test_insertion_mixture(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(40000, seed=61)
    _check(oracle, ctx, v, t)


@pytest.mark.parametrize("mod", [2, 5, 64])
def test_insertion_mod(oracle, ctx, mod):
    v, t = scenes.uniform_soup(20000, seed=62)
    _check(oracle, ctx, v, t, mod=mod)


def test_insertion_grid_ties(oracle, ctx):
    """Equal area reductions on a regular grid: the lock keys' index part decides."""
    v, t = scenes.merge(scenes.ground_grid(80), scenes.gaussian_mixture_soup(6000, seed=63))
    _check(oracle, ctx, v, t)


def test_insertion_30_bit_codes(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(30000, seed=64)
    _check(oracle, ctx, v, t, bits=30)


def test_insertion_max_leaf_size(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(20000, seed=65)
    _check(oracle, ctx, v, t, adaptive=False, max_leaf=4)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 64, 513, 4097])
def test_insertion_small_sizes(oracle, ctx, n):
    v, t = scenes.uniform_soup(n, seed=n + 7)
    _check(oracle, ctx, v, t)


def test_insertion_large(oracle, ctx):
    v, t = scenes.merge(scenes.gaussian_mixture_soup(200000, seed=66), scenes.ground_grid(128))
    _check(oracle, ctx, v, t)


def test_insertion_trace_and_refit(oracle, ctx):
    v, t = scenes.gaussian_mixture_soup(50000, seed=67)
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


def test_insertion_rejects_bad_mod(ctx):
    with pytest.raises(ValueError):
        hx.InsertionBuilder(ctx).set_mod(0)
    with pytest.raises(NotImplementedError):
        hx.InsertionBuilder(ctx).set_sbvh(True)
