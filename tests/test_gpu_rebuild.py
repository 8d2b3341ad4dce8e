"""Rebuilding reproduces the first build bit for bit, for every builder.

A rebuild reuses the BVH's device buffers; nothing a build reads may be left
over from the build before (tree kind, split positions, swapped triangle-id
buffers, climb counters).  Each builder is built, rebuilt twice, and the Bin
nodes, Woop rows and triangle indices compared with the first build's.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes

pytestmark = pytest.mark.gpu

BUILDERS = {
    "lbvh": (hx.LBVHBuilder, {}),
    "ploc": (hx.PLOCBuilder, {"ploc_radius": 16}),
    "atr": (hx.ATRBuilder, {"atr_treelet_size": 9, "atr_iterations": 2}),
    "hlbvh": (hx.HLBVHBuilder, {"hlbvh_sah_bits": 12}),
    "insertion": (hx.InsertionBuilder, {"insertion_mod": 1}),
}


@pytest.mark.parametrize("name", sorted(BUILDERS))
def test_rebuild_is_bit_identical(ctx, name):
    cls, kw = BUILDERS[name]
    v, t = scenes.merge(scenes.gaussian_mixture_soup(60000, seed=71), scenes.ground_grid(64))
    scene = hx.Scene(v, t)
    b = cls(ctx, morton_bits=60, **kw)
    bvh = b.build(scene)
    first = (bvh.nodes_host().copy(), bvh.woop_host().copy(), bvh.tri_indices_host().copy(), bvh.get_cost())
    for _ in range(2):
        b.rebuild(bvh, scene)
        assert np.array_equal(bvh.nodes_host(), first[0]), name
        assert np.array_equal(bvh.woop_host().view(np.uint32), first[1].view(np.uint32)), name
        assert np.array_equal(bvh.tri_indices_host(), first[2]), name
        assert bvh.get_cost() == first[3], name


def test_rebuild_after_other_geometry(ctx):
    """A rebuild over different vertices equals a fresh build over them."""
    v, t = scenes.gaussian_mixture_soup(40000, seed=72)
    v2 = scenes.gaussian_mixture_soup(40000, seed=73)[0]
    for name, (cls, kw) in BUILDERS.items():
        b = cls(ctx, morton_bits=60, **kw)
        bvh = b.build(hx.Scene(v, t))
        b.rebuild(bvh, hx.Scene(v2, t))
        fresh = cls(ctx, morton_bits=60, **kw).build(hx.Scene(v2, t))
        assert np.array_equal(bvh.nodes_host(), fresh.nodes_host()), name
        assert np.array_equal(bvh.tri_indices_host(), fresh.tri_indices_host()), name
