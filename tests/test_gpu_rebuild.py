"""Rebuilding reproduces the first build bit for bit, for every builder.

A rebuild reuses the BVH's device buffers; nothing a build reads may be left
over from the build before (tree kind, split positions, swapped triangle-id
buffers, climb counters).  Each builder is built, rebuilt twice, and the Bin
nodes, Woop rows and triangle indices compared with the first build's.
"""
import ctypes

import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import _lib, scenes

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


def test_rebuild_and_refit_reject_other_sizes(ctx):
    """A BVH's buffers are sized for its build's geometry: other triangle or
    vertex counts are rejected, by the host wrapper and by the C entry points."""
    v, t = scenes.gaussian_mixture_soup(5000, seed=74)
    scene = hx.Scene(v, t)
    b = hx.LBVHBuilder(ctx, morton_bits=60)
    bvh = b.build(scene)
    smaller = hx.Scene(v[:-3], t[:-1])
    with pytest.raises(ValueError):
        b.rebuild(bvh, smaller)
    with pytest.raises(ValueError):
        bvh.update(smaller)
    L = _lib.lib()
    vp, tp = ctypes.c_void_p(smaller.vertices.data_ptr()), ctypes.c_void_p(smaller.triangles.data_ptr())
    for fn in (L.hx_rebuild, L.hx_refit):
        assert fn(ctx.handle, bvh.handle, vp, smaller.n_vertices, tp, smaller.n_triangles) == _lib.HX_ERR_INVALID_ARG
        assert fn(ctx.handle, bvh.handle, vp, scene.n_vertices - 3, tp, scene.n_triangles) == _lib.HX_ERR_INVALID_ARG
    # the BVH is untouched and still refits with the right counts
    bvh.update(scene)
    assert bvh.validate() == 0
    batch = hx.HipBatch(ctx, hx.MeshBatch(v, t, [0, 2000, 5000]))
    with pytest.raises(ValueError):
        batch.refit(v[:-3])
    for fn in (L.hx_batch_refit, L.hx_batch_rebuild):
        assert fn(ctx.handle, batch.handle, vp, smaller.n_vertices, tp, smaller.n_triangles) == \
            _lib.HX_ERR_INVALID_ARG


def test_lbvh_rebuild_in_a_graph(ctx):
    """The 60-bit LBVH build of >= 2^18 triangles runs the split sort, whose
    fallback passes are gated on the device: nothing in a rebuild waits for the
    host, so a rebuild is captured in a HIP graph and its replays equal the
    eager build, for scenes whose sorts do and do not take the fallback."""
    v, t = scenes.config2_scene(n_mixture=300000, grid_cells=128, seed=75)
    for local_run in (64, 1):  # 1: every run of two equal high keys takes the fallback passes
        old = ctx.get_tuning("sort_local_run")
        ctx.set_tuning("sort_local_run", local_run)
        try:
            scene = hx.Scene(v, t)
            b = hx.LBVHBuilder(ctx, morton_bits=60)
            bvh = b.build(scene)
            eager = (bvh.nodes_host().copy(), bvh.woop_host().copy(), bvh.tri_indices_host().copy())
            torch.cuda.synchronize()
            s = torch.cuda.Stream()
            ctx.set_stream(s)
            try:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.stream(s):
                    b.rebuild(bvh, scene)  # warm-up on the capture stream
                    s.synchronize()
                    with torch.cuda.graph(g, stream=s):
                        b.rebuild(bvh, scene)
            finally:
                ctx.set_stream(torch.cuda.current_stream())
            for _ in range(3):
                scene.vertices.mul_(1.0)  # eager work between replays
                torch.cuda.synchronize()
                g.replay()
                torch.cuda.synchronize()
                assert np.array_equal(bvh.nodes_host(), eager[0]), local_run
                assert np.array_equal(bvh.woop_host().view(np.uint32), eager[1].view(np.uint32)), local_run
                assert np.array_equal(bvh.tri_indices_host(), eager[2]), local_run
        finally:
            ctx.set_tuning("sort_local_run", old)
