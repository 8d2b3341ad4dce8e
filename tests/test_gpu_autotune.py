"""The autotuner picks values from its candidate sets, caches them, and the
tuned context still produces the oracle's results."""
import json

import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes
from hippie_amd.autotune import BLOCKS_PER_CU, FETCH, LDS_STACK, SORT_DIGITS, Autotuner

pytestmark = pytest.mark.gpu


def test_autotune_sort_trace_and_cache(oracle, ctx, tmp_path):
    saved = {k: ctx.get_tuning(k) for k in ("sort_digit_bits", "trace_lds_stack", "trace_fetch",
                                             "trace_blocks_per_cu")}
    cache = tmp_path / "tune.json"
    try:
        at = Autotuner(ctx, cache_path=str(cache), repeats=2)
        s = at.tune_sort(200000, bits=60)
        assert s["sort_digit_bits"] in SORT_DIGITS
        v, t = scenes.gaussian_mixture_soup(50000, seed=41)
        scene = hx.Scene(v, t)
        bvh = hx.LBVHBuilder(ctx).build(scene)
        rays = hx.RayBuffer(scenes.random_rays(60000, seed=3))
        before = {k: ctx.get_tuning(k) for k in ("trace_lds_stack", "trace_fetch", "trace_blocks_per_cu")}
        tr = at.tune_trace(bvh, rays)
        assert tr["trace_lds_stack"] in LDS_STACK and tr["trace_fetch"] in FETCH
        assert tr["trace_blocks_per_cu"] in BLOCKS_PER_CU
        # the choice belongs to this ray set; the context keeps its values
        assert rays.launch == tr
        assert {k: ctx.get_tuning(k) for k in before} == before
        tl = at.tune_treelet(scene, rays, sizes=(5, 9))
        assert tl["atr_treelet_size"] in (5, 9)
        stored = json.loads(cache.read_text())
        assert len(stored) == 3
        # A second tuner reads the cache instead of timing again.
        at2 = Autotuner(ctx, cache_path=str(cache))
        rays2 = hx.RayBuffer(scenes.random_rays(60000, seed=3))
        assert at2.tune_trace(bvh, rays2) == tr and not at2.history and rays2.launch == tr
        # Tuned context: the build and the trace still match the oracle.
        bvh2 = hx.LBVHBuilder(ctx).build(scene)
        ref = oracle.build_lbvh(v, t)
        cg = oracle.bin_canonical(bvh2.nodes_host(), bvh2.n_interior, bvh2.n_leaves, bvh2.tri_indices_host())
        co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
        assert np.array_equal(cg, co)
        res = hx.HipTracer(ctx, bvh2).trace(rays).cpu().numpy()
        exp, _ = oracle.trace(bvh2.nodes_host(), bvh2.woop_host(), bvh2.tri_indices_host(),
                              rays.rays.cpu().numpy())
        assert np.array_equal(res[:, :2], exp[:, :2])
    finally:
        for k, val in saved.items():
            ctx.set_tuning(k, val)


def test_autotune_batch_trace_keeps_hits(ctx):
    """The batched trace's launch shape belongs to the batch; hits do not change."""
    meshes = [scenes.uniform_soup(n, seed=n) for n in (500, 3000, 12000)]
    mb = hx.MeshBatch.from_meshes(meshes)
    hb = hx.HipBatch(ctx, mb, morton_bits=60)
    import torch
    rays = torch.from_numpy(np.concatenate([scenes.random_rays(4000, seed=s) for s in range(3)])).cuda()
    owner = torch.repeat_interleave(torch.arange(3, dtype=torch.int32), 4000).cuda()
    before = hb.trace(rays, owner).clone()
    saved = {k: ctx.get_tuning(k) for k in ("trace_lds_stack", "trace_fetch", "trace_blocks_per_cu", "trace_postpone",
                                         "trace_woop_ahead")}
    best = Autotuner(ctx, repeats=1).tune_batch_trace(hb, rays, owner)
    assert hb.launch == best
    assert {k: ctx.get_tuning(k) for k in saved} == saved
    assert torch.equal(hb.trace(rays, owner), before)
