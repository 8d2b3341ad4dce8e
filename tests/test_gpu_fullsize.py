"""LBVH-60 at the benchmark's full size against the oracle's whole build.

The scene is the 16,777,216-triangle benchmark scene (8,388,608-triangle
Gaussian-mixture soup + 2048x2048 ground grid).  The device build must give
the oracle's tree (LBVHBuilder.cpp:183-263 through BVHCollapser.cpp:521-675) in
canonical form, pass the device-side HipBVH::validate, and trace a sample of
primary and diffuse rays (HipTracerKernels) to the oracle's hits, bit for bit;
the Morton-sorted trace of the same rays returns the same hits.  The
benchmark's PLOC (radius 25) and ATRBVH (treelet 9, 2 iterations) builds of
the same scene must equal the oracle's full-size builds in canonical form,
with the SAH cost within 0.1 % (the contract; the FP64 sums agree far
closer).  Every builder at this size is also checked through size-independent
properties: a valid tree and the same closest hits as the LBVH, since the
closest hit and its tie rule do not depend on the tree.
"""
import numpy as np
import pytest

import hippie_amd as hx
from hippie_amd import scenes


def _trace(ctx, bvh, rays, sort=None):
    rb = hx.RayBuffer(rays)
    tr = hx.HipTracer(ctx, bvh)
    if sort:
        tr.trace_sort(rb, method=sort, bits=60)
        return rb.results.cpu().numpy()
    return tr.trace(rb).cpu().numpy()


def _same_hits(a, b):
    """Hit ids and t bits of every ray, u and v bits of the hits."""
    h = a[:, 0] >= 0
    return np.array_equal(a[:, :2], b[:, :2]) and np.array_equal(a[h, 2:], b[h, 2:])


@pytest.fixture(scope="module")
def full_scene():
    return scenes.config2_scene()


@pytest.mark.gpu
def test_lbvh60_full_size_matches_oracle(oracle, ctx, full_scene):
    v, t = full_scene
    assert t.shape[0] == 16777216
    print(f"full-size LBVH-60: {t.shape[0]} triangles", flush=True)
    ref = oracle.build_lbvh(v, t, bits=60)
    bvh = hx.LBVHBuilder(ctx, morton_bits=60).build(hx.Scene(v, t))
    assert bvh.validate() == 0
    assert (bvh.n_interior, bvh.n_leaves) == (ref["n_interior"], ref["n_leaves"])
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    cg = oracle.bin_canonical(nodes, bvh.n_interior, bvh.n_leaves, tri)
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co), "canonical Bin tree"
    woop = bvh.woop_host()
    assert np.array_equal(woop.view(np.uint32), oracle.woopify(v, t, tri).view(np.uint32)), "Woop rows"
    del cg, co

    primary = scenes.config2_primary_rays(256, 256)
    hit_p = _trace(ctx, bvh, primary)
    ref_p, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], primary)
    assert overflow == 0
    assert _same_hits(hit_p, ref_p), "primary hits"
    assert np.count_nonzero(hit_p[:, 0] >= 0) > primary.shape[0] // 2
    diffuse = scenes.secondary_rays(v, t, primary, hit_p, seed=3)
    hit_d = _trace(ctx, bvh, diffuse)
    ref_d, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], diffuse)
    assert overflow == 0
    assert _same_hits(hit_d, ref_d), "diffuse hits"
    assert _same_hits(_trace(ctx, bvh, diffuse, sort="aila"), hit_d), "sorted trace"


@pytest.mark.gpu
@pytest.mark.parametrize("method,presplit", [("ploc", 0), ("atr", 0), ("hlbvh", 0), ("lbvh", 1)])
def test_builders_full_size_valid_and_same_hits(ctx, full_scene, method, presplit):
    v, t = full_scene
    scene = hx.Scene(v, t)
    rays = scenes.config2_primary_rays(256, 256)
    want = _trace(ctx, hx.LBVHBuilder(ctx, morton_bits=60).build(scene), rays)
    p = hx.BuildParams.from_env({"Bvh.method": method})
    p.presplitting, p.presplitter_beta = presplit, 0.3
    print(f"full-size {method} presplit={presplit}", flush=True)
    bvh = hx.BVHBuilder(ctx, params=p).build(scene)
    assert bvh.validate() == 0
    if presplit:
        assert bvh.view().n_references > t.shape[0]
    assert _same_hits(_trace(ctx, bvh, rays), want)


def _scene_area(bvh):
    sb = bvh.scene_box()
    d = sb[3:6] - sb[0:3]
    return float(np.float32(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]))


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["ploc", "atr"])
def test_config2_ploc_atr_full_size_match_oracle(oracle, ctx, full_scene, method):
    """Config 2's PLOC r = 25 and ATRBVH t = 9 x 2 at 16,777,216 triangles."""
    v, t = full_scene
    print(f"full-size {method} vs the oracle's full-size build", flush=True)
    if method == "ploc":
        bvh = hx.PLOCBuilder(ctx, morton_bits=60, ploc_radius=25).build(hx.Scene(v, t))
        ref = oracle.build_ploc(v, t, bits=60, radius=25)
    else:
        bvh = hx.ATRBuilder(ctx, morton_bits=60, atr_treelet_size=9, atr_iterations=2).build(hx.Scene(v, t))
        ref = oracle.build_atr(v, t, bits=60, treelet_size=9, iterations=2)
    assert bvh.validate() == 0
    c_ref = oracle.cost(ref["nodes"], _scene_area(bvh))
    c_gpu = bvh.get_cost()
    print(f"{method}: SAH {c_gpu:.6f} (oracle {c_ref:.6f}, rel {abs(c_gpu - c_ref) / c_ref:.2e})", flush=True)
    assert abs(c_gpu - c_ref) <= 1e-3 * abs(c_ref), "SAH within 0.1 %"
    assert (bvh.n_interior, bvh.n_leaves) == (ref["n_interior"], ref["n_leaves"])
    nodes, tri = bvh.nodes_host(), bvh.tri_indices_host()
    cg = oracle.bin_canonical(nodes, bvh.n_interior, bvh.n_leaves, tri)
    co = oracle.bin_canonical(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"])
    assert np.array_equal(cg, co), "canonical Bin tree"
    assert np.array_equal(bvh.woop_host().view(np.uint32), oracle.woopify(v, t, tri).view(np.uint32)), "Woop rows"
