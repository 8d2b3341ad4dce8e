"""The CPU SBVH (oracle/hx_oracle_sbvh.cpp), the GPU builders' quality
reference (SBVHBuilder.cpp:14-401 + BVHConverter::convertAdaptive).

Independent checks:
- validity: every triangle is referenced, every leaf box contains its
  references' pieces (validate_bin_refs), and the closest hits through the
  SBVH equal a BVH-free brute force, also on sliver soups where references are
  clipped and duplicated;
- the root's object split equals a brute-force sweep restated here (every
  axis, every split position of the references sorted by box min + max, SAH
  in float32, ties to the more balanced split) when spatial splits are off;
- spatial splits: off (alpha huge) leaves one reference per triangle; on,
  they duplicate references and lower the SAH on long thin triangles;
- quality: the SBVH's SAH is below the LBVH's and PLOC's on clustered and
  uniform soups.
"""
import numpy as np
import pytest

from hippie_amd import scenes


def _area(mn, mx):
    return 2.0 * ((mx[0] - mn[0]) * (mx[1] - mn[1]) + (mx[1] - mn[1]) * (mx[2] - mn[2]) +
                  (mx[2] - mn[2]) * (mx[0] - mn[0]))


def _scene_area(oracle, v, t):
    sb = oracle.scene_box(v, t)
    return float(_area(sb[:3].astype(np.float64), sb[3:].astype(np.float64)))


def _sliver_scene(n, seed):
    rng = np.random.default_rng(seed)
    return scenes.sliver_soup(n, rng)


@pytest.mark.parametrize("kind,n,seed", [("mix", 3000, 1), ("uniform", 3000, 2), ("sliver", 2000, 3),
                                         ("grid", 0, 4)])
def test_sbvh_valid_and_traces_like_brute_force(oracle, kind, n, seed):
    if kind == "mix":
        v, t = scenes.gaussian_mixture_soup(n, seed=seed)
    elif kind == "uniform":
        v, t = scenes.uniform_soup(n, seed=seed)
    elif kind == "sliver":
        v, t = _sliver_scene(n, seed)
    else:
        v, t = scenes.merge(scenes.ground_grid(24), scenes.gaussian_mixture_soup(1500, seed=seed))
    s = oracle.build_sbvh(v, t)
    assert oracle.validate_bin_refs(s["nodes"], s["n_interior"], s["n_leaves"], s["tri_idx"], v, t) == 0
    assert s["n_refs"] >= t.shape[0]
    rays = np.concatenate([scenes.random_rays(3000, seed=seed + 10), scenes.config1_rays(48, 48)])
    hit, overflow = oracle.trace(s["nodes"], s["woop"], s["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    # A reference's box comes from float-clipped points, so a hit within
    # rounding of a triangle's edge may go to its neighbour (as in the
    # reference): such rays must agree on t; all others on id and t.
    u = brute[:, 2].view(np.float32)
    w = brute[:, 3].view(np.float32)
    edge = (brute[:, 0] >= 0) & (np.minimum(np.minimum(u, w), 1.0 - u - w) < 1e-4)
    assert np.array_equal(hit[~edge, :2], brute[~edge, :2])
    assert np.array_equal(hit[edge, 0] >= 0, brute[edge, 0] >= 0)
    tb, th = brute[edge, 1].view(np.float32), hit[edge, 1].view(np.float32)
    assert np.all(np.abs(th - tb) <= 1e-6 * np.abs(tb))
    assert edge.sum() < 0.01 * len(rays)


def _subtree_tris(nodes, tri_idx, c):
    out = set()
    stack = [c]
    while stack:
        x = stack.pop()
        if x < 0:
            w = nodes[~x]
            out.update(int(i) for i in tri_idx[w[12]:w[13]])
        else:
            stack.extend([int(nodes[x, 12]), int(nodes[x, 13])])
    return frozenset(out)


def _brute_object_split(v, t, ct, ci):
    f32 = np.float32
    tri = v[t].astype(f32)                       # (n, 3, 3)
    mn, mx = tri.min(axis=1), tri.max(axis=1)
    n = t.shape[0]
    node_mn, node_mx = mn.min(axis=0), mx.max(axis=0)

    def area32(a, b):
        d = (b - a).astype(f32)
        if np.any(d < 0):
            return f32(0)
        return f32(2) * f32(f32(f32(d[0] * d[1]) + f32(d[1] * d[2])) + f32(d[2] * d[0]))

    node_sah = f32(area32(node_mn, node_mx) * f32(ct))
    best = (np.inf, np.inf, None)
    for axis in range(3):
        key = (mn[:, axis] + mx[:, axis]).astype(f32)
        order = sorted(range(n), key=lambda i: (key[i], i))
        for i in range(1, n):
            l, r = order[:i], order[i:]
            la = area32(mn[l].min(axis=0), mx[l].max(axis=0))
            ra = area32(mn[r].min(axis=0), mx[r].max(axis=0))
            sah = f32(f32(node_sah + f32(f32(la * f32(ci)) * f32(i))) + f32(f32(ra * f32(ci)) * f32(n - i)))
            tie = i * i + (n - i) * (n - i)
            if sah < best[0] or (sah == best[0] and tie < best[1]):
                best = (sah, tie, (frozenset(l), frozenset(r)))
    return best[2]


@pytest.mark.parametrize("seed", range(6))
def test_sbvh_root_object_split_is_the_brute_force_sweep(oracle, seed):
    v, t = scenes.gaussian_mixture_soup(14, components=3, seed=seed)
    s = oracle.build_sbvh(v, t, alpha=1e30, max_leaf_size=1)
    nodes = s["nodes"]
    a = _subtree_tris(nodes, s["tri_idx"], int(nodes[0, 12]))
    b = _subtree_tris(nodes, s["tri_idx"], int(nodes[0, 13]))
    assert {a, b} == set(_brute_object_split(v, t, 3.0, 2.0))


def test_sbvh_spatial_splits(oracle):
    v, t = _sliver_scene(3000, 7)
    area = _scene_area(oracle, v, t)
    off = oracle.build_sbvh(v, t, alpha=1e30)
    on = oracle.build_sbvh(v, t, alpha=1e-5)
    assert off["n_refs"] == t.shape[0]
    assert on["n_refs"] > t.shape[0]
    assert oracle.cost(on["nodes"], area) < oracle.cost(off["nodes"], area)


@pytest.mark.parametrize("kind", ["mix", "uniform"])
def test_sbvh_is_the_quality_reference(oracle, kind):
    v, t = (scenes.gaussian_mixture_soup(20000, seed=5) if kind == "mix" else scenes.uniform_soup(20000, seed=6))
    area = _scene_area(oracle, v, t)
    s = oracle.cost(oracle.build_sbvh(v, t)["nodes"], area)
    lb = oracle.cost(oracle.build_lbvh(v, t)["nodes"], area)
    pl = oracle.cost(oracle.build_ploc(v, t, radius=25)["nodes"], area)
    assert s < lb and s < pl
