"""Pinning the oracle's PLOC (PLOCBuilder.cpp:21-208, PLOCBuilderKernels.cu:11-389).

How each property is fixed independently of the oracle's own code:
- with a radius that covers every cluster, PLOC is agglomerative clustering by
  mutual nearest neighbours; the union-area distance is reducible
  (area(A u B u C) >= area(A u C)), so the hierarchy equals the one of greedy
  agglomerative clustering that always merges the globally closest pair
  (Walter et al. 2008) -- computed here by brute force in Python;
- one round at radius r, recomputed by brute force (nearest neighbour, mutual
  pairs, labels counting down from m-2 in cluster order);
- adaptive radius: radius k, max k+1 is non-adaptive radius k+1 from round 1;
- structure: a binary tree over all leaves, root 0, every internal box the
  exact union of its children, Bin output valid, trace equal to brute force.
"""
import numpy as np
import pytest

from hippie_amd import scenes


def _sorted_leaves(oracle, v, t, bits=60):
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    keys = oracle.morton_codes(bmin, bmax, sb, bits)
    _, sv = oracle.sort_pairs(keys, np.arange(t.shape[0], dtype=np.int32), bits)
    return sv, bmin[sv], bmax[sv]


def _area(mn, mx):
    d = (mx[:3] - mn[:3]).astype(np.float32)
    return np.float32(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _leaf_sets(left, right, n):
    """Leaf set (as frozenset of sorted positions) of every internal node."""
    memo = {}

    def rec(x):
        if x >= n - 1:
            return frozenset([x - (n - 1)])
        if x not in memo:
            memo[x] = rec(left[x]) | rec(right[x])
        return memo[x]

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10 * n))
    try:
        return {rec(i) for i in range(n - 1)}
    finally:
        sys.setrecursionlimit(old)


def _greedy_agglomerative(lmin, lmax):
    """Merge the globally closest pair (by union area) until one cluster is left."""
    clusters = [(frozenset([i]), lmin[i, :3].copy(), lmax[i, :3].copy()) for i in range(lmin.shape[0])]
    out = set()
    while len(clusters) > 1:
        best = None
        for a in range(len(clusters)):
            for b in range(a + 1, len(clusters)):
                mn = np.minimum(clusters[a][1], clusters[b][1])
                mx = np.maximum(clusters[a][2], clusters[b][2])
                d = _area(mn, mx)
                if best is None or d < best[0]:
                    best = (d, a, b, mn, mx)
        _, a, b, mn, mx = best
        s = clusters[a][0] | clusters[b][0]
        out.add(s)
        clusters = [c for k, c in enumerate(clusters) if k not in (a, b)] + [(s, mn, mx)]
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ploc_full_radius_is_greedy_agglomerative(oracle, seed):
    v, t = scenes.uniform_soup(120, seed=seed)
    n = t.shape[0]
    _, lmin, lmax = _sorted_leaves(oracle, v, t)
    left, right, parent, nmin, nmax, rounds = oracle.ploc(lmin, lmax, radius=n, max_radius=n)
    assert rounds > 0
    assert _leaf_sets(left, right, n) == _greedy_agglomerative(lmin, lmax)


def _one_round(lmin, lmax, r):
    """Brute-force first round: neighbours, mutual pairs, new labels."""
    m = lmin.shape[0]
    nb = []
    for i in range(m):
        best, bd = -1, np.float32(3.402823466e38)
        for j in range(max(0, i - r), min(m, i + r + 1)):
            if j == i:
                continue
            d = _area(np.minimum(lmin[i, :3], lmin[j, :3]), np.maximum(lmax[i, :3], lmax[j, :3]))
            if d < bd:          # strict: the leftmost minimum wins
                best, bd = j, d
        nb.append(best)
    merges = []
    for i in range(m):
        j = nb[i]
        if nb[j] == i and i < j:
            merges.append((i, j))
    return {(m - 2 - k): (i, j) for k, (i, j) in enumerate(merges)}


@pytest.mark.parametrize("r", [1, 3, 25])
def test_ploc_first_round_brute_force(oracle, r):
    """Nodes created in the first round: label m-2-k for the k-th mutual pair."""
    v, t = scenes.gaussian_mixture_soup(400, seed=5)
    n = t.shape[0]
    _, lmin, lmax = _sorted_leaves(oracle, v, t)
    left, right, parent, _, _, _ = oracle.ploc(lmin, lmax, radius=r, max_radius=max(r, 1))
    first = _one_round(lmin, lmax, r)
    assert len(first) > 0
    for node, (i, j) in first.items():
        assert left[node] == n - 1 + i and right[node] == n - 1 + j


def test_ploc_adaptive_equals_fixed_when_capped(oracle):
    v, t = scenes.uniform_soup(3000, seed=9)
    _, lmin, lmax = _sorted_leaves(oracle, v, t)
    a = oracle.ploc(lmin, lmax, radius=4, max_radius=5, adaptive=True)
    b = oracle.ploc(lmin, lmax, radius=5, max_radius=5, adaptive=False)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
    c = oracle.ploc(lmin, lmax, radius=1, max_radius=64, adaptive=True)
    d = oracle.ploc(lmin, lmax, radius=1, max_radius=64, adaptive=False)
    assert not np.array_equal(c[0], d[0])


@pytest.mark.parametrize("r", [1, 25])
def test_ploc_tree_structure(oracle, r):
    v, t = scenes.uniform_soup(5000, seed=4)
    n = t.shape[0]
    _, lmin, lmax = _sorted_leaves(oracle, v, t)
    left, right, parent, nmin, nmax, _ = oracle.ploc(lmin, lmax, radius=r)
    assert parent[0] == -1
    # every node except the root has exactly one parent that points back to it
    kids = np.concatenate([left[: n - 1], right[: n - 1]])
    assert np.array_equal(np.sort(kids), np.arange(1, 2 * n - 1))
    for i in range(n - 1):
        assert parent[left[i]] == i and parent[right[i]] == i
        mn = np.minimum(nmin[left[i], :3], nmin[right[i], :3])
        mx = np.maximum(nmax[left[i], :3], nmax[right[i], :3])
        assert np.array_equal(mn, nmin[i, :3]) and np.array_equal(mx, nmax[i, :3])


@pytest.mark.parametrize("adaptive", [True, False])
def test_build_ploc_valid_and_traces_like_brute_force(oracle, adaptive):
    v, t = scenes.uniform_soup(4000, seed=6)
    ref = oracle.build_ploc(v, t, radius=25, adaptive=adaptive, max_leaf_size=4)
    assert oracle.validate_bin(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
    rays = scenes.random_rays(2000, seed=2)
    hit, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    assert np.array_equal(hit[:, 0], brute[:, 0])
    assert np.array_equal(hit[:, 1], brute[:, 1])


def test_build_ploc_small_n(oracle):
    for n in (1, 2, 3, 7):
        v, t = scenes.uniform_soup(n, seed=n)
        ref = oracle.build_ploc(v, t, radius=1)
        assert oracle.validate_bin(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
