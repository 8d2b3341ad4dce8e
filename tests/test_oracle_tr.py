"""Pinning the oracle's TR builder (TRBuilder.cpp:140-347, TRBuilderKernels.cu:11-476).

Independent checks:
- a tree of exactly 7 leaves with gamma = 7: only the root qualifies, its
  treelet holds every leaf, and the kept tree must reach the minimum, over all
  10395 binary trees on those leaves (enumerated here), of the cost
  C(S) = min(ct A(S) + C(left) + C(right), ci A(S) R); a rejected treelet
  means no tree beats the climb cost;
- on larger trees the whole optimisation pass is restated here with a plain
  minimum over every bipartition (no partition enumeration tricks, no lane
  reductions), in float32 in the same operation order, and the resulting
  trees must hold the same sets of leaves under every interior node.  The
  restated pass keeps the reference's cost bookkeeping, which later treelets
  of the same pass read when they reach into a restructured one: a new
  interior node's cost is costs[partition] of its parent's split, on the
  right side too (TRBuilderKernels.cu:264 and :295), and the treelet root
  keeps its climb cost.  That makes a split's orientation matter: every
  partition the reference enumerates leaves the subset's lowest leaf on the
  right, and so does the plain minimum here;
- the binary SAH never rises, the structure stays a tree with union boxes,
  iterations = 0 is the LBVH and the Bin output traces like brute force.
"""
import itertools
import math

import numpy as np
import pytest

from hippie_amd import scenes

f32 = np.float32


def _karras(oracle, v, t, bits=60):
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    keys = oracle.morton_codes(bmin, bmax, sb, bits)
    sk, sv = oracle.sort_pairs(keys, np.arange(t.shape[0], dtype=np.int32), bits)
    left, right, parent = oracle.construct(sk)
    nmin, nmax = oracle.refit(sv, bmin, bmax, left, right, parent)
    return left, right, parent, nmin, nmax


def _area32(mn, mx):
    x, y, z = (f32(mx[i]) - f32(mn[i]) for i in range(3))
    return f32(2.0) * f32(f32(f32(x * y) + f32(x * z)) + f32(y * z))


def _area64(mn, mx):
    d = np.asarray(mx[:3], np.float64) - np.asarray(mn[:3], np.float64)
    return 2.0 * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2])


def _leaf_sets(left, right, n):
    memo = {}

    def rec(x):
        if x >= n - 1:
            return frozenset([x - (n - 1)])
        if x not in memo:
            memo[x] = rec(left[x]) | rec(right[x])
        return memo[x]

    return {rec(i) for i in range(n - 1)}


def _all_trees(items):
    """Every full binary tree over the leaf tuple (nested pairs)."""
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for k in range(len(rest)):
        for other in itertools.combinations(rest, k):
            a = (first,) + other
            b = tuple(x for x in rest if x not in other)
            for ta in _all_trees(a):
                for tb in _all_trees(b):
                    yield (ta, tb)


def _tree_cost(tree, box, leaf_cost, R, ct, ci):
    """FP64 cost of a nested-pair tree with the collapse option at every
    interior node; returns (cost, leaf list)."""
    if not isinstance(tree, tuple):
        return leaf_cost[tree], [tree]
    ca, la = _tree_cost(tree[0], box, leaf_cost, R, ct, ci)
    cb, lb = _tree_cost(tree[1], box, leaf_cost, R, ct, ci)
    ls = la + lb
    mn = np.min([box[i][0] for i in ls], axis=0)
    mx = np.max([box[i][1] for i in ls], axis=0)
    A = _area64(mn, mx)
    return min(ct * A + ca + cb, ci * A * R), ls


def _nested(left, right, n, x):
    if x >= n - 1:
        return x - (n - 1)
    return (_nested(left, right, n, left[x]), _nested(left, right, n, right[x]))


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("ct,ci", [(3.0, 2.0), (1.0, 0.05)])
def test_tr_whole_tree_treelet_is_optimal(oracle, seed, ct, ci):
    n = 7
    v, tr = scenes.gaussian_mixture_soup(n, components=3, seed=seed)
    left, right, parent, nmin, nmax = _karras(oracle, v, tr)
    box = {i: (nmin[n - 1 + i, :3].astype(np.float64), nmax[n - 1 + i, :3].astype(np.float64)) for i in range(n)}
    leaf_cost = {i: ci * _area64(*box[i]) for i in range(n)}
    best = min(_tree_cost(t, box, leaf_cost, n, ct, ci)[0] for t in _all_trees(tuple(range(n))))
    # the climb cost of the tree as built (no collapse option)
    def climb(x):
        if x >= n - 1:
            return ci * _area64(nmin[x], nmax[x])
        return ct * _area64(nmin[x], nmax[x]) + climb(left[x]) + climb(right[x])
    before = climb(0)
    l2, r2, p2, mn2, mx2, acc = oracle.tr(left, right, parent, nmin, nmax, iterations=1, ct=ct, ci=ci)
    if acc:
        got = _tree_cost(_nested(l2, r2, n, 0), box, leaf_cost, n, ct, ci)[0]
        assert got == pytest.approx(best, rel=1e-5)
        assert best < before * (1 + 1e-6)
    else:
        assert np.array_equal(l2, left) and np.array_equal(r2, right)
        assert best >= before * (1 - 1e-5)
    if ct == 3.0:
        assert acc == 1 or seed > 0          # Karras trees of clustered soups are rarely optimal


def _py_tr(left, right, nmin, nmax, n, iterations, ct, ci):
    """The TR pass restated with a plain minimum over all bipartitions."""
    L, R = list(map(int, left)), list(map(int, right))
    mn = [np.asarray(nmin[i, :3], f32) for i in range(2 * n - 1)]
    mx = [np.asarray(nmax[i, :3], f32) for i in range(2 * n - 1)]
    area = [_area32(mn[i], mx[i]) for i in range(2 * n - 1)]
    ct, ci = f32(ct), f32(ci)
    gamma = 7
    for _ in range(iterations):
        cost = [f32(0)] * (2 * n - 1)
        refs = [1] * (2 * n - 1)
        for p in range(n - 1, 2 * n - 1):
            cost[p] = f32(ci * area[p])

        def visit(x):
            if x >= n - 1:
                return
            visit(L[x])
            visit(R[x])
            refs[x] = refs[L[x]] + refs[R[x]]
            cost[x] = f32(f32(f32(ct * area[x]) + cost[L[x]]) + cost[R[x]])
            if refs[x] >= gamma:
                treelet(x)

        def treelet(X):
            internal, leaves = [X], [L[X], R[X]]
            for _ in range(5):
                cand = [(area[c], -i) for i, c in enumerate(leaves) if c < n - 1]
                e = -max(cand)[1]
                rep = leaves[e]
                internal.append(rep)
                leaves[e] = L[rep]
                leaves.append(R[rep])
            S = 1 << 7
            smn, smx, C, best = {}, {}, {}, {}
            for s in range(1, S):
                ls = [leaves[i] for i in range(7) if s >> i & 1]
                smn[s] = np.min([mn[c] for c in ls], axis=0)
                smx[s] = np.max([mx[c] for c in ls], axis=0)
            for i in range(7):
                C[1 << i] = cost[leaves[i]]
            Rf = f32(refs[X])
            for s in sorted(range(1, S), key=lambda s: bin(s).count("1")):
                if bin(s).count("1") < 2:
                    continue
                # the split's left side is the part without the subset's lowest
                # leaf, as every partition the reference enumerates is
                rest = s & (s - 1)
                opt, arg = None, None
                p = rest
                while p:
                    c = f32(C[p] + C[s ^ p])
                    if opt is None or c < opt:
                        opt, arg = c, p
                    p = (p - 1) & rest
                A = _area32(smn[s], smx[s])
                C[s] = min(f32(f32(ct * A) + opt), f32(f32(ci * A) * Rf))
                best[s] = arg
            if not C[S - 1] < cost[X]:
                return
            nxt = iter(internal[1:])

            def place(root, s):
                p = best[s]
                for side, sub in ((0, p), (1, s ^ p)):
                    if bin(sub).count("1") > 1:
                        c = next(nxt)
                        mn[c], mx[c] = smn[sub], smx[sub]
                        area[c] = _area32(mn[c], mx[c])
                        cost[c] = C[p]          # both sides: the reference's costs[partition]
                        place(c, sub)
                    else:
                        c = leaves[sub.bit_length() - 1]
                    if side == 0:
                        L[root] = c
                    else:
                        R[root] = c

            place(X, S - 1)

        visit(0)
        gamma *= 2
        if gamma > n:
            break
    return L, R


@pytest.mark.parametrize("n,seed,iters", [(12, 1, 1), (40, 2, 2), (150, 3, 2), (400, 4, 3)])
def test_tr_pass_matches_plain_restatement(oracle, n, seed, iters):
    v, t = scenes.gaussian_mixture_soup(n, components=4, seed=seed)
    left, right, parent, nmin, nmax = _karras(oracle, v, t)
    l2, r2, p2, mn2, mx2, acc = oracle.tr(left, right, parent, nmin, nmax, iterations=iters)
    assert acc > 0
    L, R = _py_tr(left, right, nmin, nmax, n, iters, 3.0, 2.0)
    assert _leaf_sets(l2, r2, n) == _leaf_sets(np.array(L), np.array(R), n)


def _sah(left, right, nmin, nmax, ct=3.0, ci=2.0):
    n = (left.shape[0] + 1) // 2
    return sum((ci if v >= n - 1 else ct) * _area64(nmin[v], nmax[v]) for v in range(2 * n - 1))


def test_tr_sah_never_rises_and_structure(oracle):
    v, t = scenes.gaussian_mixture_soup(20000, seed=7)
    n = t.shape[0]
    left, right, parent, nmin, nmax = _karras(oracle, v, t)
    before = _sah(left, right, nmin, nmax)
    for it in (1, 2, 4):
        l2, r2, p2, mn2, mx2, acc = oracle.tr(left, right, parent, nmin, nmax, iterations=it)
        assert acc > 0
        assert _sah(l2, r2, mn2, mx2) < before
        kids = np.concatenate([l2[: n - 1], r2[: n - 1]])
        assert np.array_equal(np.sort(kids), np.arange(1, 2 * n - 1))
        assert p2[0] == -1
        idx = np.arange(n - 1)
        assert np.all(p2[l2[idx]] == idx) and np.all(p2[r2[idx]] == idx)
        assert np.array_equal(np.minimum(mn2[l2[idx], :3], mn2[r2[idx], :3]), mn2[idx, :3])
        assert np.array_equal(np.maximum(mx2[l2[idx], :3], mx2[r2[idx], :3]), mx2[idx, :3])


def test_tr_zero_iterations_is_lbvh(oracle):
    v, t = scenes.uniform_soup(5000, seed=3)
    a = oracle.build_tr(v, t, iterations=0)
    b = oracle.build_lbvh(v, t)
    ca = oracle.bin_canonical(a["nodes"], a["n_interior"], a["n_leaves"], a["tri_idx"])
    cb = oracle.bin_canonical(b["nodes"], b["n_interior"], b["n_leaves"], b["tri_idx"])
    assert np.array_equal(ca, cb)


def test_build_tr_valid_and_traces_like_brute_force(oracle):
    v, t = scenes.gaussian_mixture_soup(6000, seed=8)
    ref = oracle.build_tr(v, t, iterations=2)
    assert ref["accepted"] > 0
    assert oracle.validate_bin(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
    rays = scenes.random_rays(2000, seed=4)
    hit, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    assert np.array_equal(hit[:, 0], brute[:, 0])
    assert np.array_equal(hit[:, 1], brute[:, 1])
    lb = oracle.build_lbvh(v, t)
    sb = oracle.scene_box(v, t)
    d = sb[3:] - sb[:3]
    area = float(2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]))
    assert oracle.cost(ref["nodes"], area) < oracle.cost(lb["nodes"], area)


def test_build_tr_small_n(oracle):
    for n in (1, 2, 3, 6, 7, 8, 14, 15):
        v, t = scenes.uniform_soup(n, seed=n)
        ref = oracle.build_tr(v, t, iterations=3)
        assert oracle.validate_bin(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
        assert (ref["accepted"] == 0) or n >= 7
