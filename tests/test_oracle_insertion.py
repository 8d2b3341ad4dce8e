"""Pinning the oracle's insertion optimiser (InsertionBuilder.cpp:22-274,
InsertionBuilderKernels.cu:11-515).

Independent checks, on trees whose leaf boxes have small integer corners so
that every area and every sum of areas is exact in float32:
- findBestNode equals a brute force over every candidate sibling, in the
  search's visiting order (preorder of the sibling's subtree, then for each
  ancestor the preorder of its other child followed by the ancestor itself,
  the root excluded), with the area reduction written in closed form: the
  parent's area minus the new parent's area, plus the shrinkage of the
  ancestors below the common ancestor, plus the growth of the nodes on the
  path down to the new sibling;
- the whole loop equals a re-statement built on that brute force, in which a
  move's lock set is derived as the set of nodes whose links the move writes,
  the kept moves are applied in reverse order (their write sets are
  disjoint), boxes are refitted exactly and the cost is the exact fixed-point
  sum of float32 node terms;
- on float scenes: every step keeps a valid tree over the same leaves, the
  result is never worse than the start, and the built Bin tree validates and
  traces like brute force.
"""
import numpy as np
import pytest

from hippie_amd import scenes

F32 = np.float32


def _random_tree(n, rng):
    """Random binary tree with the builders' numbering: root 0, interiors 0..n-2, leaves n-1..2n-2."""
    N = 2 * n - 1
    left = np.full(N, -1, np.int32)
    right = np.full(N, -1, np.int32)
    parent = np.full(N, -1, np.int32)
    ids = list(rng.permutation(np.arange(1, n - 1))) if n > 2 else []
    roots = [n - 1 + i for i in range(n)]
    while len(roots) > 1:
        a, b = rng.choice(len(roots), 2, replace=False)
        x, y = roots[a], roots[b]
        roots = [r for k, r in enumerate(roots) if k not in (a, b)]
        v = ids.pop() if len(roots) > 0 else 0
        left[v], right[v] = x, y
        parent[x] = parent[y] = v
        roots.append(v)
    return left, right, parent


def _int_leaf_boxes(n, rng, span=12):
    lo = rng.integers(0, span, (n, 3))
    ext = rng.integers(0, 5, (n, 3))
    return lo, lo + ext


def _refit_exact(left, right, n, lo, hi):
    N = 2 * n - 1
    mn = np.zeros((N, 3), np.int64)
    mx = np.zeros((N, 3), np.int64)
    mn[n - 1:] = lo
    mx[n - 1:] = hi

    def rec(v):
        if v >= n - 1:
            return
        rec(left[v])
        rec(right[v])
        mn[v] = np.minimum(mn[left[v]], mn[right[v]])
        mx[v] = np.maximum(mx[left[v]], mx[right[v]])
    rec(0)
    return mn, mx


def _as4(a):
    out = np.zeros((a.shape[0], 4), np.float32)
    out[:, :3] = a
    return out


def _area(mn, mx):
    d = [int(mx[k]) - int(mn[k]) for k in range(3)]
    if min(d) < 0:
        return 0
    return 2 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _union(a, b):
    return np.minimum(a[0], b[0]), np.maximum(a[1], b[1])


def _brute_find(i, left, right, parent, mn, mx, n):
    box = lambda v: (mn[v], mx[v])
    p = parent[i]
    sib = right[p] if left[p] == i else left[p]
    inb = box(i)
    pa = _area(*box(p))
    best = [0, -1]

    def consider(x, val):
        if val > best[0]:
            best[0], best[1] = val, x

    def rec(x, red, skip):
        if x != skip:
            consider(x, pa - _area(*_union(inb, box(x))) + red)
        if x < n - 1:
            red2 = red + _area(*box(x)) - _area(*_union(box(x), inb))
            rec(left[x], red2, skip)
            rec(right[x], red2, skip)

    rec(sib, 0, sib)
    H = box(sib)
    acc = 0
    below, a = p, parent[p]
    while a >= 0:
        s = right[a] if left[a] == below else left[a]
        rec(s, acc, -1)
        H = _union(H, box(s))
        if a != 0:
            consider(a, pa - _area(*_union(inb, H)) + acc)
            acc += _area(*box(a)) - _area(*H)
        below, a = a, parent[a]
    return best[1], best[0]


@pytest.mark.parametrize("seed", range(6))
def test_find_best_equals_brute_force(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 40))
    left, right, parent = _random_tree(n, rng)
    lo, hi = _int_leaf_boxes(n, rng)
    mn, mx = _refit_exact(left, right, n, lo, hi)
    out, red = oracle.ins_find(left, right, parent, _as4(mn), _as4(mx))
    assert out[0] == -1
    for i in range(1, 2 * n - 1):
        bo, br = _brute_find(i, left, right, parent, mn, mx, n)
        assert out[i] == bo, (seed, i)
        assert (red[i] == F32(br)) if bo >= 0 else True, (seed, i)


def _write_set_move(i, out, left, right, parent):
    """Apply the move of node i beside `out` to the dict tree; returns the nodes whose links it wrote."""
    wrote = set()
    p = parent[i]
    sib = right[p] if left[p] == i else left[p]
    if p != 0:
        pp = parent[p]
        if left[pp] == p:
            left[pp] = sib
        else:
            right[pp] = sib
        parent[sib] = pp
        wrote |= {pp, sib}
        newp = p
    else:
        a, b = left[sib], right[sib]
        left[0], right[0] = a, b
        parent[a] = parent[b] = 0
        wrote |= {0, a, b}
        newp = sib
    op = parent[out]
    if left[op] == out:
        left[op] = newp
    else:
        right[op] = newp
    parent[newp] = op
    left[newp], right[newp] = out, i
    parent[out] = parent[i] = newp
    wrote |= {op, newp, out, i}
    return wrote


def _fixed_term(c, area, scene_area):
    P = F32(F32(area) / F32(scene_area))
    t = F32(F32(c) * P)
    return int(np.float64(t) * 4294967296.0)


def _py_insertion(left, right, parent, lo, hi, n, mod, ct, ci):
    left, right, parent = (list(map(int, a)) for a in (left, right, parent))
    N = 2 * n - 1
    mn, mx = _refit_exact(left, right, n, lo, hi)
    scene = _area(mn[0], mx[0])
    cost, best_cost = 3.402823466e38, 3.402823466e38
    best = (left[:], right[:], parent[:])
    steps, mod_cur, total = 0, max(mod, 1), 0
    while True:
        prev = cost
        rem = steps % mod_cur
        moves = {}
        for i in range(rem, N, mod_cur):
            if i == 0:
                continue
            o, r = _brute_find(i, left, right, parent, mn, mx, n)
            if o >= 0:
                moves[i] = (o, (int(F32(r).view(np.uint32)) << 32) | i)
        # Lock sets: the nodes a move writes, found on a scratch copy.
        sets = {}
        for i, (o, _) in moves.items():
            sets[i] = _write_set_move(i, o, left[:], right[:], parent[:])
        owner = {}
        for i, (o, key) in moves.items():
            for v in sets[i]:
                owner[v] = max(owner.get(v, 0), key)
        kept = [i for i, (o, key) in moves.items() if all(owner[v] == key for v in sets[i])]
        for i in reversed(kept):
            _write_set_move(i, moves[i][0], left, right, parent)
        total += len(kept)
        mn, mx = _refit_exact(left, right, n, lo, hi)
        cost = sum(_fixed_term(ci if v >= n - 1 else ct, _area(mn[v], mx[v]), scene) for v in range(N))
        cost = cost / 4294967296.0
        if best_cost > cost:
            best_cost = cost
            best = (left[:], right[:], parent[:])
        steps += 1
        if (abs(prev - cost) <= 0.01 or not kept) and mod_cur == 1:
            break
        if abs(prev - cost) <= 0.01:
            mod_cur = max(1, mod_cur - 1)
    return [np.array(a, np.int32) for a in best], total, steps


@pytest.mark.parametrize("seed,mod", [(0, 1), (1, 1), (2, 2), (3, 3), (4, 1), (5, 4)])
def test_insertion_loop_equals_restatement(oracle, seed, mod):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(8, 48))
    left, right, parent = _random_tree(n, rng)
    lo, hi = _int_leaf_boxes(n, rng, span=16)
    mn, mx = _refit_exact(left, right, n, lo, hi)
    (el, er, ep), total, steps = _py_insertion(left, right, parent, lo, hi, n, mod, 3.0, 2.0)
    l2, r2, p2, mn2, mx2, k, st = oracle.insertion(left, right, parent, _as4(mn), _as4(mx), mod=mod)
    assert total > 0
    assert (k, st) == (total, steps)
    assert np.array_equal(l2, el) and np.array_equal(r2, er) and np.array_equal(p2, ep)
    emn, emx = _refit_exact(el, er, n, lo, hi)
    assert np.array_equal(mn2[:, :3], emn.astype(F32)) and np.array_equal(mx2[:, :3], emx.astype(F32))


def _check_tree(left, right, parent, n):
    seen = np.zeros(2 * n - 1, int)
    st = [0]
    assert parent[0] == -1
    while st:
        v = st.pop()
        seen[v] += 1
        if v < n - 1:
            for c in (left[v], right[v]):
                assert parent[c] == v
                st.append(c)
    assert np.all(seen == 1)


def _sah(nmin, nmax, n, ct=3.0, ci=2.0):
    d = (nmax[:, :3] - nmin[:, :3]).astype(np.float64)
    a = 2 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
    return (ct * a[: n - 1].sum() + ci * a[n - 1:].sum()) / a[0]


@pytest.mark.parametrize("mod", [1, 3])
def test_insertion_improves_lbvh_on_float_scene(oracle, mod):
    v, t = scenes.merge(scenes.gaussian_mixture_soup(3000, seed=9), scenes.ground_grid(20))
    n = t.shape[0]
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    keys = oracle.morton_codes(bmin, bmax, sb, 60)
    sk, sv = oracle.sort_pairs(keys, np.arange(n, dtype=np.int32), 60)
    left, right, parent = oracle.construct(sk)
    nmin, nmax = oracle.refit(sv, bmin, bmax, left, right, parent)
    l2, r2, p2, mn2, mx2, k, steps = oracle.insertion(left, right, parent, nmin, nmax, mod=mod)
    _check_tree(l2, r2, p2, n)
    assert k > 0 and steps >= 1
    assert _sah(mn2, mx2, n) < _sah(nmin, nmax, n)
    # Leaf boxes never move.
    assert np.array_equal(mn2[n - 1:], nmin[n - 1:]) and np.array_equal(mx2[n - 1:], nmax[n - 1:])


def test_build_insertion_valid_and_traces_like_brute_force(oracle):
    v, t = scenes.merge(scenes.gaussian_mixture_soup(4000, seed=5), scenes.ground_grid(16))
    ref = oracle.build_insertion(v, t)
    assert ref["steps"] >= 1
    assert oracle.validate_bin(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
    rays = scenes.random_rays(2000, seed=6)
    hit, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    assert np.array_equal(hit[:, 0], brute[:, 0]) and np.array_equal(hit[:, 1], brute[:, 1])
