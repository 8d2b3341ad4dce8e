"""Pins the CPU oracle's LBVH build against independent checks (no GPU).

The reference ships no golden vectors, so each oracle function is checked
against what the mathematics fixes: a bit-loop Morton interleave, numpy's
stable sort, a recursive radix-tree split, brute-force box unions, a direct
top-down collapse and the FP64 Woop frame property.
"""
import numpy as np
import pytest

from hippie_amd import scenes


# ---------------------------------------------------------------- helpers --

def interleave_bits(x, y, z, nbits):
    """Bit loop: bit 3k = x_k, 3k+1 = y_k, 3k+2 = z_k (rt/HipUtil.h:35-49)."""
    code = np.zeros(x.shape, np.uint64)
    for k in range(nbits):
        code |= ((x >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
        code |= ((y >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k + 1)
        code |= ((z >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k + 2)
    return code


def expected_morton60(bmin, bmax, sb):
    """computeMortonCodes60 in float32 numpy with the interleave as a bit loop."""
    f = np.float32
    mn, mx = bmin[:, :3].astype(f), bmax[:, :3].astype(f)
    c = (mn + mx) * f(0.5)
    scale = f(1.0) / (sb[3:6] - sb[0:3]).astype(f)
    q = (c - sb[0:3].astype(f)) * scale
    xyz = np.floor(q * f(1048575.0)).astype(np.uint64)
    lo = interleave_bits(xyz[:, 0] & np.uint64(1023), xyz[:, 1] & np.uint64(1023), xyz[:, 2] & np.uint64(1023), 10)
    hi = interleave_bits(xyz[:, 0] >> np.uint64(10), xyz[:, 1] >> np.uint64(10), xyz[:, 2] >> np.uint64(10), 10)
    return (hi << np.uint64(30)) | lo


def karras_ranges(left, right, n):
    """(first, last) of every internal node from the child arrays."""
    first = np.zeros(2 * n - 1, np.int64)
    last = np.zeros(2 * n - 1, np.int64)
    first[n - 1:] = np.arange(n)
    last[n - 1:] = np.arange(n)
    order = []
    stack = [0]
    while stack:
        i = stack.pop()
        order.append(i)
        for c in (left[i], right[i]):
            if c < n - 1:
                stack.append(c)
    for i in reversed(order):
        first[i] = first[left[i]]
        last[i] = last[right[i]]
    return first, last


def radix_split(keys, lo, hi):
    """Split of [lo, hi] at the highest differing bit of the (key, index) pair."""
    if keys[lo] != keys[hi]:
        bit = int(keys[lo] ^ keys[hi]).bit_length() - 1
        aug = lambda i: (int(keys[i]) >> bit) & 1  # noqa: E731
    else:
        bit = (lo ^ hi).bit_length() - 1
        aug = lambda i: (i >> bit) & 1  # noqa: E731
    # last index in [lo, hi] whose augmented bit is 0 (binary search)
    a, b = lo, hi
    while a < b:
        m = (a + b + 1) // 2
        if aug(m) == 0:
            a = m
        else:
            b = m - 1
    return a


def brute_tree(keys):
    """All (lo, hi, split) triples of the radix tree over sorted keys."""
    out = set()
    stack = [(0, len(keys) - 1)]
    while stack:
        lo, hi = stack.pop()
        if lo == hi:
            continue
        g = radix_split(keys, lo, hi)
        out.add((lo, hi, g))
        stack.append((lo, g))
        stack.append((g + 1, hi))
    return out


def build_pieces(oracle, v, t, bits):
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    keys = oracle.morton_codes(bmin, bmax, sb, bits)
    sk, sv = oracle.sort_pairs(keys, np.arange(t.shape[0], dtype=np.int32), bits)
    left, right, parent = oracle.construct(sk)
    nmin, nmax = oracle.refit(sv, bmin, bmax, left, right, parent)
    return sb, bmin, bmax, keys, sk, sv, left, right, parent, nmin, nmax


# ------------------------------------------------------------------ tests --

def test_scene_box_and_setup_boxes(oracle):
    v, t = scenes.uniform_soup(5000, seed=31)
    tv = v[t]
    assert np.array_equal(oracle.scene_box(v, t), np.concatenate([tv.min(axis=(0, 1)), tv.max(axis=(0, 1))]))
    bmin, bmax = oracle.setup_boxes(v, t)
    assert np.array_equal(bmin[:, :3], tv.min(axis=1)) and np.array_equal(bmax[:, :3], tv.max(axis=1))
    assert np.all(bmin[:, 3] == 0) and np.all(bmax[:, 3] == 0)


def test_morton_interleave_all_10bit_values(oracle):
    """Every 10-bit value on each axis: place centroids exactly on lattice points."""
    vals = np.arange(1024, dtype=np.float32)
    # scene box [0, 1048575]: q * 1048575 reproduces the lattice value exactly
    sb = np.array([0, 0, 0, 1048575, 1048575, 1048575], np.float32)
    for axis in range(3):
        c = np.zeros((1024, 3), np.float32)
        c[:, axis] = vals * 1024.0 + vals  # hi and lo halves both = vals
        b4 = np.zeros((1024, 4), np.float32)
        b4[:, :3] = c
        keys = oracle.morton_codes(b4, b4, sb, 60)
        assert np.array_equal(keys, expected_morton60(b4, b4, sb))
        xyz = [np.zeros(1024, np.uint64)] * 3
        xyz = [vals.astype(np.uint64) if k == axis else np.zeros(1024, np.uint64) for k in range(3)]
        part = interleave_bits(*xyz, 10)
        assert np.array_equal(keys, (part << np.uint64(30)) | part)


def test_morton_random_and_shift(oracle):
    v, t = scenes.uniform_soup(20000, seed=32)
    sb = oracle.scene_box(v, t)
    bmin, bmax = oracle.setup_boxes(v, t)
    k60 = oracle.morton_codes(bmin, bmax, sb, 60)
    assert np.array_equal(k60, expected_morton60(bmin, bmax, sb))
    for bits in (6, 30, 45):
        assert np.array_equal(oracle.morton_codes(bmin, bmax, sb, bits), k60 >> np.uint64(60 - bits))


def test_morton_end_points(oracle):
    sb = np.array([-1, 2, 3, 5, 7, 11], np.float32)
    b = np.array([[-1, 2, 3, 0], [5, 7, 11, 0]], np.float32)
    keys = oracle.morton_codes(b, b, sb, 60)
    assert keys[0] == 0
    assert keys[1] == (1 << 60) - 1


@pytest.mark.parametrize("kind", ["random", "dups", "equal", "sorted", "reversed"])
def test_sort_is_stable_sort(oracle, kind):
    rng = np.random.default_rng(33)
    n = 30001
    k = {"random": rng.integers(0, 1 << 60, n, dtype=np.uint64),
         "dups": rng.integers(0, 7, n, dtype=np.uint64) << np.uint64(40),
         "equal": np.full(n, 5, np.uint64),
         "sorted": np.arange(n, dtype=np.uint64),
         "reversed": np.arange(n, dtype=np.uint64)[::-1].copy()}[kind]
    vals = rng.permutation(n).astype(np.int32)
    sk, sv = oracle.sort_pairs(k, vals, 60)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(sk, k[order]) and np.array_equal(sv, vals[order])


@pytest.mark.parametrize("n,bits,dup", [(2, 60, 0), (8, 30, 0), (257, 60, 0), (3000, 30, 0), (2000, 60, 1)])
def test_construct_matches_radix_tree(oracle, n, bits, dup):
    rng = np.random.default_rng(n)
    if dup:
        keys = np.sort(rng.integers(0, 40, n, dtype=np.uint64))
    else:
        keys = np.sort(rng.integers(0, 1 << bits, n, dtype=np.uint64))
    left, right, parent = oracle.construct(keys)
    assert parent[0] == -1
    first, last = karras_ranges(left, right, n)
    got = set()
    for i in range(n - 1):
        lc, rc = left[i], right[i]
        assert parent[lc] == i and parent[rc] == i
        g = last[lc]
        assert first[rc] == g + 1
        assert i in (first[i], last[i])  # Karras: one end of the range is the node index
        got.add((int(first[i]), int(last[i]), int(g)))
    assert got == brute_tree(keys)


def test_karras_paper_example(oracle):
    """The 8-key example of Karras 2012 (Fig. 3), tree derived by brute force."""
    keys = np.array([0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110], np.uint64)
    left, right, parent = oracle.construct(keys)
    first, last = karras_ranges(left, right, 8)
    triples = {(int(first[i]), int(last[i]), int(last[left[i]])) for i in range(7)}
    assert triples == brute_tree(keys)
    assert (first[0], last[0], last[left[0]]) == (0, 7, 3)


def test_refit_is_union_of_leaf_boxes(oracle):
    v, t = scenes.uniform_soup(4000, seed=34)
    sb, bmin, bmax, keys, sk, sv, left, right, parent, nmin, nmax = build_pieces(oracle, v, t, 60)
    n = t.shape[0]
    first, last = karras_ranges(left, right, n)
    tv = v[t[sv]]  # sorted triangles' vertices
    lo = tv.min(axis=1)
    hi = tv.max(axis=1)
    for i in range(2 * n - 1):
        a, b = first[i], last[i] + 1
        assert np.array_equal(nmin[i, :3], lo[a:b].min(axis=0))
        assert np.array_equal(nmax[i, :3], hi[a:b].max(axis=0))


def _area32(mn, mx):
    f = np.float32
    d = (mx - mn).astype(f)
    return f(2.0) * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


@pytest.mark.parametrize("adaptive,max_leaf", [(True, 8), (False, 1), (False, 4), (False, 8)])
def test_collapse_matches_top_down(oracle, adaptive, max_leaf):
    """The reference's multi-pass collapse equals a direct top-down one."""
    v, t = scenes.uniform_soup(3000, seed=35)
    sb, bmin, bmax, keys, sk, sv, left, right, parent, nmin, nmax = build_pieces(oracle, v, t, 60)
    n = t.shape[0]
    first, last = karras_ranges(left, right, n)
    f = np.float32
    ct, ci = f(3.0), f(2.0)
    state = np.zeros(2 * n - 1, np.int32)
    cost = np.zeros(2 * n - 1, np.float32)
    # post-order, children before parents
    order = []
    stack = [0]
    while stack:
        i = stack.pop()
        order.append(i)
        for c in (left[i], right[i]):
            if c < n - 1:
                stack.append(c)
    for leaf in range(n - 1, 2 * n - 1):
        cost[leaf] = ci * _area32(nmin[leaf, :3], nmax[leaf, :3])
    for i in reversed(order):
        area = _area32(nmin[i, :3], nmax[i, :3])
        size = last[i] - first[i] + 1
        if adaptive:
            c_int = f(f(ct * area) + cost[left[i]]) + cost[right[i]]
            c_leaf = f(ci * area) * f(size)
            state[i] = 0 if c_leaf < c_int else 1
            cost[i] = c_leaf if state[i] == 0 else c_int
        else:
            state[i] = 0 if size <= max_leaf else 1
    nodes, tri, ni, nl = oracle.collapse(sv, left, right, parent, nmin, nmax, adaptive, max_leaf)
    assert oracle.validate_bin(nodes, ni, nl, tri, v, t) == 0
    # Direct top-down emission from the root (always interior): a child in
    # state 1 is an interior node, anything else a leaf over its whole range.
    # Serialised like hxo_bin_canonical.
    ref = []
    stack = [("i", 0)]
    while stack:
        kind, i = stack.pop()
        if kind == "i":
            ref.append(0)
            l, r = left[i], right[i]
            words = [nmin[l, 0], nmax[l, 0], nmin[l, 1], nmax[l, 1], nmin[r, 0], nmax[r, 0], nmin[r, 1], nmax[r, 1],
                     nmin[l, 2], nmax[l, 2], nmin[r, 2], nmax[r, 2]]
            ref += [int(np.float32(w).view(np.uint32)) for w in words]
            kids = [("i" if (c < n - 1 and state[c] == 1) else "l", c) for c in (l, r)]
            for k in reversed(kids):
                stack.append(k)
        else:
            ref.append(1)
            ref += [int(np.float32(w).view(np.uint32)) for w in
                    (nmin[i, 0], nmax[i, 0], nmin[i, 1], nmax[i, 1], nmin[i, 2], nmax[i, 2])]
            ids = np.sort(sv[first[i]:last[i] + 1])
            ref.append(len(ids))
            ref += [int(x) for x in ids]
    got = oracle.bin_canonical(nodes, ni, nl, tri)
    assert np.array_equal(got, np.array(ref, np.uint32))
    if not adaptive:
        leaves = nodes[ni:]
        assert np.all(~leaves[:, 14] <= max_leaf)


def test_single_and_two_references(oracle):
    for n in (1, 2):
        v, t = scenes.uniform_soup(n, seed=36)
        b = oracle.build_lbvh(v, t, bits=60)
        assert oracle.validate_bin(b["nodes"], b["n_interior"], b["n_leaves"], b["tri_idx"], v, t) == 0
        assert b["nodes"][0, 15] == -1 and b["nodes"][0, 14] >= 0  # the root is interior


def test_woop_frame_fp64(oracle):
    """The Woop rows map v0 -> (u,v)=(1,0), v1 -> (0,1), v2 -> (0,0), all on z = 0."""
    v, t = scenes.uniform_soup(5000, seed=37)
    w = oracle.woopify(v, t, np.arange(t.shape[0], dtype=np.int32)).astype(np.float64)
    tv = v[t].astype(np.float64)
    for k, (eu, ev) in enumerate([(1, 0), (0, 1), (0, 0)]):
        p = tv[:, k]
        z = np.sum(w[:, 0:3] * p, axis=1) - w[:, 3]
        u = np.sum(w[:, 4:7] * p, axis=1) + w[:, 7]
        vv = np.sum(w[:, 8:11] * p, axis=1) + w[:, 11]
        assert np.max(np.abs(u - eu)) < 2e-3 and np.max(np.abs(vv - ev)) < 2e-3
        scale = np.abs(w[:, 0:3]).sum(axis=1) * np.abs(p).sum(axis=1)
        assert np.max(np.abs(z) / scale) < 1e-5


def test_woop_matches_fp64_inverse(oracle):
    """FP32 cofactor inverse rows agree with the FP64 inverse to FP32 accuracy."""
    v, t = scenes.uniform_soup(2000, seed=38)
    w = oracle.woopify(v, t, np.arange(t.shape[0], dtype=np.int32)).astype(np.float64)
    tv = v[t].astype(np.float64)
    for i in range(0, 2000, 97):
        v0, v1, v2 = tv[i]
        M = np.eye(4)
        M[:3, 0] = v0 - v2
        M[:3, 1] = v1 - v2
        M[:3, 2] = np.cross(v0 - v2, v1 - v2)
        M[:3, 3] = v2
        inv = np.linalg.inv(M)
        ref = np.concatenate([[inv[2, 0], inv[2, 1], inv[2, 2], -inv[2, 3]], inv[0], inv[1]])
        assert np.allclose(w[i], ref, rtol=1e-3, atol=1e-3 * np.abs(ref).max())


def test_cost_closed_form(oracle):
    v, t = scenes.uniform_soup(3000, seed=39)
    b = oracle.build_lbvh(v, t, bits=60)
    sb = oracle.scene_box(v, t)
    sa = float(_area32(sb[0:3], sb[3:6]))
    nodes = b["nodes"]
    f = nodes.view(np.float32)
    total = 0.0
    for nd, fl in zip(nodes, f):
        mn = np.minimum(fl[[0, 2, 8]], fl[[4, 6, 10]])
        mx = np.maximum(fl[[1, 3, 9]], fl[[5, 7, 11]])
        d = (mx - mn).astype(np.float64)
        area = 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])
        p = area / sa
        total += 2.0 * p * (~nd[14]) if nd[14] < 0 else 3.0 * p
    c = oracle.cost(nodes, sa)
    assert abs(c - total) <= 1e-5 * total
    # the root spans the scene box: its term is ct
    assert total > 3.0


def test_validate_catches_planted_mistakes(oracle):
    """Mistakes planted in an oracle-built tree are detected by the checkers."""
    v, t = scenes.uniform_soup(2000, seed=40)
    b = oracle.build_lbvh(v, t, bits=60)
    nodes, tri, ni, nl = b["nodes"], b["tri_idx"], b["n_interior"], b["n_leaves"]
    assert oracle.validate_bin(nodes, ni, nl, tri, v, t) == 0
    base = oracle.bin_canonical(nodes, ni, nl, tri)
    bad = nodes.copy()
    bad[0, 12], bad[0, 13] = bad[0, 13], bad[0, 12]            # swapped children
    assert not np.array_equal(oracle.bin_canonical(bad, ni, nl, tri), base)
    bad = nodes.copy()
    bad[ni + 3, 0] = np.float32(bad[ni + 3, 0].view(np.float32) + 0.5).view(np.int32)  # shrunk leaf box
    assert oracle.validate_bin(bad, ni, nl, tri, v, t) != 0
    bad_tri = tri.copy()
    bad_tri[0] = bad_tri[1]                                       # a triangle lost
    assert oracle.validate_bin(nodes, ni, nl, bad_tri, v, t) != 0
    bad = nodes.copy()
    bad[1, 15] = 7                                                # wrong parent
    assert oracle.validate_bin(bad, ni, nl, tri, v, t) != 0


@pytest.mark.parametrize("seed,scale", [(1, (3.0, 1.0, 0.5)), (2, (0.2, 5.0, 1.0)), (3, (1.0, 1.0, 7.0))])
def test_normalize_closed_form(oracle, seed, scale):
    """SceneLoader.cpp:134-141: the vertex box, then (v - mn) * (1 / largest
    extent) with each operation rounded to float32, whichever axis is longest."""
    rng = np.random.default_rng(seed)
    v = (rng.normal(size=(5000, 3)) * np.array(scale) - 2.5).astype(np.float32)
    got, box = oracle.normalize(v)
    mn, mx = v.min(0), v.max(0)
    assert np.array_equal(box, np.concatenate([mn, mx]))
    d = mx - mn
    s = np.float32(1.0) / np.float32(d.max())
    exp = (v - mn).astype(np.float32) * s
    assert np.array_equal(got.view(np.uint32), exp.astype(np.float32).view(np.uint32))
    assert got.min() == 0.0 and np.isclose(got.max(), 1.0, rtol=1e-6)
