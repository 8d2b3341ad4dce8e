"""Pins the oracle's final-layout refit (HipBVH::refit) against brute force (no GPU)."""
import numpy as np

from hippie_amd import scenes


def _subtree_refs(nodes, idx):
    """Reference slots under node idx (leaf ranges follow the reference's
    node-index order, so a subtree's slots need not be contiguous)."""
    nd = nodes[idx]
    if nd[14] < 0:
        return list(range(nd[12], nd[13]))
    c0 = nd[12] if nd[12] >= 0 else ~nd[12]
    c1 = nd[13] if nd[13] >= 0 else ~nd[13]
    return _subtree_refs(nodes, c0) + _subtree_refs(nodes, c1)


def test_refit_final_is_union_of_triangles(oracle):
    v, t = scenes.uniform_soup(3000, seed=51)
    b = oracle.build_lbvh(v, t, bits=60)
    nodes, tri, ni, nl = b["nodes"], b["tri_idx"], b["n_interior"], b["n_leaves"]
    rng = np.random.default_rng(52)
    v2 = (v + rng.normal(0, 0.01, v.shape)).astype(np.float32)
    r = oracle.refit_final(nodes, ni, nl, tri, v2, t)
    tv = v2[t[tri]]
    lo, hi = tv.min(axis=1), tv.max(axis=1)
    f = r.view(np.float32)
    for i in range(ni):
        for c, (wlo, whi) in enumerate((([0, 2, 8], [1, 3, 9]), ([4, 6, 10], [5, 7, 11]))):
            child = r[i, 12 + c] if r[i, 12 + c] >= 0 else ~r[i, 12 + c]
            slots = _subtree_refs(r, child)
            assert np.array_equal(f[i, wlo], lo[slots].min(axis=0))
            assert np.array_equal(f[i, whi], hi[slots].max(axis=0))
    # topology and leaf ranges untouched
    assert np.array_equal(r[:, 12:16], nodes[:, 12:16])
    assert oracle.validate_bin(r, ni, nl, tri, v2, t) == 0


def test_refit_final_same_vertices_is_identity(oracle):
    v, t = scenes.uniform_soup(4000, seed=53)
    b = oracle.build_lbvh(v, t, bits=30)
    r = oracle.refit_final(b["nodes"], b["n_interior"], b["n_leaves"], b["tri_idx"], v, t)
    assert np.array_equal(r, b["nodes"])


def test_icosphere_counts():
    for level in range(4):
        v, f = scenes.icosphere(level)
        assert f.shape[0] == 20 * 4 ** level and v.shape[0] == 10 * 4 ** level + 2
        # closed manifold: every edge shared by exactly two faces
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)
