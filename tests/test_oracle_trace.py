"""Pins the oracle's trace and ray keys (no GPU): BVH-free brute force, FP64
Moller-Trumbore, the closest-hit tie rule and a table-driven ray-key packing."""
import numpy as np
import pytest

from hippie_amd import scenes


def _mt64(v, t, rays, ids):
    tri = t[ids]
    v0, v1, v2 = (v[tri[:, k]].astype(np.float64) for k in range(3))
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=1)
    q = np.cross(o - v0, e1)
    return np.sum(e2 * q, axis=1) / det


@pytest.fixture(scope="module")
def soup(oracle):
    v, t = scenes.uniform_soup(6000, seed=41)
    b = oracle.build_lbvh(v, t, bits=60)
    return v, t, b


def test_trace_matches_brute_force(oracle, soup):
    v, t, b = soup
    rays = scenes.random_rays(20000, seed=42)
    res, ov = oracle.trace(b["nodes"], b["woop"], b["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert ov == 0
    assert np.array_equal(res[:, 0], brute[:, 0])
    assert np.array_equal(res[:, 1], brute[:, 1])
    h = res[:, 0] >= 0
    assert h.sum() > 1000
    t64 = _mt64(v, t, rays[h], res[h, 0])
    assert np.all(np.abs(res[h, 1].view(np.float32) - t64) <= 1e-5 * np.abs(t64))
    assert np.all(res[~h, 1].view(np.float32) == rays[~h, 7])


def test_any_hit_is_a_hit(oracle, soup):
    v, t, b = soup
    rays = scenes.random_rays(20000, seed=43)
    res, _ = oracle.trace(b["nodes"], b["woop"], b["tri_idx"], rays, any_hit=True)
    brute = oracle.trace_brute(v, t, rays)
    assert np.array_equal(res[:, 0] >= 0, brute[:, 0] >= 0)
    h = res[:, 0] >= 0
    tv = res[h, 1].view(np.float32)
    assert np.all(tv > 0) and np.all(tv < rays[h, 7])
    assert np.all(tv >= brute[h, 1].view(np.float32))


def test_closest_hit_tie_goes_to_smaller_id(oracle):
    """Identical triangles give bit-identical t; the smaller id must win in
    any storage order (the reference keeps the first one found)."""
    rng = np.random.default_rng(44)
    base = np.array([[0.2, 0.2, 0.5], [0.8, 0.2, 0.5], [0.5, 0.8, 0.5]], np.float32)
    filler_v, filler_t = scenes.uniform_soup(200, seed=45)
    filler_v[:, 2] += np.float32(2.0)  # behind the z = 0.5 plane
    for perm_seed in range(5):
        ids = [17, 3, 120] if perm_seed % 2 else [120, 3, 17]
        v = filler_v.copy()
        t = filler_t.copy()
        for i in ids:
            v[3 * i:3 * i + 3] = base
        rays = np.zeros((100, 8), np.float32)
        rays[:, 0:2] = rng.uniform(0.4, 0.6, (100, 2))
        rays[:, 2] = -1.0
        rays[:, 6] = 1.0
        rays[:, 7] = 5.0
        b = oracle.build_lbvh(v, t, bits=60)
        res, _ = oracle.trace(b["nodes"], b["woop"], b["tri_idx"], rays)
        assert np.all(res[:, 0] == 3)
        brute = oracle.trace_brute(v, t, rays)
        assert np.all(brute[:, 0] == 3)


def test_stack_depth_and_stats(oracle, soup):
    v, t, b = soup
    rays = scenes.random_rays(5000, seed=46)
    res, stats, ov = oracle.trace(b["nodes"], b["woop"], b["tri_idx"], rays, want_stats=True)
    assert ov == 0
    assert np.all(stats[:, 0] >= 1)
    assert np.all(stats[res[:, 0] >= 0, 1] >= 1)


# ----------------------------------------------------------- ray keys --

def aila_table():
    """(key bit, source, axis, source bit) of computeMortonCodesAila64
    (RayBufferKernels.cu:578-608): origin bits 12..10 on top, then origin
    bits 9..1 interleaved 6 per level with direction bits 12..4."""
    tab = []
    pos = 63
    for b in (12, 11, 10):
        for ax in range(3):
            tab.append((pos, "o", ax, b))
            pos -= 1
    for k in range(9, 0, -1):
        for ax in range(3):
            tab.append((6 * k - ax, "o", ax, k))
    for k in range(12, 3, -1):
        for ax in range(3):
            tab.append((6 * k - 21 - ax, "d", ax, k))
    return tab


def test_aila_table_is_a_bit_permutation():
    tab = aila_table()
    pos = [p for p, *_ in tab]
    # 9 top origin bits + 27 origin bits + 27 direction bits; bit 0 unused
    assert len(pos) == len(set(pos)) == 63
    assert max(pos) == 63 and min(pos) == 1


def test_ray_keys_aila(oracle):
    rays = scenes.random_rays(30000, seed=47)
    f = np.float32
    o = rays[:, 0:3].astype(f)
    d = rays[:, 4:7].astype(f)
    ls = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    rl = f(1.0) / np.sqrt(ls).astype(f)
    dn = ((d * rl[:, None]) + f(1.0)) * f(0.5)
    ia = np.clip(np.trunc(o * f(8191.0)), -2 ** 31, 2 ** 31 - 1).astype(np.int64)
    ib = np.trunc(dn * f(8191.0)).astype(np.int64)
    key = np.zeros(rays.shape[0], np.uint64)
    for pos, src, ax, b in aila_table():
        val = ia[:, ax] if src == "o" else ib[:, ax]
        key |= (((val >> b) & 1).astype(np.uint64)) << np.uint64(pos)
    for bits in (30, 60, 64):
        assert np.array_equal(oracle.ray_keys_aila64(rays, bits), key >> np.uint64(64 - bits))


def test_ray_keys_twopoint_properties(oracle):
    """TwoPoint64: the origin half depends only on the origin, identical rays
    give identical keys, and the key is the documented bit permutation."""
    rays = scenes.random_rays(10000, seed=48)
    k = oracle.ray_keys_twopoint64(rays, 64, 0.25)
    r2 = rays.copy()
    r2[:, 4:7] *= -1.0
    k2 = oracle.ray_keys_twopoint64(r2, 64, 0.25)
    # bits from the origin (positions 6i-9 .. 6i-11, i = 12..2) agree
    omask = np.uint64(0)
    for i in range(12, 1, -1):
        for s in (9, 10, 11):
            omask |= np.uint64(1) << np.uint64(6 * i - s)
    assert np.array_equal(k & omask, k2 & omask)
    assert np.array_equal(oracle.ray_keys_twopoint64(rays.copy(), 64, 0.25), k)
