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


def twopoint_table():
    """(key bit, source, axis, source bit) of computeMortonCodesTwoPoint64
    (RayBufferKernels.cu:530-547): origin bits 12..2 at 6i-9, 6i-10, 6i-11;
    end-point bits 12..3 at 6i-12, 6i-13, 6i-14; end-point x bit 0 at 0."""
    tab = []
    for i in range(12, 1, -1):
        for ax in range(3):
            tab.append((6 * i - 9 - ax, "a", ax, i))
    for i in range(12, 2, -1):
        for ax in range(3):
            tab.append((6 * i - 12 - ax, "b", ax, i))
    tab.append((0, "b", 0, 0))
    return tab


def _f2i(x):
    """float32 -> int32 as the device convert does: truncate, saturate, NaN -> 0."""
    x = np.asarray(x, np.float32)
    r = np.where(np.isnan(x), np.float32(0), np.trunc(x)).astype(np.float64)
    return np.clip(r, -2.0 ** 31, 2.0 ** 31 - 1).astype(np.int64)


def _pack(table, ia, ib):
    key = np.zeros(ia.shape[0], np.uint64)
    for pos, src, ax, b in table:
        val = ia[:, ax] if src == "a" else ib[:, ax]
        key |= (((val >> b) & 1).astype(np.uint64)) << np.uint64(pos)
    return key


def twopoint_closed_form(rays, bits, ray_length, plant=None):
    """numpy float32 restatement of the TwoPoint64 key, vectorised over rays.

    Box exit of the [0,1]^3 ray box by slabs (intersect, RayBufferKernels.cu:
    12-56: per axis the near/far plane by the direction's sign, t0 = fmaxf
    over axes, t1 = fminf, -1 when t0 > t1, else (1 - 1e-3) * (t0 >= 0 ? t0 :
    t1)); len = fminf(fminf(rayLength, fmaxf(tmax, 0)), boxT) (:516-521);
    points o and o + len * d (each product and sum rounded to float32), scaled
    by 1/(1-0), quantised by 8191 and packed by twopoint_table().  `plant`
    introduces one known mistake so that the test can show it is caught."""
    f = np.float32
    o = rays[:, 0:3].astype(f)
    d = rays[:, 4:7].astype(f)
    with np.errstate(divide="ignore", invalid="ignore"):
        lo = np.where(d >= 0, (f(0) - o) / d, (f(1) - o) / d).astype(f)
        hi = np.where(d >= 0, (f(1) - o) / d, (f(0) - o) / d).astype(f)
    t0 = np.fmax(lo[:, 2], np.fmax(lo[:, 0], lo[:, 1]))
    t1 = np.fmin(hi[:, 2], np.fmin(hi[:, 0], hi[:, 1]))
    eps = f(0) if plant == "no_eps" else f(1.0e-3)
    box_t = np.where(t0 > t1, f(-1), (f(1) - eps) * np.where(t0 >= 0, t0, t1)).astype(f)
    tmax = rays[:, 7].astype(f)
    ln = np.full(rays.shape[0], f(ray_length), f)
    if plant != "no_tmax":
        ln = np.fmin(ln, np.fmax(tmax, f(0)))
    if plant != "no_box":
        ln = np.fmin(ln, box_t)
    b = (o + (ln[:, None] * d).astype(f)).astype(f)
    scale = f(1) / (f(1) - f(0))
    ia = _f2i(((o - f(0)) * scale).astype(f) * f(8191.0))
    ib = _f2i(((b - f(0)) * scale).astype(f) * f(8191.0))
    table = twopoint_table()
    if plant == "no_bit0":
        table = table[:-1]
    key = _pack(table, ia, ib)
    return key >> np.uint64(64 - bits)


def test_twopoint_table_is_a_bit_permutation():
    pos = [p for p, *_ in twopoint_table()]
    assert len(pos) == len(set(pos)) == 64 and set(pos) == set(range(64))


@pytest.mark.parametrize("ray_length", [0.25, 2.0])
def test_ray_keys_twopoint_closed_form(oracle, ray_length):
    rays = scenes.ray_key_edge_rays(30000, seed=48)
    for bits in (30, 60, 64):
        assert np.array_equal(oracle.ray_keys_twopoint64(rays, bits, ray_length),
                              twopoint_closed_form(rays, bits, ray_length))


def _hand_key(o, d, ln):
    """Key of one ray whose clamped length ln is worked out by hand."""
    f = np.float32
    o = np.asarray(o, f)
    d = np.asarray(d, f)
    b = (o + (f(ln) * d).astype(f)).astype(f)
    return _pack(twopoint_table(), _f2i(o[None] * f(8191.0)), _f2i(b[None] * f(8191.0)))[0]


def test_ray_keys_twopoint_each_clamp(oracle):
    """One ray per clamp of the length: rayLength, tmax, tmax < 0, the box exit
    (with its 1e-3 shortening), a ray entering the box from outside, and one
    missing it (boxT = -1 gives len = -1)."""
    f = np.float32
    eps = f(1) - f(1.0e-3)
    cases = [  # origin, direction, tmax, rayLength, expected len
        ((0.5, 0.5, 0.5), (1, 0, 0), 10.0, 0.25, f(0.25)),                       # rayLength
        ((0.5, 0.5, 0.5), (0, 1, 0), 0.125, 0.25, f(0.125)),                     # tmax
        ((0.5, 0.5, 0.5), (0, 0, 1), -1.0, 0.25, f(0.0)),                        # tmax < 0 -> 0
        ((0.9, 0.5, 0.5), (1, 0, 0), 10.0, 0.25, eps * (f(1) - f(0.9))),         # box exit (far plane)
        ((0.5, 0.25, 0.5), (0, -1, 0), 10.0, 2.0, eps * ((f(0) - f(0.25)) / f(-1))),  # box exit, negative d
        ((-1.0, 0.5, 0.5), (1, 0, 0), 10.0, 5.0, eps * f(1.0)),                  # enters from outside: near plane
        ((-1.0, 2.0, 0.5), (1, 0, 0), 10.0, 0.25, f(-1.0)),                     # misses the box
    ]
    rays = np.zeros((len(cases), 8), np.float32)
    for i, (o, d, tmax, _, _) in enumerate(cases):
        rays[i, 0:3], rays[i, 4:7], rays[i, 7] = o, d, tmax
    for i, (o, d, tmax, rl, ln) in enumerate(cases):
        want = _hand_key(o, d, ln)
        got = oracle.ray_keys_twopoint64(rays[i:i + 1], 64, rl)[0]
        assert got == want, (i, hex(int(got)), hex(int(want)))
        assert twopoint_closed_form(rays[i:i + 1], 64, rl)[0] == want


@pytest.mark.parametrize("plant", ["no_eps", "no_tmax", "no_box", "no_bit0"])
def test_ray_keys_twopoint_planted_mistakes_are_caught(oracle, plant):
    """A restatement with one step dropped disagrees with the oracle, so the
    closed-form test above would catch that mistake in the oracle."""
    rays = scenes.ray_key_edge_rays(30000, seed=48)
    assert not np.array_equal(oracle.ray_keys_twopoint64(rays, 64, 0.25),
                              twopoint_closed_form(rays, 64, 0.25, plant=plant))
