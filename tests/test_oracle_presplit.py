"""Pinning the oracle's presplitter (Presplitter.cpp:124-326, PresplitterKernels.cu:11-419).

Independent checks:
- the power used for the priority agrees with numpy's float64 power to one
  float32 ulp over eight decades;
- priorities equal a float64 numpy evaluation of (2^i (A_box - A_ideal))^(1/3)
  with i from the highest differing bit of the 63-bit corner codes (computed
  with Python integers), and a hand-computed triangle;
- one split of a hand-computed triangle gives the boxes worked out by hand;
- D leaves the rounded budget within beta * n and each triangle gets
  floor(D p) + 1 references;
- the references of every triangle lie inside its box and cover it: points
  sampled on the triangle each fall in one of its reference boxes;
- builds over presplit references validate and trace like brute force.
"""
import numpy as np
import pytest

from hippie_amd import scenes

F32 = np.float32


def _scene():
    """Large triangles across the box (a coarse grid) among a small-triangle soup."""
    return scenes.merge(scenes.gaussian_mixture_soup(3000, seed=4), scenes.ground_grid(4))


def test_pow_matches_float64(oracle):
    rng = np.random.default_rng(1)
    xs = np.concatenate([10.0 ** rng.uniform(-7, 1, 2000), [1.0, 0.5, 8.0, 2.0 ** -21, 3.0]]).astype(F32)
    y = F32(1.0) / F32(3.0)
    for x in xs:
        got = F32(oracle.ps_pow(x, y))
        exp = F32(np.float64(x) ** np.float64(y))
        assert abs(int(got.view(np.int32)) - int(exp.view(np.int32))) <= 1, x
    assert oracle.ps_pow(0.0, y) == 0.0


def _code63(q):
    c = 0
    for i in range(21):
        for a in range(3):
            c |= ((int(q[a]) >> i) & 1) << (3 * i + a)
    return c


def _priority64(v0, v1, v2, sb):
    mn, mx = np.minimum(np.minimum(v0, v1), v2), np.maximum(np.maximum(v0, v1), v2)
    smn, d = sb[:3].astype(np.float64), (sb[3:] - sb[:3]).astype(np.float64)
    qmn = np.floor((mn - smn) / d * 2097151.9).astype(np.int64)
    qmx = np.floor((mx - smn) / d * 2097151.9).astype(np.int64)
    x = _code63(qmn) ^ _code63(qmx)
    diff = x.bit_length() - 1
    i = int(diff / 3) - 21 if diff >= 0 else -21
    e = (mx - mn).astype(np.float64)
    area = 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0])
    c = np.cross((v0 - v1).astype(np.float64), (v0 - v2).astype(np.float64))
    return (2.0 ** i * (area - np.abs(c).sum())) ** (1.0 / 3.0)


def test_priorities_closed_form(oracle):
    v, t = _scene()
    sb = oracle.scene_box(v, t)
    ps = oracle.presplit(v, t, 0.3)
    p = ps["priorities"]
    for k in range(0, t.shape[0], 7):
        e = _priority64(v[t[k, 0]], v[t[k, 1]], v[t[k, 2]], sb)
        assert abs(p[k] - e) <= 2e-5 * max(e, 1e-3), k


def test_priority_hand_case(oracle):
    # Right triangle in the unit box: box area 2, ideal 1, codes differ first
    # at y's bit 20 (code bit 61): i = 61 // 3 - 21 = -1, p = 0.5^(1/3).
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [1, 1, 1]], F32)
    t = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    p = oracle.presplit(v, t, 0.0, scene_box=[0, 0, 0, 1, 1, 1])["priorities"]
    assert abs(p[0] - 0.5 ** (1.0 / 3.0)) <= 1e-6


def test_single_split_hand_case(oracle):
    # Box [0.1,0.9]^2 x {0.5}: the codes first differ at y's bit 20, so the cut
    # is y = 0.5; the left part keeps v0, v1 and the edge points (0.5, 0.5),
    # (0.1, 0.5); the right part keeps v2 and the same edge points.
    tv = [0.1, 0.1, 0.5, 0.9, 0.1, 0.5, 0.1, 0.9, 0.5]
    lb, lc, rb, rc = oracle.ps_split(tv, [0.1, 0.1, 0.5, 0.9, 0.9, 0.5], 5, [0, 0, 0, 1, 1, 1])
    assert np.allclose(lb, [0.1, 0.1, 0.5, 0.9, 0.5, 0.5], atol=1e-6)
    assert np.allclose(rb, [0.1, 0.5, 0.5, 0.5, 0.9, 0.5], atol=1e-6)
    # Both halves have extent, so the remaining budget 4 is halved (rounded).
    assert (lc, rc) == (2, 2)
    # A box whose corners share every code bit is cut at the middle of its
    # longest side.
    tv2 = [0.3, 0.3, 0.3, 0.3000001, 0.3, 0.3, 0.3, 0.3000002, 0.3]
    b2 = [0.3, 0.3, 0.3, 0.3000001, 0.3000002, 0.3]
    lb, lc, rb, rc = oracle.ps_split(tv2, b2, 1, [0, 0, 0, 1, 1, 1])
    mid = F32(F32(0.3) + F32(0.3000002)) * F32(0.5)
    assert lb[4] == mid and rb[1] == mid and (lc, rc) == (0, 0)


@pytest.mark.parametrize("beta", [0.1, 0.3, 1.0])
def test_budget_and_counts(oracle, beta):
    v, t = _scene()
    n = t.shape[0]
    ps = oracle.presplit(v, t, beta)
    D, p = F32(ps["D"]), ps["priorities"]
    counts = np.array([int(F32(D * x)) for x in p])
    smax = int(F32(beta) * F32(n))
    assert counts.sum() > 0
    assert counts.sum() <= smax
    assert ps["ref_tri"].shape[0] == n + counts.sum()
    assert np.array_equal(np.bincount(ps["ref_tri"], minlength=n), counts + 1)
    # References are grouped by triangle, in triangle order.
    assert np.all(np.diff(ps["ref_tri"]) >= 0)


@pytest.mark.parametrize("kind", ["mixed", "slivers"])
def test_references_cover_their_triangles(oracle, kind):
    # Slivers alone at beta 0.3 get no splits: their priority sum (~540) is
    # close to Smax = 600, so the reference's integer Smax / S gives D = 1 and
    # floor(p) = 0 everywhere; beta 1.0 gives them a budget.
    v, t = _scene() if kind == "mixed" else scenes.sliver_soup(2000, np.random.default_rng(3))
    ps = oracle.presplit(v, t, 0.3 if kind == "mixed" else 1.0)
    rt, rmin, rmax = ps["ref_tri"], ps["ref_min"][:, :3], ps["ref_max"][:, :3]
    many = np.flatnonzero(np.bincount(rt) > 1)
    assert many.size > 0
    g = np.linspace(0.0, 1.0, 13)
    bary = np.array([(a, b, 1 - a - b) for a in g for b in g if a + b <= 1.0 + 1e-9])
    for k in many[:40]:
        tv = v[t[k]].astype(np.float64)
        lo, hi = tv.min(0), tv.max(0)
        sel = rt == k
        bmn, bmx = rmin[sel].astype(np.float64), rmax[sel].astype(np.float64)
        assert np.all(bmn >= lo - 1e-7) and np.all(bmx <= hi + 1e-7)
        pts = bary @ tv
        eps = 1e-6
        inside = np.all((pts[:, None, :] >= bmn[None] - eps) & (pts[:, None, :] <= bmx[None] + eps), axis=2)
        assert np.all(inside.any(axis=1)), k


@pytest.mark.parametrize("builder", ["lbvh", "ploc", "atr", "hlbvh", "insertion"])
def test_presplit_builds_validate_and_trace(oracle, builder):
    v, t = _scene()
    ref = getattr(oracle, "build_" + builder)(v, t, presplit_beta=0.3)
    n_refs = ref["tri_idx"].shape[0]
    assert n_refs > t.shape[0]
    assert oracle.validate_bin_refs(ref["nodes"], ref["n_interior"], ref["n_leaves"], ref["tri_idx"], v, t) == 0
    rays = scenes.random_rays(3000, seed=8)
    hit, overflow = oracle.trace(ref["nodes"], ref["woop"], ref["tri_idx"], rays)
    brute = oracle.trace_brute(v, t, rays)
    assert overflow == 0
    assert np.array_equal(hit[:, 0], brute[:, 0]) and np.array_equal(hit[:, 1], brute[:, 1])
