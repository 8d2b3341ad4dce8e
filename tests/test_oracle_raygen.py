"""Pins of the oracle's ray generation, occlusion reconstruction and vertex
interpolation (oracle/hx_oracle_raygen.cpp) by inverses, closed forms and
geometry computed independently here."""
from fractions import Fraction

import numpy as np
import pytest

from hippie_amd import raygen as RG

M32 = np.uint64(0xFFFFFFFF)


def _u32(x):
    return np.uint32(int(x) & 0xFFFFFFFF)


# --- hashes and low-discrepancy sequences ------------------------------------

def _tea_decrypt(v0, v1, rounds=16):
    """TEA decryption (Wheeler & Needham 1994) with the seed generator's key
    (0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e)."""
    k = (0xa341316c, 0xc8013ea4, 0xad90777d, 0x7e95761e)
    delta = 0x9e3779b9
    s = (delta * rounds) & 0xFFFFFFFF
    v0, v1 = int(v0), int(v1)
    for _ in range(rounds):
        v1 = (v1 - ((((v0 << 4) + k[2]) ^ (v0 + s) ^ ((v0 >> 5) + k[3])) & 0xFFFFFFFF)) & 0xFFFFFFFF
        v0 = (v0 - ((((v1 << 4) + k[0]) ^ (v1 + s) ^ ((v1 >> 5) + k[1])) & 0xFFFFFFFF)) & 0xFFFFFFFF
        s = (s - delta) & 0xFFFFFFFF
    return v0, v1


def test_tea16_is_undone_by_tea_decryption(oracle):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 2**32, 300, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, 300, dtype=np.uint64).astype(np.uint32)
    a[:4] = [0, 1, 0xFFFFFFFF, 7]
    b[:4] = [0, 0, 0xFFFFFFFF, 1 << 31]
    enc = oracle.tea16(a, b)
    for i in range(300):
        assert _tea_decrypt(enc[i, 0], enc[i, 1]) == (int(a[i]), int(b[i]))
    assert len(set(enc[:, 0].tolist())) == 300
    seeds = oracle.init_seeds(1000, 3)
    assert np.array_equal(seeds, oracle.tea16(np.arange(1000), np.full(1000, 3))[:, 0])


def _jenkins_inverse(a, b, c):
    m = 0xFFFFFFFF
    steps = [("a", "b", "c", 13, ">"), ("b", "c", "a", 8, "<"), ("c", "a", "b", 13, ">"),
             ("a", "b", "c", 12, ">"), ("b", "c", "a", 16, "<"), ("c", "a", "b", 5, ">"),
             ("a", "b", "c", 3, ">"), ("b", "c", "a", 10, "<"), ("c", "a", "b", 15, ">")]
    v = {"a": int(a), "b": int(b), "c": int(c)}
    # Each forward step is x -= y; x -= z; x ^= (z shift k), z being the
    # variable updated last; undo the steps in reverse.
    for x, y, z, k, d in reversed(steps):
        sh = (v[z] >> k) if d == ">" else ((v[z] << k) & m)
        v[x] ^= sh
        v[x] = (v[x] + v[z]) & m
        v[x] = (v[x] + v[y]) & m
    return v["a"], v["b"], v["c"]


def test_jenkins_mix_is_undone_by_its_inverse(oracle):
    rng = np.random.default_rng(2)
    abc = rng.integers(0, 2**32, (200, 3), dtype=np.uint64).astype(np.uint32)
    abc[0] = [0, 0x9e3779b9, 0x9e3779b9]
    out = oracle.jenkins_mix(abc)
    for i in range(200):
        assert _jenkins_inverse(*out[i]) == tuple(int(x) for x in abc[i])


def _radical_inverse(i, base):
    r, f = Fraction(0), Fraction(1, base)
    while i:
        r += f * (i % base)
        i //= base
        f /= base
    return r


def test_halton_is_the_radical_inverse(oracle):
    idx = np.arange(0, 5000, dtype=np.int32)
    for base in (2, 3, 5):
        h = oracle.halton(idx, base)
        exact = np.array([float(_radical_inverse(int(i), base)) for i in idx])
        if base == 2:
            assert np.array_equal(h.astype(np.float64), exact)          # every operation exact
        else:
            assert np.max(np.abs(h - exact)) < 4e-7


def test_sobol2d_first_points_form_nets(oracle):
    # Each coordinate of the first 2^m points is stratified in 2^m intervals,
    # and the pairs form a (0,m,2)-net: every 2^a x 2^b box (a + b = m) holds one point.
    for m in range(1, 11):
        n = 1 << m
        p = oracle.sobol2d(np.arange(n, dtype=np.int32)).astype(np.float64)
        assert np.all((p >= 0) & (p < 1))
        for a in range(m + 1):
            b = m - a
            cell = np.floor(p[:, 0] * (1 << a)).astype(np.int64) * (1 << b) + np.floor(p[:, 1] * (1 << b)).astype(np.int64)
            assert np.array_equal(np.sort(cell), np.arange(n)), (m, a)


# --- camera and primary rays --------------------------------------------------------

def _ndc_dir64(cam, px, py):
    """Direction of the ray through pixel corner (px, py) from float64 camera
    geometry: NDC (2x/w - 1, 2y/h - 1, 0) unprojected by inv(P V)."""
    P = cam.projection_matrix().astype(np.float64)
    V = cam.view_matrix().astype(np.float64)
    inv = np.linalg.inv(P @ V)
    w, h = cam.size
    q = inv @ np.array([2 * px / w - 1, 2 * py / h - 1, 0.0, 1.0])
    d = q[:3] / q[3] - cam.position.astype(np.float64)
    return d / np.linalg.norm(d)


def test_camera_matrices():
    cam = RG.Camera(position=(0.2, -0.1, 0.3), direction=(0.3, -0.2, 1.0), fov=60.0, size=(64, 48), far=5.0)
    d = np.array([0.3, -0.2, 1.0])
    assert np.allclose(cam.direction(), d / np.linalg.norm(d), atol=2e-6)
    V = cam.view_matrix().astype(np.float64)
    assert np.allclose(V @ np.r_[cam.position, 1.0], [0, 0, 0, 1], atol=1e-6)      # camera at the origin
    assert np.allclose((V[:3, :3] @ (d / np.linalg.norm(d))), [0, 0, -1], atol=2e-6)   # looking down -z
    M = cam.projection_view_matrix().astype(np.float64)
    inv = RG.inverse4(cam.projection_view_matrix()).astype(np.float64)
    assert np.allclose(inv @ M, np.eye(4), atol=1e-4)
    # fovY: the top and bottom edge rays through the image centre column meet at 60 degrees.
    a, b = _ndc_dir64(cam, 32, 0), _ndc_dir64(cam, 32, 48)
    assert abs(np.degrees(np.arccos(a @ b)) - 60.0) < 1e-3


@pytest.mark.parametrize("sample", [0, 1, 5])
def test_primary_rays_follow_the_pinhole(oracle, sample):
    cam = RG.Camera(position=(0.0, 0.0, 0.0), direction=(0.1, 0.2, -1.0), fov=45.0, size=(40, 30), far=3.0)
    itp, pti = RG.pixel_table(40, 30)
    rays = oracle.primary_rays(sample, cam.position, cam.screen_to_world(), 40, 30, 3.0, itp)
    hx_, hy_ = float(_radical_inverse(sample, 2)), float(_radical_inverse(sample, 3))
    for i in range(0, 1200, 7):
        p = int(itp[i])
        d = _ndc_dir64(cam, p % 40 + hx_, p // 40 + hy_)
        assert np.allclose(rays[i, 4:7], d, atol=1e-5), i
    assert np.all(rays[:, 0:4] == 0.0) and np.all(rays[:, 7] == np.float32(3.0))
    assert np.allclose(np.linalg.norm(rays[:, 4:7], axis=1), 1.0, atol=1e-6)


def test_pixel_table_is_a_permutation_in_tiles():
    for w, h in ((64, 64), (40, 30), (13, 7), (1024, 768)):
        itp, pti = RG.pixel_table(w, h)
        assert np.array_equal(np.sort(itp), np.arange(w * h))
        assert np.array_equal(itp[pti], np.arange(w * h))
    itp, _ = RG.pixel_table(64, 64)
    # Each run of 64 rays covers one 8x8 tile; tile order is Morton.
    tiles = (itp // 64 // 8) * 8 + (itp % 64) // 8
    assert np.all(tiles.reshape(-1, 64) == tiles.reshape(-1, 64)[:, :1])
    first = itp[::64]
    assert [(int(p) % 64 // 8, int(p) // 64 // 8) for p in first[:4]] == [(0, 0), (1, 0), (0, 1), (1, 1)]


# --- secondary rays ----------------------------------------------------------------

def _scene(rng, n=64):
    c = rng.random((n, 3)).astype(np.float32)
    v = (np.repeat(c, 3, 0) + (rng.random((3 * n, 3)) - 0.5) * 0.2).astype(np.float32)
    t = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return v, t


def _hits(rng, n, v, t, miss_every=5):
    rays = np.zeros((n, 8), np.float32)
    res = np.zeros((n, 4), np.int32)
    rays[:, 0:3] = rng.random((n, 3)) - 1.0
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1)[:, None]
    rays[:, 7] = 5.0
    res[:, 0] = rng.integers(0, t.shape[0], n)
    res[:, 1] = rng.random(n).astype(np.float32).view(np.int32)
    res[::miss_every, 0] = -1
    return rays, res


def test_shadow_rays_end_on_the_light(oracle):
    rng = np.random.default_rng(3)
    v, t = _scene(rng)
    rays, res = _hits(rng, 50, v, t)
    seeds = oracle.init_seeds(50, 0)
    light, radius, ns = np.array([0.5, 2.0, 0.25], np.float32), np.float32(0.1), 8
    out = oracle.shadow_rays(0, 50, ns, seeds, radius, light, rays, res)
    for task in range(50):
        o = out[task * ns:(task + 1) * ns]
        if res[task, 0] == -1:
            assert np.all(o[:, 0:3] == 0) and np.all(o[:, 7] == -1) and np.all(o[:, 4:7] == [1, 0, 0])
            continue
        tt = res[task, 1:2].view(np.float32)[0]
        org = rays[task, 0:3] + rays[task, 4:7] * np.float32(max(tt - np.float32(1e-5), 0))
        assert np.array_equal(o[:, 0:3], np.broadcast_to(org.astype(np.float32), (ns, 3)))
        end = o[:, 0:3].astype(np.float64) + o[:, 4:7] * o[:, 7:8]
        p = (end - light) / radius                         # points of [-1, 1]^3
        assert np.all(np.abs(p) <= 1 + 1e-4)
        # Undo the shift: the unshifted points are (sobol2d(i), hammersley(i, ns)).
        base = np.c_[oracle.sobol2d(np.arange(ns, dtype=np.int32)), (np.arange(ns) + 0.5) / ns]
        shift = np.mod((p + 1) / 2 - base, 1.0)
        shift = np.where(shift > 1 - 1e-4, shift - 1, shift)
        assert np.max(np.ptp(shift, axis=0)) < 1e-4      # one Cranley-Patterson shift per input ray


def test_ao_rays_are_halton_points_on_the_hemisphere(oracle):
    rng = np.random.default_rng(4)
    v, t = _scene(rng)
    rays, res = _hits(rng, 60, v, t)
    seeds = oracle.init_seeds(60, 1)
    ns = 8
    out = oracle.ao_rays(0, 60, ns, seeds, np.float32(0.25), rays, res, v, t)
    h2 = np.array([float(_radical_inverse(i + 1, 2)) for i in range(ns)])
    h3 = np.array([float(_radical_inverse(i + 1, 3)) for i in range(ns)])
    for task in range(60):
        o = out[task * ns:(task + 1) * ns]
        tt = res[task, 1:2].view(np.float32)[0]
        org = (rays[task, 0:3] + rays[task, 4:7] * np.float32(max(tt - np.float32(1e-4), 0))).astype(np.float32)
        assert np.array_equal(o[:, 0:3], np.broadcast_to(org, (ns, 3)))
        assert np.allclose(np.linalg.norm(o[:, 4:7], axis=1), 1.0, atol=2e-6)
        if res[task, 0] == -1:
            n = np.array([1.0, 0, 0])
            assert np.all(o[:, 7] == -1)
        else:
            a, b, c = v[t[res[task, 0]]].astype(np.float64)
            n = np.cross(b - a, c - a)
            n /= np.linalg.norm(n)
            assert np.all(o[:, 7] == np.float32(0.25))
        if n @ rays[task, 4:7] > 0:
            n = -n
        cz = o[:, 4:7] @ n
        assert np.allclose(cz, np.sqrt(1 - h2), atol=2e-5)       # r = sqrt(x): height sqrt(1 - x)
        # Azimuth differences are 2 pi (h3_i - h3_0), whatever the rotation.
        e1 = np.cross(n, [0.0, 0.0, 1.0] if abs(n[2]) < 0.9 else [1.0, 0.0, 0.0])
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        phi = np.arctan2(o[:, 4:7] @ e2, o[:, 4:7] @ e1)
        sign = np.sign(np.sin(phi[1] - phi[0])) * np.sign(np.sin(2 * np.pi * (h3[1] - h3[0])))
        dphi = np.mod(sign * (phi - phi[0]) - 2 * np.pi * (h3 - h3[0]) + np.pi, 2 * np.pi) - np.pi
        ok = h2 < 0.999                                  # a point at the pole has no azimuth
        assert np.max(np.abs(dphi[ok])) < 1e-3


def _lcg(seed):
    s = (1103515245 * int(seed) + 12345) & 0xFFFFFFFF
    return s, np.float32(s & 0xFFFFFF) / np.float32(0x01000000)


def test_path_rays_survival_seeds_and_diffuse_weights(oracle):
    rng = np.random.default_rng(5)
    v, t = _scene(rng, 40)
    normals = rng.normal(size=v.shape).astype(np.float32)
    normals /= np.linalg.norm(normals, axis=1)[:, None]
    n = 200
    rays, res = _hits(rng, n, v, t, miss_every=7)
    res[:, 2] = (rng.random(n) * 0.5).astype(np.float32).view(np.int32)
    res[:, 3] = (rng.random(n) * 0.5).astype(np.float32).view(np.int32)
    mats = np.zeros(3, RG.MATERIAL_DTYPE)
    mats["diffuse"] = [[0.8, 0.5, 0.2], [0.3, 0.3, 0.3], [0.9, 0.9, 0.9]]
    mats["shininess"] = 10.0
    mats["tex_index"] = -1                               # purely diffuse: specular 0
    mi = rng.integers(0, 3, t.shape[0]).astype(np.int32)
    seeds = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    pix = rng.permutation(n).astype(np.int32)
    dec = np.ones((n, 3), np.float32)
    for rr in (False, True):
        out, op, seeds2, dec2 = oracle.path_rays(rr, seeds, pix, mi, t, normals, mats, rays, res, dec)
        live = []
        for task in range(n):
            if res[task, 0] < 0:
                assert seeds2[task] == seeds[task]
                continue
            s = int(seeds[task])
            s, r0 = _lcg(s)
            s, r1 = _lcg(s)
            s, r2 = _lcg(s)
            assert seeds2[task] == s
            kd = mats["diffuse"][mi[res[task, 0]]].astype(np.float32)
            pd = np.float32(kd.max())
            if pd > 1:
                kd, pd = (kd * (np.float32(1) / pd)).astype(np.float32), np.float32(pd * (np.float32(1) / pd))
            total = pd if rr else np.float32(pd * (np.float32(1) / pd))
            if r2 <= total:
                live.append(task)
                w = kd / np.float32(total)
                k = len(live) - 1
                assert np.allclose(dec2[pix[task]], w, rtol=1e-5), task
        assert np.array_equal(op, pix[live])
        for k, task in enumerate(live):
            tt = res[task, 1:2].view(np.float32)[0]
            hitp = rays[task, 0:3] + rays[task, 4:7] * tt
            d = out[k, 4:7].astype(np.float64)
            assert abs(np.linalg.norm(d) - 1) < 1e-5
            uu, vv = res[task, 2:4].view(np.float32)
            a, b, c = normals[t[res[task, 0]]]
            nn = uu * a + vv * b + (1 - uu - vv) * c
            nn /= np.linalg.norm(nn)
            if nn @ rays[task, 4:7] > 0:
                nn = -nn
            assert d @ nn >= -1e-6
            assert np.allclose(out[k, 0:3], hitp + d * 1e-4, atol=1e-6)
            assert out[k, 3] == 0 and out[k, 7] == np.float32(3.402823466e38)
        if not rr:
            assert len(live) == int((res[:, 0] >= 0).sum())   # normalised: every diffuse path continues


def test_occlusion_reconstruction_counts_misses(oracle):
    rng = np.random.default_rng(6)
    npx, ns = 30, 4
    itp = rng.permutation(npx).astype(np.int32)
    res = np.zeros((npx * ns, 4), np.int32)
    res[:, 0] = np.where(rng.random(npx * ns) < 0.4, -1, rng.integers(0, 9, npx * ns))
    inp = rng.random((npx, 4)).astype(np.float32)
    outp = rng.random((npx, 4)).astype(np.float32)
    for replace in (True, False):
        got = oracle.reconstruct_occlusion(0, npx, ns, replace, res, itp, inp, outp)
        exp = outp.copy()
        for task in range(npx):
            f = np.float32((res[task * ns:(task + 1) * ns, 0] == -1).sum()) / np.float32(ns)
            v = np.r_[(inp[itp[task], :3] * f).astype(np.float32), np.float32(1)]
            exp[itp[task]] = v if replace else (exp[itp[task]] + v).astype(np.float32)
        assert np.array_equal(got, exp)
    slot = rng.permutation(npx * ns).astype(np.int32)
    got = oracle.reconstruct_occlusion(5, 10, ns, True, res, itp, inp, outp, index_to_slot=slot)
    for task in range(10):
        f = np.float32((res[slot[task * ns:(task + 1) * ns], 0] == -1).sum()) / np.float32(ns)
        assert np.array_equal(got[itp[5 + task], :3], (inp[itp[5 + task], :3] * f).astype(np.float32))


def test_interpolation_is_float32_mix(oracle):
    rng = np.random.default_rng(7)
    a = rng.normal(size=(100, 3)).astype(np.float32)
    b = rng.normal(size=(100, 3)).astype(np.float32)
    out = np.full((110, 3), 9.0, np.float32)
    for t in (np.float32(0.0), np.float32(0.3), np.float32(1.0)):
        got = oracle.interpolate(10, a, b, t, out)
        exp = (a * (np.float32(1) - t)).astype(np.float32) + (b * t).astype(np.float32)
        assert np.array_equal(got[10:], exp) and np.all(got[:10] == 9.0)
        gn = oracle.interpolate(10, a, b, t, out, normalize=True)
        ln = np.sqrt(((exp[:, 0] * exp[:, 0] + exp[:, 1] * exp[:, 1]).astype(np.float32)
                      + exp[:, 2] * exp[:, 2]).astype(np.float32)).astype(np.float32)
        assert np.array_equal(gn[10:], ((np.float32(1) / ln)[:, None] * exp).astype(np.float32))


def test_interpolator_frame_pairs():
    it = RG.Interpolator(None, [None] * 5, frame_rate=4.0, looped=True)
    seen = [it.frame_pair() for _ in range(4)]
    assert [(a, b) for a, b, _ in seen] == [(1, 2), (2, 3), (3, 4), (0, 1)]
    assert all(abs(t) < 1e-6 for _, _, t in seen)
    it2 = RG.Interpolator(None, [None] * 3, frame_rate=10.0, looped=False)
    for _ in range(30):
        a, b, t = it2.frame_pair()
    assert (a, b) == (1, 2) and abs(float(t) - 1.0) < 1e-4
