"""Ray generation, occlusion reconstruction and interpolation on the GPU
against the oracle: seeds, primary and shadow rays, reconstruction and
interpolation bit for bit; AO and path directions (cosf/sinf/powf) within a
few float32 ulps; a whole ambient-occlusion frame (primary rays -> trace -> AO
rays -> trace -> reconstruction) against the oracle's frame."""
import numpy as np
import pytest
import torch

import hippie_amd as hx
from hippie_amd import raygen as RG
from hippie_amd import scenes

pytestmark = pytest.mark.gpu


def _d(x, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    return t if dtype is None else t.to(dtype)


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def soup():
    v, t = scenes.uniform_soup(20000, seed=4)
    return v, t


@pytest.mark.parametrize("n,frame", [(1, 0), (1000, 7), (300001, 2)])
def test_seeds(oracle, ctx, n, frame):
    got = RG.init_seeds(ctx, n, frame).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, oracle.init_seeds(n, frame))


@pytest.mark.parametrize("sample", [0, 3])
@pytest.mark.parametrize("size", [(64, 48), (333, 97)])
def test_primary_rays(oracle, ctx, sample, size):
    cam = RG.Camera(position=(0.5, 0.5, -1.5), direction=(0.0, 0.0, 1.0), fov=60.0, size=size, far=3.0)
    rays, itp = RG.RayGen(ctx).primary(cam, sample)
    exp = oracle.primary_rays(sample, cam.position, cam.screen_to_world(), size[0], size[1], cam.far,
                              itp.cpu().numpy())
    assert _bits_equal(rays.cpu().numpy(), exp)


def _first_hits(ctx, v, t, w=128, h=96):
    scene = hx.Scene(v, t)
    bvh = hx.LBVHBuilder(ctx, morton_bits=30).build(scene)
    cam = RG.Camera(position=(0.5, 0.5, -1.5), direction=(0.0, 0.0, 1.0), fov=40.0, size=(w, h), far=3.0)
    rays, itp = RG.RayGen(ctx).primary(cam)
    rb = hx.RayBuffer(rays)
    res = hx.HipTracer(ctx, bvh).trace(rb)
    return scene, bvh, rays, itp, res


def test_shadow_rays(oracle, ctx, soup):
    v, t = soup
    scene, bvh, rays, itp, res = _first_hits(ctx, v, t)
    n = rays.shape[0]
    seeds = RG.init_seeds(ctx, n, 1)
    light = (0.48, 1.5, 0.17)
    for begin, size, ns in ((0, n, 2), (100, 500, 5)):
        out = RG.RayGen(ctx).shadow(rays, res, seeds, ns, light, 0.05,
                                    begin=begin, size=size)
        exp = oracle.shadow_rays(begin, size, ns, seeds.cpu().numpy().view(np.uint32), 0.05,
                                 np.array(light, np.float32), rays.cpu().numpy(), res.cpu().numpy())
        assert _bits_equal(out.cpu().numpy(), exp), (begin, size, ns)


def test_ao_rays(oracle, ctx, soup):
    v, t = soup
    scene, bvh, rays, itp, res = _first_hits(ctx, v, t)
    n = rays.shape[0]
    seeds = RG.init_seeds(ctx, n, 5)
    out = RG.RayGen(ctx).ao(rays, res, seeds, 8, 0.25, scene).cpu().numpy()
    exp = oracle.ao_rays(0, n, 8, seeds.cpu().numpy().view(np.uint32), 0.25, rays.cpu().numpy(),
                         res.cpu().numpy(), v, t)
    assert _bits_equal(out[:, [0, 1, 2, 3, 7]], exp[:, [0, 1, 2, 3, 7]])     # origins, tmin, tmax
    err = np.abs(out[:, 4:7] - exp[:, 4:7]).max()
    assert err < 2e-6, err                                                     # cosf/sinf of two libraries


def test_path_rays(oracle, ctx, soup):
    v, t = soup
    scene, bvh, rays, itp, res = _first_hits(ctx, v, t)
    rng = np.random.default_rng(8)
    nrm = rng.normal(size=v.shape).astype(np.float32)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    mats = np.zeros(4, RG.MATERIAL_DTYPE)
    mats["diffuse"] = [[0.7, 0.6, 0.5], [0.2, 0.2, 0.2], [0.5, 0.1, 0.1], [0.1, 0.3, 0.9]]
    mats["specular"] = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.3, 0.3, 0.3], [0.9, 0.9, 0.9]]
    mats["shininess"] = [1.0, 20.0, 5.0, 60.0]
    mats["tex_index"] = -1
    mi = rng.integers(0, 4, t.shape[0]).astype(np.int32)
    n = rays.shape[0]
    for rr in (False, True):
        seeds0 = RG.init_seeds(ctx, n, 11)
        seeds_h = seeds0.cpu().numpy().view(np.uint32).copy()
        dec = torch.ones((n, 3), dtype=torch.float32, device="cuda")
        out, op = RG.RayGen(ctx).path(rays, res, seeds0, itp, _d(nrm), scene.triangles, _d(mi), mats, dec,
                                      russian_roulette=rr)
        e_out, e_op, e_seeds, e_dec = oracle.path_rays(rr, seeds_h, itp.cpu().numpy(), mi, t, nrm, mats,
                                                       rays.cpu().numpy(), res.cpu().numpy(),
                                                       np.ones((n, 3), np.float32))
        assert np.array_equal(op.cpu().numpy(), e_op)                          # survivors, in input order
        assert np.array_equal(seeds0.cpu().numpy().view(np.uint32), e_seeds)
        got = out.cpu().numpy()
        assert _bits_equal(got[:, [3, 7]], e_out[:, [3, 7]])
        derr = np.abs(got[:, 4:7] - e_out[:, 4:7]).max(axis=1)
        # powf/cosf/sinf may move a glossy sample across the cosine-lobe fallback on a few rays.
        flips = derr > 1e-3
        assert flips.mean() < 2e-3, flips.mean()
        assert derr[~flips].max() < 2e-5
        assert np.abs(got[~flips, 0:3] - e_out[~flips, 0:3]).max() < 1e-5
        dg = dec.cpu().numpy()
        pix = e_op[~flips]
        assert np.allclose(dg[pix], e_dec[pix], rtol=2e-4, atol=1e-6)


def test_occlusion_reconstruction_and_interpolation(oracle, ctx):
    rng = np.random.default_rng(9)
    npx, ns = 5000, 8
    itp = rng.permutation(npx).astype(np.int32)
    res = np.zeros((npx * ns, 4), np.int32)
    res[:, 0] = np.where(rng.random(npx * ns) < 0.3, -1, 5)
    inp = rng.random((npx, 4)).astype(np.float32)
    outp = rng.random((npx, 4)).astype(np.float32)
    slot = rng.permutation(npx * ns).astype(np.int32)
    for replace, its in ((True, None), (False, slot)):
        g_out = _d(outp)
        RG.reconstruct_occlusion(ctx, _d(res), _d(itp), ns, _d(inp), g_out, begin=100, size=4000, replace=replace,
                                 index_to_slot=None if its is None else _d(its))
        exp = oracle.reconstruct_occlusion(100, 4000, ns, replace, res, itp, inp, outp, index_to_slot=its)
        assert _bits_equal(g_out.cpu().numpy(), exp)
    a = rng.normal(size=(70000, 3)).astype(np.float32)
    b = rng.normal(size=(70000, 3)).astype(np.float32)
    out0 = rng.normal(size=(70100, 3)).astype(np.float32)
    for tt, nz in ((0.25, False), (0.7, True)):
        g = _d(out0)
        RG.interpolate_vertices(ctx, _d(a), _d(b), tt, g, n_static=100, normalize=nz)
        assert _bits_equal(g.cpu().numpy(), oracle.interpolate(100, a, b, np.float32(tt), out0, normalize=nz))


def test_ambient_occlusion_frame(oracle, ctx, soup):
    """Primary rays -> trace -> AO rays -> trace -> reconstruction, against the
    oracle's frame from the same primary hits; pixels whose AO samples' hits
    differ (a direction a few ulps off grazing an edge) are counted, not failed."""
    v, t = soup
    scene, bvh, rays, itp, res = _first_hits(ctx, v, t, 160, 120)
    n, ns = rays.shape[0], 4
    seeds = RG.init_seeds(ctx, n, 0)
    ao = RG.RayGen(ctx).ao(rays, res, seeds, ns, 0.1, scene)
    ab = hx.RayBuffer(ao)
    ares = hx.HipTracer(ctx, bvh).trace(ab)
    white = torch.ones((n, 4), dtype=torch.float32, device="cuda")
    pix = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    RG.reconstruct_occlusion(ctx, ares, itp, ns, white, pix)
    nodes, woop, tri = bvh.nodes_host(), bvh.woop_host(), bvh.tri_indices_host()
    e_rays = oracle.ao_rays(0, n, ns, seeds.cpu().numpy().view(np.uint32), 0.1, rays.cpu().numpy(),
                            res.cpu().numpy(), v, t)
    e_res, _ = oracle.trace(nodes, woop, tri, e_rays)
    e_pix = oracle.reconstruct_occlusion(0, n, ns, True, e_res, itp.cpu().numpy(), np.ones((n, 4), np.float32),
                                         np.zeros((n, 4), np.float32))
    got = pix.cpu().numpy()
    differ = np.any(got != e_pix, axis=1)
    assert differ.mean() < 1e-3, differ.mean()
    hits = res.cpu().numpy()[:, 0] >= 0
    assert 0.0 < got[itp.cpu().numpy()[hits], 0].mean() < 1.0                # some occlusion on hit pixels
