"""The FP64 adjoint reference used by the GPU tests, pinned on the CPU by
central differences of the ray-plane hit distance."""
import numpy as np
import torch

from adjoint_ref import adjoint_terms_fp64, assert_scatter_within_bound


def _hit_t(v0, v1, v2, o, d):
    e1, e2 = v0 - v2, v1 - v2
    n = np.cross(e1, e2)
    return np.dot(v2 - o, n) / np.dot(d, n)


def test_adjoint_reference_matches_central_differences():
    rng = np.random.default_rng(0)
    verts = rng.random((5, 3)) * 2 - 1
    verts[:, 2] += 3.0
    tris = np.array([[0, 1, 2], [2, 3, 4], [4, 1, 0]], np.int64)
    R = 40
    o = np.zeros((R, 3))
    tri_of = rng.integers(0, 3, R)
    d = np.empty((R, 3))
    for r in range(R):
        a, b = rng.random(2) * 0.5 + 0.1
        v0, v1, v2 = verts[tris[tri_of[r]]]
        p = v2 + a * (v0 - v2) + b * (v1 - v2)
        d[r] = p / np.linalg.norm(p)
    dl_dt = rng.random(R) + 0.5
    rays = torch.zeros((R, 8), dtype=torch.float64)
    rays[:, 0:3] = torch.as_tensor(o)
    rays[:, 4:7] = torch.as_tensor(d)
    res = torch.zeros((R, 4), dtype=torch.int32)
    res[:, 0] = torch.as_tensor(tri_of, dtype=torch.int32)
    g, S, K = adjoint_terms_fp64(torch.as_tensor(verts), torch.as_tensor(tris), rays, res, torch.as_tensor(dl_dt))

    def loss(vv):
        return sum(dl_dt[r] * _hit_t(*vv[tris[tri_of[r]]], o[r], d[r]) for r in range(R))

    h = 1e-6
    fd = np.zeros(15)
    for i in range(15):
        vp, vm = verts.copy(), verts.copy()
        vp.flat[i] += h
        vm.flat[i] -= h
        fd[i] = (loss(vp) - loss(vm)) / (2 * h)
    assert np.allclose(g.numpy(), fd, rtol=1e-6, atol=1e-8)
    # K counts the rays of each vertex's triangles; S bounds |g|.
    assert np.all(S.numpy() >= np.abs(g.numpy()) - 1e-12)
    cnt = np.zeros(5, np.int64)
    for r in range(R):
        for vid in tris[tri_of[r]]:
            cnt[vid] += 1
    assert np.array_equal(K.numpy().reshape(5, 3), np.repeat(cnt[:, None], 3, axis=1))
    # The bound accepts float32 reorderings and rejects a wrong gradient.
    assert_scatter_within_bound(g.float(), g, S, K)
    bad = g.float().clone()
    bad[4] += 1e-3 * (abs(float(bad[4])) + 1)
    try:
        assert_scatter_within_bound(bad, g, S, K)
    except AssertionError:
        pass
    else:
        raise AssertionError("a wrong coordinate passed the bound")
