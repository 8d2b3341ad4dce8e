"""Seeded batched workloads, generated on the device: many meshes and their rays.

Used by the benchmark and by the full-size tests, so that both run exactly the
same geometry and rays.

  batched_build_trace   4096 meshes of log-uniform 64..262,144 triangles (half
                        displaced icospheres topped up with soup, half soups),
                        each in the unit box; 256 rays per 1k triangles per
                        mesh from a sphere around the box aimed inside it.
  batched_presplit_diff 512 meshes of log-uniform 1k..200k triangles, 30% long
                        thin slivers (aspect up to 1:200); 64M rays over the
                        meshes in proportion to their triangles.
"""
import math

import numpy as np
import torch

from . import scenes


def batched_build_trace(device, meshes=4096, lo=64, hi=262144, rays_per_1k=256, seed=3):
    """Meshes of round(exp(U[ln lo, ln hi])) triangles; returns a dict with the
    concatenated vertices V and triangles T (global vertex indices), triangle
    offsets `offs` (M+1), vertex offsets `vbase` (M+1), per-mesh `counts`, the
    mesh-local triangles `tris_local`, rays (R, 8), the mesh of each ray
    `owner` and ray offsets `roffs` (rays are grouped by mesh)."""
    rng = np.random.default_rng(seed)
    M = meshes
    counts = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), M))).astype(np.int64)
    sphere = rng.random(M) < 0.5
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    ico = {}
    verts, tris_local, offs, vbase = [], [], [0], [0]

    def soup(c):
        cen = torch.rand((c, 1, 3), generator=g, device=device)
        sc = 0.001 + 0.019 * torch.rand((c, 1, 1), generator=g, device=device)
        v = cen + sc * (torch.rand((c, 3, 3), generator=g, device=device) - 0.5)
        return v.reshape(-1, 3), torch.arange(3 * c, device=device, dtype=torch.int32).reshape(-1, 3)

    for m in range(M):
        c = int(counts[m])
        if sphere[m] and c >= 20:
            level = 0
            while 20 * 4 ** (level + 1) <= c and level < 7:
                level += 1
            if level not in ico:
                v0, f0 = scenes.icosphere(level)
                ico[level] = (torch.as_tensor(v0, dtype=torch.float32, device=device),
                              torch.as_tensor(f0, dtype=torch.int32, device=device))
            v0, f0 = ico[level]
            k = torch.randint(1, 6, (3,), generator=g, device=device).float()
            ph = 6.283 * torch.rand(3, generator=g, device=device)
            r = 1.0 + 0.05 * torch.prod(torch.sin(k * v0 + ph), dim=1)
            v = 0.5 + 0.45 * v0 * r[:, None]
            f = f0
            extra = c - f0.shape[0]
            if extra > 0:
                sv, sf = soup(extra)
                f = torch.cat([f0, sf + v.shape[0]])
                v = torch.cat([v, sv])
        else:
            v, f = soup(c)
        verts.append(v)
        tris_local.append(f)
        offs.append(offs[-1] + c)
        vbase.append(vbase[-1] + v.shape[0])
    V = torch.cat(verts).contiguous()
    T = torch.cat([f + vb for f, vb in zip(tris_local, vbase[:-1])]).contiguous()
    # Rays: per mesh round(rays_per_1k * c / 1000) (at least 1), grouped by mesh.
    rc = np.maximum(1, np.rint(rays_per_1k * counts / 1000.0)).astype(np.int64)
    owner = torch.as_tensor(np.repeat(np.arange(M, dtype=np.int32), rc), device=device)
    R = int(rc.sum())
    u = torch.randn((R, 3), generator=g, device=device)
    u = u / u.norm(dim=1, keepdim=True)
    rad = 0.75 * math.sqrt(3.0) + 1e-3
    o = 0.5 + rad * u
    tgt = torch.rand((R, 3), generator=g, device=device)
    d = tgt - o
    d = d / d.norm(dim=1, keepdim=True)
    rays = torch.empty((R, 8), dtype=torch.float32, device=device)
    rays[:, 0:3] = o
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = 3.0 * rad
    roffs = np.concatenate([[0], np.cumsum(rc)])
    return dict(V=V, T=T, offs=np.array(offs, np.int64), vbase=np.array(vbase, np.int64), counts=counts,
                tris_local=tris_local, rays=rays, owner=owner, roffs=roffs)


def batched_presplit_diff(device, meshes=512, rays=1 << 26, lo=1000, hi=200000, sliver_fraction=0.3, seed=5):
    """Meshes of round(exp(U[ln lo, ln hi])) triangles, a `sliver_fraction` of
    them long thin slivers (scenes.sliver_soup), the rest config-3 meshes
    (scenes.batch_mesh); `rays` rays spread over the meshes in proportion to
    their triangles (at least one each), from a radius-1.5 sphere around the
    unit box aimed at uniform points inside it, grouped by mesh.  Returns a
    dict: host `meshes` [(vertices, triangles)], `counts`, `sliver` (which
    meshes are slivers), rays (R, 8) and `owner` on the device, and ray
    offsets `roffs`."""
    rng = np.random.default_rng(seed)
    M = meshes
    counts = np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), M))).astype(np.int64)
    out, sliver = [], np.zeros(M, dtype=bool)
    for m, c in enumerate(counts):
        if rng.random() < sliver_fraction:
            out.append(scenes.sliver_soup(int(c), rng))
            sliver[m] = True
        else:
            out.append(scenes.batch_mesh(int(c), rng, bool(rng.random() < 0.5)))
    N = int(counts.sum())
    per = np.maximum(1, np.rint(counts * (rays / N))).astype(np.int64)
    R = int(per.sum())
    owner = torch.repeat_interleave(torch.arange(M, device=device, dtype=torch.int32),
                                    torch.as_tensor(per, device=device))
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    d = torch.randn((R, 3), generator=g, device=device)
    o = 0.5 + 1.5 * d / d.norm(dim=1, keepdim=True)
    tgt = torch.rand((R, 3), generator=g, device=device)
    dirn = tgt - o
    dirn = dirn / dirn.norm(dim=1, keepdim=True)
    r = torch.zeros((R, 8), dtype=torch.float32, device=device)
    r[:, 0:3] = o
    r[:, 4:7] = dirn
    r[:, 7] = 4.0
    return dict(meshes=out, counts=counts, sliver=sliver, rays=r, owner=owner,
                roffs=np.concatenate([[0], np.cumsum(per)]))
