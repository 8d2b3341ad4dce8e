"""Seeded synthetic scenes and rays for the benchmark configurations.

Harness counterpart of the reference's scene loading and ray generation
(src/hippie/rt/scene/SceneLoader.cpp, rt/ray/RayGenKernels.cu:106-153,232-349):
geometry and rays are produced once on the host, in float32, from numpy's
PCG64 generator, and fed identically to the GPU path and to the oracle.
"""
import math

import numpy as np


def _u(rng, shape, lo=0.0, hi=1.0):
    return (np.float32(lo) + np.float32(hi - lo) * rng.random(shape, dtype=np.float32)).astype(np.float32)


def soup_triangles(centroids, rng, smin=0.001, smax=0.02):
    """Triangle soup around centroids: v_k = c + s (u_k - 0.5), s ~ U[smin, smax]."""
    n = centroids.shape[0]
    s = _u(rng, (n, 1, 1), smin, smax)
    u = rng.random((n, 3, 3), dtype=np.float32)
    v = centroids[:, None, :].astype(np.float32) + s * (u - np.float32(0.5))
    verts = v.reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return verts, tris


def uniform_soup(n, seed=1):
    """Config 1 geometry: n triangles, centroids U[0,1]^3, edge scale U[0.001,0.02]."""
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3), dtype=np.float32)
    return soup_triangles(c, rng)


def gaussian_mixture_soup(n, components=64, seed=2):
    """Triangles whose centroids come from a Gaussian mixture (mu ~ U[0,1]^3, sigma ~ U[0.01,0.1])."""
    rng = np.random.default_rng(seed)
    mu = rng.random((components, 3), dtype=np.float32)
    sigma = _u(rng, (components,), 0.01, 0.1)
    comp = rng.integers(0, components, n)
    c = mu[comp] + sigma[comp][:, None] * rng.standard_normal((n, 3), dtype=np.float32)
    return soup_triangles(c.astype(np.float32), rng)


def ground_grid(cells):
    """cells x cells quads (2 triangles each) at y = 0 over [0,1]^2 in x, z; shared vertices."""
    g = np.linspace(0.0, 1.0, cells + 1, dtype=np.float32)
    xx, zz = np.meshgrid(g, g, indexing="xy")
    verts = np.stack([xx.ravel(), np.zeros(xx.size, np.float32), zz.ravel()], axis=1).astype(np.float32)
    i = np.arange(cells, dtype=np.int32)
    ix, iz = np.meshgrid(i, i, indexing="xy")
    v00 = (iz * (cells + 1) + ix).ravel()
    v10 = v00 + 1
    v01 = v00 + (cells + 1)
    v11 = v01 + 1
    t0 = np.stack([v00, v10, v11], axis=1)
    t1 = np.stack([v00, v11, v01], axis=1)
    tris = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.int32)
    return verts, tris


def merge(*meshes):
    verts, tris, off = [], [], 0
    for v, t in meshes:
        verts.append(v)
        tris.append(t + off)
        off += v.shape[0]
    return np.concatenate(verts).astype(np.float32), np.concatenate(tris).astype(np.int32)


def config2_scene(n_mixture=8388608, grid_cells=2048, seed=2):
    """Config 2: Gaussian-mixture soup plus a ground-plane grid (n = n_mixture + 2 cells^2)."""
    return merge(gaussian_mixture_soup(n_mixture, seed=seed), ground_grid(grid_cells))


def _camera_basis(forward, up=(0.0, 1.0, 0.0)):
    f = np.asarray(forward, np.float64)
    f = f / np.linalg.norm(f)
    r = np.cross(np.asarray(up, np.float64), f)
    if np.linalg.norm(r) < 1e-8:
        r = np.cross((1.0, 0.0, 0.0), f)
    r = r / np.linalg.norm(r)
    u = np.cross(f, r)
    return r, u, f


def pinhole_rays(width, height, origin, forward, fov_y_deg=60.0, tmin=0.0, tmax=3.0):
    """Primary rays through pixel corners (Halton sample 0, RayGenKernels.cu:128-132).

    Returns float32 (width*height, 8): origin, tmin, normalised direction, tmax;
    ray index = y * width + x.
    """
    r, u, f = _camera_basis(forward)
    tan = math.tan(math.radians(fov_y_deg) * 0.5)
    aspect = width / height
    x = np.arange(width, dtype=np.float64)
    y = np.arange(height, dtype=np.float64)
    xx, yy = np.meshgrid(x, y, indexing="xy")
    nx = (2.0 * xx / width - 1.0) * tan * aspect
    ny = (2.0 * yy / height - 1.0) * tan
    d = nx[..., None] * r + ny[..., None] * u + f
    d = d / np.linalg.norm(d, axis=-1, keepdims=True)
    n = width * height
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = np.asarray(origin, np.float32)
    rays[:, 3] = tmin
    rays[:, 4:7] = d.reshape(n, 3).astype(np.float32)
    rays[:, 7] = tmax
    return rays


def config1_rays(width=1024, height=1024):
    """Config 1: pinhole at (0.5,0.5,-1.5) looking +z, 60 deg vertical FOV, tmax 3.0."""
    return pinhole_rays(width, height, (0.5, 0.5, -1.5), (0.0, 0.0, 1.0), 60.0, 0.0, 3.0)


def config2_primary_rays(width=4096, height=4096):
    """Config 2 primaries: camera inside the unit box (the Aila key assumes [0,1]^3 origins)."""
    return pinhole_rays(width, height, (0.5, 0.6, 0.02), (0.0, -0.35, 1.0), 60.0, 0.0, 3.0)


def secondary_rays(verts, tris, rays, results, seed=3, offset=1e-4, tmax=3.0):
    """Cosine-hemisphere diffuse rays from primary hits (A20 conventions).

    origin = hit - offset * direction; the geometric normal is flipped to face
    the incoming ray; misses become dead rays (tmax = -1).
    """
    rng = np.random.default_rng(seed)
    n = rays.shape[0]
    ids = results[:, 0]
    hit = ids >= 0
    t = results[:, 1].view(np.float32)
    o = rays[:, 0:3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    out = np.empty((n, 8), np.float32)
    out[:] = rays
    out[~hit, 7] = -1.0
    idx = np.nonzero(hit)[0]
    tri = tris[ids[idx]]
    v0, v1, v2 = verts[tri[:, 0]].astype(np.float64), verts[tri[:, 1]].astype(np.float64), verts[tri[:, 2]].astype(
        np.float64)
    nrm = np.cross(v1 - v0, v2 - v0)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    flip = np.sum(nrm * d[idx], axis=1) > 0
    nrm[flip] *= -1.0
    p = o[idx] + (t[idx].astype(np.float64) - offset)[:, None] * d[idx]
    r1 = rng.random(idx.size)
    r2 = rng.random(idx.size)
    phi = 2.0 * np.pi * r1
    sr = np.sqrt(r2)
    lx, ly, lz = sr * np.cos(phi), sr * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - r2))
    a = np.where(np.abs(nrm[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    tx = np.cross(a, nrm)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(nrm, tx)
    dirs = lx[:, None] * tx + ly[:, None] * ty + lz[:, None] * nrm
    out[idx, 0:3] = p.astype(np.float32)
    out[idx, 3] = 0.0
    out[idx, 4:7] = dirs.astype(np.float32)
    out[idx, 7] = tmax
    return out


def ray_key_edge_rays(n, seed):
    """Random rays plus the cases the ray keys' slab test special-cases:
    origins outside the unit box, rays leaving or missing it, axis-parallel
    directions (zero components), origins on box faces, negative and zero tmax."""
    rng = np.random.default_rng(seed)
    r = random_rays(n, seed=seed).copy()
    k = n // 6
    r[:k, 0:3] = rng.uniform(-1.0, 2.0, (k, 3))
    ax = rng.integers(0, 3, 2 * k)
    r[k:3 * k, 4:7] = 0.0
    r[np.arange(k, 3 * k), 4 + ax] = rng.choice([-1.0, 1.0], 2 * k)
    r[3 * k:4 * k, rng.integers(0, 3)] = rng.choice([0.0, 1.0], k)
    r[4 * k:5 * k, 7] = rng.uniform(-1.0, 0.3, k)
    r[5 * k:5 * k + 16, 7] = 0.0
    return r.astype(np.float32)


def icosphere(level):
    """Unit icosphere: 20 * 4^level triangles, 10 * 4^level + 2 shared vertices."""
    p = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
                  [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv + v.shape[0]
        nf = f.shape[0]
        a, b, c = m[:nf], m[nf:2 * nf], m[2 * nf:]
        v = np.concatenate([v, mid])
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    return v, f.astype(np.int32)


def value_noise3(p, seed, freq=4.0):
    """Seeded 3D value noise in [-1, 1]: lattice values, smoothstep trilinear blend."""
    rng = np.random.default_rng(seed)
    size = 64
    lattice = rng.uniform(-1.0, 1.0, (size, size, size))
    q = (p + 2.0) * freq
    i0 = np.floor(q).astype(np.int64)
    fr = q - i0
    s = fr * fr * (3.0 - 2.0 * fr)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = ((s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) *
                     (s[:, 2] if dz else 1 - s[:, 2]))
                out += w * lattice[(i0[:, 0] + dx) % size, (i0[:, 1] + dy) % size, (i0[:, 2] + dz) % size]
    return out


def displaced_icosphere(level=8, seed=1, amplitude=0.05):
    """Config 4 mesh: icosphere with radial displacement amplitude * noise(seed)."""
    v, f = icosphere(level)
    r = 1.0 + amplitude * value_noise3(v, seed)
    return (v * r[:, None]).astype(np.float32), f


def fibonacci_cameras(count=16, radius=3.0):
    """Camera positions on a Fibonacci sphere, all looking at the origin."""
    i = np.arange(count) + 0.5
    phi = np.arccos(1.0 - 2.0 * i / count)
    theta = np.pi * (1.0 + 5 ** 0.5) * i
    return radius * np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)


def orbit_rays(count=16, res=512, radius=3.0, fov_y_deg=30.0, tmax=10.0):
    """Config 4 rays: res x res pinhole rays from each of `count` cameras."""
    out = []
    for c in fibonacci_cameras(count, radius):
        up = (0.0, 1.0, 0.0) if abs(c[1]) < 0.99 * radius else (1.0, 0.0, 0.0)
        r, u, f = _camera_basis(-c, up)
        tan = math.tan(math.radians(fov_y_deg) * 0.5)
        x = (2.0 * (np.arange(res) + 0.5) / res - 1.0) * tan
        xx, yy = np.meshgrid(x, x, indexing="xy")
        d = xx[..., None] * r + yy[..., None] * u + f
        d = (d / np.linalg.norm(d, axis=-1, keepdims=True)).reshape(-1, 3)
        rays = np.empty((res * res, 8), np.float32)
        rays[:, 0:3] = c
        rays[:, 3] = 0.0
        rays[:, 4:7] = d
        rays[:, 7] = tmax
        out.append(rays)
    return np.concatenate(out)


def random_rays(n, seed=5, box=(0.0, 1.0), tmax=3.0):
    """Rays from uniform origins around the unit box aimed at uniform targets inside it."""
    rng = np.random.default_rng(seed)
    lo, hi = box
    o = rng.uniform(lo - 0.5, hi + 0.5, (n, 3))
    tgt = rng.uniform(lo, hi, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


def box_rays(bmin, bmax, count, rng, tmax_scale=3.0):
    """Config 3 rays for one mesh: origins on a sphere around its box, aimed at
    uniform points inside it."""
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    c = 0.5 * (bmin + bmax)
    r = 0.75 * np.linalg.norm(bmax - bmin) + 1e-3
    u = rng.standard_normal((count, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = c + r * u
    tgt = bmin + rng.random((count, 3)) * (bmax - bmin)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((count, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = tmax_scale * r
    return rays


def batch_mesh(count, rng, sphere):
    """One config 3 mesh of `count` triangles in the unit box: a displaced
    icosphere (largest level that fits, topped up with soup) or a soup."""
    if sphere and count >= 20:
        level = 0
        while 20 * 4 ** (level + 1) <= count and level < 7:
            level += 1
        v, f = icosphere(level)
        k = rng.integers(1, 6, 3)
        ph = rng.random(3) * 6.283
        r = 1.0 + 0.05 * (np.sin(k[0] * v[:, 0] + ph[0]) * np.sin(k[1] * v[:, 1] + ph[1]) *
                          np.sin(k[2] * v[:, 2] + ph[2]))
        v = 0.5 + 0.45 * v * r[:, None]
        extra = count - f.shape[0]
        if extra > 0:
            sv, sf = soup_triangles(rng.random((extra, 3), dtype=np.float32), rng)
            f = np.concatenate([f, sf + v.shape[0]])
            v = np.concatenate([v, sv])
        return v.astype(np.float32), f.astype(np.int32)
    return soup_triangles(rng.random((count, 3), dtype=np.float32), rng)


def sliver_soup(count, rng, max_aspect=200.0):
    """Long thin triangles in the unit box (config 5): each spans a random
    direction with length U[0.05, 0.4] and width length / U[1, max_aspect]."""
    c = rng.random((count, 3)).astype(np.float64) * 0.6 + 0.2
    d = rng.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.normal(size=(count, 3))
    w -= (w * d).sum(1, keepdims=True) * d
    w /= np.maximum(np.linalg.norm(w, axis=1, keepdims=True), 1e-12)
    length = rng.uniform(0.05, 0.4, (count, 1))
    width = length / rng.uniform(1.0, max_aspect, (count, 1))
    v0 = c - 0.5 * length * d
    v1 = c + 0.5 * length * d
    v2 = c + width * w
    v = np.stack([v0, v1, v2], 1).reshape(-1, 3).clip(0.0, 1.0).astype(np.float32)
    return v, np.arange(3 * count, dtype=np.int32).reshape(-1, 3)


def batch_meshes(n_meshes, seed=3, lo=64, hi=262144):
    """Config 3 geometry: counts round(exp(U[ln lo, ln hi])), half icospheres."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), n_meshes))).astype(np.int64)
    return [batch_mesh(int(c), rng, bool(rng.random() < 0.5)) for c in counts]


def batch_rays(meshes, per_1k=256, seed=4):
    """Rays for a list of meshes: per_1k per 1000 triangles (at least 1) per mesh;
    returns rays (R, 8) and the mesh of each ray (R,)."""
    rng = np.random.default_rng(seed)
    rays, owner = [], []
    for m, (v, f) in enumerate(meshes):
        cnt = max(1, int(round(per_1k * f.shape[0] / 1000.0)))
        used = v[f.reshape(-1)]
        rays.append(box_rays(used.min(0), used.max(0), cnt, rng))
        owner.append(np.full(cnt, m, np.int32))
    return np.concatenate(rays), np.concatenate(owner)
