"""The steps either side of a trace: ray generation, occlusion reconstruction
and keyframe interpolation, over libhx's device kernels.

Host-side mirrors (src/hippie/):
  Camera           rt/scene/Camera.cpp:11-215 (position, direction angles,
                   projection * view, reset from Camera.* options)
  pixel_table      rt/ray/PixelTable.cpp:10-82 (8x8 Morton-ordered tiles)
  RayGen           rt/ray/RayGen.cpp:35-200 (primary / shadow / AO / path)
  Interpolator     rt/scene/Interpolator.cpp:21-95 (frame pair and t)
The 4x4 matrix helpers follow util/Math.h / Math.cpp (perspective, rotate,
translate, cofactor inverse) in float32 with the reference's operation order;
the host's tan/cos/sin/asin/atan2 are evaluated in double and rounded to
float.  Device kernels: hx_init_seeds, hx_primary_rays, hx_shadow_rays,
hx_ao_rays, hx_path_rays, hx_reconstruct_occlusion, hx_interpolate_vertices.
"""
import ctypes
import math

import numpy as np
import torch

from ._lib import call
from .core import _ptr

F = np.float32
PI_F = F(math.pi)
MATERIAL_DTYPE = np.dtype([("tex_index", "<i4"), ("shininess", "<f4"), ("diffuse", "<f4", (3,)),
                           ("specular", "<f4", (3,))])


def radians(a):
    """a * M_PIf / 180.0f (util/Math.h:16)."""
    return F(F(F(a) * PI_F) / F(180.0))


def degrees(a):
    """a * 180.0f / M_PIf (util/Math.h:17)."""
    return F(F(F(a) * F(180.0)) / PI_F)


def _fn(f, x):
    return F(f(float(x)))


def mat_identity():
    return np.eye(4, dtype=np.float32)


def mat_mul(a, b):
    """MatBase::operator* (util/Math.h:390-401): each element a sum from 0."""
    r = np.zeros((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            rr = F(0.0)
            for k in range(4):
                rr = F(rr + F(a[i, k] * b[k, j]))
            r[i, j] = rr
    return r


def perspective(fovy, aspect, near, far):
    """util/Math.cpp:10-20 (fovy in radians)."""
    t = _fn(math.tan, F(fovy * F(0.5)))
    fy = F(F(1.0) / t) if t != 0 else F(0.0)
    fx = F(fy / F(aspect))
    nf = F(F(near) - F(far))
    d = F(F(1.0) / nf) if nf != 0 else F(0.0)
    r = np.zeros((4, 4), np.float32)
    r[0, 0] = fx
    r[1, 1] = fy
    r[2, 2] = F(F(F(near) + F(far)) * d)
    r[2, 3] = F(F(F(F(2.0) * F(near)) * F(far)) * d)
    r[3, 2] = F(-1.0)
    return r


def rotate(angle, axis):
    """util/Math.cpp:22-30."""
    x, y, z = (F(a) for a in axis)
    c = _fn(math.cos, angle)
    s = _fn(math.sin, angle)
    omc = F(F(1.0) - c)
    R = mat_identity()
    R[0, 0] = F(c + F(F(x * x) * omc))
    R[0, 1] = F(F(F(x * y) * omc) - F(z * s))
    R[0, 2] = F(F(F(x * z) * omc) + F(y * s))
    R[1, 0] = F(F(F(x * y) * omc) + F(z * s))
    R[1, 1] = F(c + F(F(y * y) * omc))
    R[1, 2] = F(F(F(y * z) * omc) - F(x * s))
    R[2, 0] = F(F(F(z * x) * omc) - F(y * s))
    R[2, 1] = F(F(F(z * y) * omc) + F(x * s))
    R[2, 2] = F(c + F(F(z * z) * omc))
    return R


def translate(xyz):
    """util/Math.cpp:32-38."""
    R = mat_identity()
    R[0, 3], R[1, 3], R[2, 3] = (F(a) for a in xyz)
    return R


def _det(m):
    """detImpl (util/Math.h:318-345): closed forms for 1..3, cofactor expansion
    along column 0 above."""
    n = m.shape[0]
    if n == 1:
        return F(m[0, 0])
    if n == 2:
        return F(F(m[0, 0] * m[1, 1]) - F(m[0, 1] * m[1, 0]))
    if n == 3:
        t1 = F(F(m[0, 0] * m[1, 1]) * m[2, 2])
        t2 = F(F(m[0, 0] * m[1, 2]) * m[2, 1])
        t3 = F(F(m[1, 0] * m[2, 1]) * m[0, 2])
        t4 = F(F(m[1, 0] * m[2, 2]) * m[0, 1])
        t5 = F(F(m[2, 0] * m[0, 1]) * m[1, 2])
        t6 = F(F(m[2, 0] * m[0, 2]) * m[1, 1])
        return F(F(F(F(F(t1 - t2) + t3) - t4) + t5) - t6)
    r, s = F(0.0), F(1.0)
    for i in range(n):
        rows = [j if j < i else j + 1 for j in range(n - 1)]
        sub = m[np.ix_(rows, list(range(1, n)))]
        r = F(r + F(F(_det(sub) * m[i, 0]) * s))
        s = F(-s)
    return r


def inverse4(m):
    """inverse (util/Math.h:350-372): cofactors, d = sum of cofactor * entry
    over every (i, j) (= 4 det), then r * rcp(d) * 4."""
    m = np.asarray(m, np.float32)
    r = np.zeros((4, 4), np.float32)
    d = F(0.0)
    si = F(1.0)
    for i in range(4):
        sj = si
        for j in range(4):
            rows = [k if k < j else k + 1 for k in range(3)]
            cols = [l if l < i else l + 1 for l in range(3)]
            dd = F(_det(m[np.ix_(rows, cols)]) * sj)
            r[i, j] = dd
            d = F(d + F(dd * m[j, i]))
            sj = F(-sj)
        si = F(-si)
    rd = F(F(1.0) / d) if d != 0 else F(0.0)
    return ((r * rd).astype(np.float32) * F(4.0)).astype(np.float32)


def _normalize(v):
    v = np.asarray(v, np.float32)
    s = F(0.0)
    for c in v:
        s = F(s + F(c * c))
    ln = F(math.sqrt(float(s)))
    rc = F(F(1.0) / ln) if ln != 0 else F(0.0)
    return (rc * v).astype(np.float32)


class Camera:
    """rt/scene/Camera.cpp: a pinhole camera from a position, a direction
    (stored as view and elevation angles in degrees), a wheel angle, near and
    far planes, a vertical field of view and an image size."""

    def __init__(self, position=(0.0, 0.0, 0.0), direction=(0.0, 0.0, -1.0), near=0.001, far=3.0, fov=45.0,
                 wheel_angle=0.0, size=(1024, 768)):
        self.position = np.asarray(position, np.float32)
        self.near, self.far = F(near), F(far)
        self.field_of_view = radians(fov)
        self.wheel_angle = F(wheel_angle)
        self.size = (int(size[0]), int(size[1]))
        self.set_direction(direction)

    @classmethod
    def from_env(cls, env):
        """Camera::reset (Camera.cpp:187-209) plus the Resolution.* size."""
        return cls(env.get("Camera.position"), env.get("Camera.direction"), env.get("Camera.nearPlane"),
                   env.get("Camera.farPlane"), env.get("Camera.fieldOfView"), env.get("Camera.wheelAngle"),
                   (env.get("Resolution.width"), env.get("Resolution.height")))

    def set_direction(self, d):
        """Camera::setDirection (Camera.cpp:58-62)."""
        d = _normalize(d)
        self.elev_angle = degrees(_fn(math.asin, d[1]))
        self.view_angle = degrees(F(math.atan2(float(d[0]), float(-d[2]))))

    def direction(self):
        """Camera::getDirection (Camera.cpp:52-56)."""
        r = mat_mul(rotate(radians(F(-self.elev_angle)), (1.0, 0.0, 0.0)), rotate(radians(self.view_angle),
                                                                               (0.0, 1.0, 0.0)))
        # rotMatrix[c][2]: column c (Mat4f is column-major), row 2.
        return _normalize(-np.array([r[2, 0], r[2, 1], r[2, 2]], np.float32))

    def projection_matrix(self):
        return perspective(self.field_of_view, F(F(self.size[0]) / F(self.size[1])), self.near, self.far)

    def view_matrix(self):
        """Camera::getViewMatrix (Camera.cpp:99-105)."""
        m = rotate(radians(self.wheel_angle), (0.0, 0.0, 1.0))
        m = mat_mul(m, rotate(radians(F(-self.elev_angle)), (1.0, 0.0, 0.0)))
        m = mat_mul(m, rotate(radians(self.view_angle), (0.0, 1.0, 0.0)))
        return mat_mul(m, translate(-self.position))

    def projection_view_matrix(self):
        return mat_mul(self.projection_matrix(), self.view_matrix())

    def screen_to_world(self):
        """inverse(projection * view), as RayGen::primary passes it (RayGen.cpp:47)."""
        return inverse4(self.projection_view_matrix())


def pixel_table(width, height):
    """PixelTable::recalculate (PixelTable.cpp:10-82): the ray -> pixel order,
    8x8 tiles in Morton order with a Morton swizzle inside each tile, then the
    bottom stripe and the right stripe.  Returns (index_to_pixel,
    pixel_to_index) int32 arrays."""
    w, h = int(width), int(height)
    itp = np.empty(w * h, np.int32)
    pti = np.empty(w * h, np.int32)
    bh, bw = h & ~7, w & ~7
    maxdim = max(bw, bh)
    for s in (1, 2, 4, 8, 16):
        maxdim |= maxdim >> s
    maxdim = (maxdim + 1) >> 1
    w8, h8 = bw >> 3, bh >> 3
    inner = np.arange(64)
    ix = ((inner & 1) >> 0) | ((inner & 4) >> 1) | ((inner & 16) >> 2)
    iy = ((inner & 2) >> 1) | ((inner & 8) >> 2) | ((inner & 32) >> 3)
    idx = 0
    for i in range(maxdim * maxdim):
        tx = ty = 0
        val, bit = i, 1
        while val:
            if val & 1:
                tx |= bit
            if val & 2:
                ty |= bit
            bit += bit
            val >>= 2
        if tx < w8 and ty < h8:
            pos = (ty * 8 + iy) * w + (tx * 8 + ix)
            itp[idx:idx + 64] = pos
            idx += 64
    for px in range(bw):
        for py in range(bh, h):
            itp[idx] = px + py * w
            idx += 1
    for py in range(h):
        for px in range(bw, w):
            itp[idx] = px + py * w
            idx += 1
    assert idx == w * h
    pti[itp] = np.arange(w * h, dtype=np.int32)
    return itp, pti


def _dev(x, dtype, device):
    return torch.as_tensor(x).to(device=device, dtype=dtype).contiguous()


def _check_dev(name, t, dtype, shape=None):
    if not t.is_cuda or not t.is_contiguous() or t.dtype != dtype:
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor")
    if shape is not None and tuple(t.shape[1:]) != tuple(shape):
        raise ValueError(f"{name} must have rows of shape {shape}")


def init_seeds(ctx, n, frame_index=0, device="cuda"):
    """initSeeds: seeds[i] = tea<16>(i, frame_index), as int32 bit patterns."""
    out = torch.empty(int(n), dtype=torch.int32, device=device)
    call("hx_init_seeds", ctx.handle, int(n), int(frame_index), _ptr(out))
    return out


class RayGen:
    """RayGen (rt/ray/RayGen.cpp) on a context: every method writes its rays
    on the device in the reference's layout ((n, 8) float32)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._tables = {}

    def pixel_table(self, width, height, device="cuda"):
        key = (width, height, str(device))
        if key not in self._tables:
            itp, pti = pixel_table(width, height)
            self._tables[key] = (_dev(itp, torch.int32, device), _dev(pti, torch.int32, device))
        return self._tables[key]

    def primary(self, camera, sample_index=0, device="cuda", out=None):
        """RayGen::primary: one ray per pixel in pixel-table order; returns
        (rays, index_to_pixel)."""
        w, h = camera.size
        itp, _ = self.pixel_table(w, h, device)
        rays = out if out is not None else torch.empty((w * h, 8), dtype=torch.float32, device=device)
        _check_dev("rays", rays, torch.float32, (8,))
        o = (ctypes.c_float * 3)(*[float(x) for x in camera.position])
        m = (ctypes.c_float * 16)(*[float(x) for x in camera.screen_to_world().reshape(16)])
        call("hx_primary_rays", self.ctx.handle, int(sample_index), o, m, w, h, float(camera.far), _ptr(itp),
             _ptr(rays))
        return rays, itp

    def shadow(self, in_rays, in_results, seeds, n_samples, light, light_radius, begin=0, size=None, out=None):
        """RayGen::shadow: n_samples rays per input ray of [begin, begin+size)."""
        size = in_rays.shape[0] - begin if size is None else size
        out = out if out is not None else torch.empty((size * n_samples, 8), dtype=torch.float32,
                                                      device=in_rays.device)
        L = (ctypes.c_float * 3)(*[float(x) for x in light])
        call("hx_shadow_rays", self.ctx.handle, int(begin), int(size), int(n_samples), _ptr(seeds),
             float(light_radius), L, _ptr(in_rays), _ptr(in_results), _ptr(out))
        return out

    def ao(self, in_rays, in_results, seeds, n_samples, max_dist, scene, begin=0, size=None, out=None):
        """RayGen::ao: n_samples cosine-weighted rays about each hit's normal."""
        size = in_rays.shape[0] - begin if size is None else size
        out = out if out is not None else torch.empty((size * n_samples, 8), dtype=torch.float32,
                                                      device=in_rays.device)
        call("hx_ao_rays", self.ctx.handle, int(begin), int(size), int(n_samples), _ptr(seeds), float(max_dist),
             _ptr(in_rays), _ptr(in_results), _ptr(scene.triangles), _ptr(scene.vertices), _ptr(out))
        return out

    def path(self, in_rays, in_results, seeds, in_pixel, normals, triangles, mat_indices, materials, decreases,
             russian_roulette=False):
        """RayGen::path: the continuation of every live path, in input order;
        advances seeds and multiplies decreases (float3 per pixel) in place.
        Returns (rays, index_to_pixel) of the surviving paths."""
        n = int(in_rays.shape[0])
        mats = materials
        if isinstance(materials, np.ndarray):
            mats = torch.as_tensor(np.ascontiguousarray(materials, MATERIAL_DTYPE).view(np.uint8),
                                   device=in_rays.device)
        out = torch.empty((max(n, 1), 8), dtype=torch.float32, device=in_rays.device)
        out_pixel = torch.empty(max(n, 1), dtype=torch.int32, device=in_rays.device)
        n_out = torch.zeros(1, dtype=torch.int32, device=in_rays.device)
        call("hx_path_rays", self.ctx.handle, int(bool(russian_roulette)), n, _ptr(seeds), _ptr(in_pixel),
             _ptr(mat_indices), _ptr(triangles), _ptr(normals), _ptr(mats), _ptr(in_rays), _ptr(in_results),
             _ptr(out), _ptr(out_pixel), _ptr(decreases), _ptr(n_out))
        k = int(n_out.item())
        return out[:k], out_pixel[:k]


def reconstruct_occlusion(ctx, results, index_to_pixel, n_samples, in_pixels, out_pixels, begin=0, size=None,
                          replace=True, index_to_slot=None):
    """Renderer's reconstructAO / reconstructShadow: out pixel = in colour x the
    fraction of the pixel's n_samples rays that miss (replace, or add)."""
    size = index_to_pixel.shape[0] - begin if size is None else size
    call("hx_reconstruct_occlusion", ctx.handle, int(begin), int(size), int(n_samples), int(bool(replace)),
         _ptr(results), _ptr(index_to_pixel), _ptr(index_to_slot), _ptr(in_pixels), _ptr(out_pixels))
    return out_pixels


def interpolate_vertices(ctx, frame0, frame1, t, out, n_static=0, normalize=False):
    """interpolateVertices / interpolateNormals: out[n_static + i] = mix(frame0[i], frame1[i], t)."""
    n = int(frame0.shape[0])
    if frame1.shape != frame0.shape or out.shape[0] < n_static + n:
        raise ValueError("frames must match and out must hold n_static + n entries")
    call("hx_interpolate_vertices", ctx.handle, int(n_static), n, float(t), _ptr(frame0), _ptr(frame1), _ptr(out),
         int(bool(normalize)))
    return out


class Interpolator:
    """Interpolator::updateAdaptive (Interpolator.cpp:21-95): advance a dynamic
    scene's time, pick the bounding keyframes and t, and interpolate the
    dynamic vertices (after the static ones) and normals on the device; the
    result is the input of HipBVH.update / BVHBuilder.rebuild."""

    EPSILON = 1.0e-5

    def __init__(self, ctx, frames, frame_rate=30.0, looped=True):
        """frames: list of (vertices (Vd,3), normals (Vd,3) or None) device tensors."""
        self.ctx = ctx
        self.frames = frames
        self.frame_rate = F(frame_rate)
        self.looped = looped
        self.time = F(0.0)

    def frame_pair(self, time=None):
        """Advance the scene time (one frame, or by `time` seconds rounded up
        to whole frames) and return (frame0, frame1, t)."""
        if time is None or time < 0:
            st = F(self.time + F(F(1.0) / self.frame_rate))
        else:
            st = F(self.time + F(F(math.ceil(float(F(F(time) * self.frame_rate)))) / self.frame_rate))
        if self.looped:
            while st < 0:
                st = F(st + F(1.0))
            while st >= 1:
                st = F(st - F(1.0))
        else:
            st = F(min(max(st, F(0.0)), F(1.0) - F(self.EPSILON)))
        self.time = st
        nf = len(self.frames)
        fi = F(st * F(nf - 1))
        f0 = int(fi)
        f1 = (f0 + 1) % nf
        return f0, f1, F(fi - F(f0))

    def update(self, vertices, normals=None, n_static=0, time=None):
        f0, f1, t = self.frame_pair(time)
        v0, n0 = self.frames[f0]
        v1, n1 = self.frames[f1]
        interpolate_vertices(self.ctx, v0, v1, t, vertices, n_static)
        if normals is not None and n0 is not None:
            interpolate_vertices(self.ctx, n0, n1, t, normals, n_static, normalize=True)
        return t
