"""Host-side mirror of hippie's builder / BVH / tracer interface over libhx.

Class and method names follow the reference (src/hippie/):
  Scene                   rt/scene/Scene.h (vertices, triangles, scene box)
  BVHBuilder, LBVHBuilder rt/bvh/BVHBuilder.h:16-68, rt/bvh/LBVHBuilder.cpp
  HipBVH                  rt/bvh/HipBVH.h:19-106 (nodes, woop triangles,
                          triangle indices, getCost, update)
  HipTracer               rt/tracer/HipTracer.h:16-44 (setBVH, trace,
                          traceStats, traceSort)
  RayBuffer               rt/ray/RayBuffer.h:16-81 (rays, results, stats, indices)
Invalid parameters raise ValueError instead of the reference's log-and-ignore.
PyTorch only provides device memory and the stream; all work runs in libhx.
"""
import contextlib
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HxBuildParams, HxBvhView, HxBvhDebugView, call, lib

METHODS = {
    "lbvh": _lib.HX_METHOD_LBVH,
    "hlbvh": _lib.HX_METHOD_HLBVH,
    "ploc": _lib.HX_METHOD_PLOC,
    "atr": _lib.HX_METHOD_ATRBVH,
    "insertion": _lib.HX_METHOD_INSERTION,
    "tr": _lib.HX_METHOD_TR,
}
LAYOUTS = {"bin": _lib.HX_LAYOUT_BIN, "quad": _lib.HX_LAYOUT_QUAD, "oct": _lib.HX_LAYOUT_OCT}
RAY_KEY_METHODS = {"twopoint": _lib.HX_RAYKEY_TWOPOINT, "aila": _lib.HX_RAYKEY_AILA}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """One libhx context per device (HipModule::staticInit)."""

    _default = {}

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("hippie_amd needs a ROCm GPU: torch.cuda.is_available() is False")
        self.device = int(device)
        h = ctypes.c_void_p()
        call("hx_context_create", self.device, ctypes.byref(h))
        self.handle = h
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream().cuda_stream
        call("hx_context_set_stream", self.handle, ctypes.c_void_p(stream))

    @classmethod
    def default(cls, device=0):
        if device not in cls._default:
            cls._default[device] = Context(device)
        return cls._default[device]

    def set_stream(self, stream):
        call("hx_context_set_stream", self.handle, ctypes.c_void_p(stream.cuda_stream))

    def synchronize(self):
        call("hx_context_synchronize", self.handle)

    def set_profiling(self, on=True):
        call("hx_context_set_profiling", self.handle, int(bool(on)))

    def profile(self):
        """{kernel name: (launches, total ms)} since the last reset."""
        cnt = ctypes.c_int32()
        call("hx_context_profile_count", self.handle, ctypes.byref(cnt))
        out = {}
        buf = ctypes.create_string_buffer(64)
        for i in range(cnt.value):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            call("hx_context_profile_entry", self.handle, i, buf, 64, ctypes.byref(n), ctypes.byref(ms))
            out[buf.value.decode()] = (n.value, ms.value)
        return out

    def profile_reset(self):
        call("hx_context_profile_reset", self.handle)

    def set_tuning(self, name, value):
        call("hx_context_set_tuning", self.handle, name.encode(), int(value))

    def get_tuning(self, name):
        v = ctypes.c_int32()
        call("hx_context_get_tuning", self.handle, name.encode(), ctypes.byref(v))
        return v.value

    def memcpy(self, dst_ptr, src_ptr, nbytes):
        call("hx_memcpy", self.handle, ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr), int(nbytes))

    def overflow_count(self):
        v = ctypes.c_int64()
        call("hx_trace_overflow_count", self.handle, ctypes.byref(v))
        return v.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().hx_context_destroy(h)
            except Exception:
                pass
            self.handle = None


class Scene:
    """Triangle geometry resident on the device (float32 Vec3f, int32 Vec3i)."""

    def __init__(self, vertices, triangles, device=0):
        dev = torch.device("cuda", device)
        v = torch.as_tensor(vertices)
        t = torch.as_tensor(triangles)
        if v.ndim != 2 or v.shape[1] != 3 or t.ndim != 2 or t.shape[1] != 3:
            raise ValueError("vertices must be (V,3) and triangles (F,3)")
        self.vertices = v.to(device=dev, dtype=torch.float32).contiguous()
        self.triangles = t.to(device=dev, dtype=torch.int32).contiguous()

    @property
    def n_triangles(self):
        return int(self.triangles.shape[0])

    @property
    def n_vertices(self):
        return int(self.vertices.shape[0])


def _same_geometry(old, new):
    """A BVH's buffers are sized for the geometry it was built from."""
    if new.n_triangles != old.n_triangles or tuple(new.vertices.shape) != tuple(old.vertices.shape):
        raise ValueError(f"geometry of {new.n_triangles} triangles / {new.n_vertices} vertices does not match the "
                         f"BVH's {old.n_triangles} triangles / {old.n_vertices} vertices")


class BuildParams:
    """hx_build_params with the reference's Bvh.* defaults (AppEnvironment.cpp:52-75)."""

    ENV_KEYS = {
        "Bvh.method": "method", "Bvh.maxLeafSize": "max_leaf_size",
        "Bvh.adaptiveLeafSize": "adaptive_leaf_size", "Bvh.ct": "ct", "Bvh.ci": "ci",
        "Bvh.presplitting": "presplitting", "Bvh.presplitterBeta": "presplitter_beta",
        "Bvh.plocRadius": "ploc_radius", "Bvh.plocMaxRadius": "ploc_max_radius",
        "Bvh.plocAdaptive": "ploc_adaptive", "Bvh.atrIterations": "atr_iterations",
        "Bvh.trIterations": "tr_iterations",
        "Bvh.hlbvhMortonCodeSAHBits": "hlbvh_sah_bits", "Bvh.insertionMod": "insertion_mod",
    }
    BITS_KEYS = {"lbvh": "Bvh.lbvhMortonCodeBits", "ploc": "Bvh.plocMortonCodeBits",
                 "atr": "Bvh.atrMortonCodeBits", "tr": "Bvh.trMortonCodeBits", "hlbvh": "Bvh.hlbvhMortonCodeBits",
                 "insertion": "Bvh.insertionMortonCodeBits"}

    def __init__(self, **kw):
        self.c = HxBuildParams()
        call("hx_build_params_default", ctypes.byref(self.c))
        for k, v in kw.items():
            setattr(self, k, v)

    def __setattr__(self, name, value):
        if name == "c":
            object.__setattr__(self, name, value)
            return
        if name == "method" and isinstance(value, str):
            value = METHODS[value]
        if name == "layout" and isinstance(value, str):
            value = LAYOUTS[value]
        if name == "scene_box":
            arr = (ctypes.c_float * 6)(*[float(x) for x in value])
            self.c.scene_box = arr
            self.c.use_scene_box = 1
            return
        if not hasattr(self.c, name):
            raise AttributeError(f"unknown build parameter {name}")
        setattr(self.c, name, value)

    def __getattr__(self, name):
        return getattr(object.__getattribute__(self, "c"), name)

    @classmethod
    def from_env(cls, env):
        """Map the Bvh.* options of a reference .env file onto parameters.
        ``env`` is a formats.Environment (typed values and the reference's
        defaults) or a dict of Bvh.* key/value pairs."""
        from .formats import Environment
        if isinstance(env, Environment):
            env = env.as_dict("Bvh.")
        layout = str(env.get("Bvh.layout", "bin")).lower()
        p = cls()
        if layout != "bin":
            if layout not in LAYOUTS:
                raise ValueError(f"unknown Bvh.layout '{layout}'")
            p.layout = LAYOUTS[layout]
        method = str(env.get("Bvh.method", "lbvh"))
        p.method = method
        if method in cls.BITS_KEYS and cls.BITS_KEYS[method] in env:
            p.morton_bits = int(env[cls.BITS_KEYS[method]])
        if method == "insertion" and str(env.get("Bvh.insertionSbvh", "0")).strip().lower() in ("1", "true"):
            raise NotImplementedError("Bvh.insertionSbvh: the SBVH start of the insertion builder is not available")
        for key, field in cls.ENV_KEYS.items():
            if key in env and key != "Bvh.method":
                cur = getattr(p.c, field)
                setattr(p, field, type(cur)(env[key]))
        return p


class HipBVH:
    """A built Bin BVH owned by libhx (HipBVH)."""

    def __init__(self, ctx, handle, scene, params):
        self.ctx = ctx
        self.handle = handle
        self.scene = scene
        self.params = params

    def view(self):
        v = HxBvhView()
        call("hx_bvh_get_view", self.handle, ctypes.byref(v))
        return v

    @property
    def n_interior(self):
        return self.view().n_interior

    @property
    def n_leaves(self):
        return self.view().n_leaves

    @property
    def n_nodes(self):
        v = self.view()
        return v.n_interior + v.n_leaves

    def _copy(self, ptr, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes:
            self.ctx.memcpy(out.ctypes.data, ptr, out.nbytes)
        return out

    def nodes_host(self):
        v = self.view()
        # 8N words per node: 16 (Bin), 32 (Quad), 64 (Oct)
        return self._copy(v.nodes, (v.n_interior + v.n_leaves, 8 * v.layout), np.int32)

    @property
    def layout(self):
        return self.view().layout

    def woop_host(self):
        v = self.view()
        return self._copy(v.woop, (v.n_references, 12), np.float32)

    def tri_indices_host(self):
        v = self.view()
        return self._copy(v.tri_indices, (v.n_references,), np.int32)

    def scene_box(self):
        return np.array(self.view().scene_box[:], np.float32)

    def debug_host(self):
        """Intermediate arrays of the build (sorted keys, Karras arrays, boxes)."""
        d = HxBvhDebugView()
        call("hx_bvh_debug", self.handle, ctypes.byref(d))
        n = self.view().n_references
        N = max(2 * n - 1, 1)
        kd = np.uint64 if d.keys_are_64bit else np.uint32
        out = dict(
            sorted_keys=self._copy(d.sorted_keys, (n,), kd).astype(np.uint64),
            sorted_tri=self._copy(d.sorted_tri, (n,), np.int32),
            left=self._copy(d.left, (N,), np.int32),
            right=self._copy(d.right, (N,), np.int32),
            parent=self._copy(d.parent, (N,), np.int32),
            node_min=self._copy(d.node_min, (N, 4), np.float32),
            node_max=self._copy(d.node_max, (N, 4), np.float32),
            state=self._copy(d.state, (max(n - 1, 0),), np.int32),
        )
        return out

    def validate(self):
        """HipBVH::validate on the device: 0 if the BVH is valid for the geometry
        of its last build or refit, else the OR of the failed checks (hx.h)."""
        f = ctypes.c_int32()
        call("hx_bvh_validate", self.handle, ctypes.byref(f))
        return f.value

    def leaf_stats(self):
        """(counts of leaves of size 0..32, sum of leaf sizes) -- the inputs of
        HipBVH::getLeafSizeHistogram / getAvgLeafSize."""
        h = np.zeros(33, np.int32)
        s = ctypes.c_int64()
        call("hx_bvh_leaf_stats", self.handle, h.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s))
        return h, s.value

    def avg_leaf_size(self):
        """HipBVH::getAvgLeafSize (HipBVH.cpp:380-404)."""
        return self.leaf_stats()[1] / float(self.n_leaves)

    def leaf_size_histogram(self):
        """HipBVH::getLeafSizeHistogram (HipBVH.cpp:406-442): fraction of leaves of each size 0..32."""
        return self.leaf_stats()[0] / float(self.n_leaves)

    def get_cost(self):
        c = ctypes.c_double()
        call("hx_bvh_cost", self.handle, ctypes.byref(c))
        return c.value

    def info(self, key):
        """Build statistic: "ploc_rounds", "atr_accepted" or "method"."""
        v = ctypes.c_int64()
        call("hx_bvh_info", self.handle, key.encode(), ctypes.byref(v))
        return v.value

    def update(self, scene=None):
        """HipBVH::update -- refit + re-woopify with unchanged topology.  The
        scene must have the triangle and vertex counts of the build."""
        s = scene or self.scene
        _same_geometry(self.scene, s)
        call("hx_refit", self.ctx.handle, self.handle, _ptr(s.vertices), s.n_vertices, _ptr(s.triangles),
             s.n_triangles)
        self.scene = s  # the BVH keeps the refit geometry's device pointers

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().hx_bvh_destroy(h)
            except Exception:
                pass
            self.handle = None


class BVHBuilder:
    """BVHBuilder::build / rebuild over libhx (rt/bvh/BVHBuilder.h:37-39)."""

    method = "lbvh"

    def __init__(self, ctx=None, params=None, **kw):
        self.ctx = ctx or Context.default()
        self.params = params or BuildParams(method=self.method)
        for k, v in kw.items():
            setattr(self.params, k, v)

    def set_max_leaf_size(self, n):
        if not (0 < n <= 1024):
            raise ValueError("maximum leaf size must be in range (0,1024]")
        self.params.max_leaf_size = n

    def set_morton_code_bits(self, bits):
        if not (6 <= bits <= 60):
            raise ValueError("Morton code bits must be in range [6,60]")
        self.params.morton_bits = bits

    def build(self, scene):
        h = ctypes.c_void_p()
        call("hx_build", self.ctx.handle, ctypes.byref(self.params.c), _ptr(scene.vertices), scene.n_vertices,
             _ptr(scene.triangles), scene.n_triangles, ctypes.byref(h))
        return HipBVH(self.ctx, h, scene, self.params)

    def rebuild(self, bvh, scene=None):
        """BVHBuilder::rebuild: the reference rebuilds from the BVH's own scene;
        a new scene must have the same triangle and vertex counts."""
        s = scene or bvh.scene
        _same_geometry(bvh.scene, s)
        call("hx_rebuild", self.ctx.handle, bvh.handle, _ptr(s.vertices), s.n_vertices, _ptr(s.triangles),
             s.n_triangles)
        bvh.scene = s


class LBVHBuilder(BVHBuilder):
    method = "lbvh"


class PLOCBuilder(BVHBuilder):
    """PLOCBuilder (rt/bvh/PLOCBuilder.h, PLOCBuilder.cpp:210-273).  The
    reference warns and keeps the old value on an out-of-range radius; here the
    setters raise."""

    method = "ploc"

    def set_radius(self, r):
        if not (1 <= r <= self.params.ploc_max_radius):
            raise ValueError("radius must be in range [1,%d]" % self.params.ploc_max_radius)
        self.params.ploc_radius = r

    def set_max_radius(self, r):
        if r < 1:
            raise ValueError("max. radius must be positive")
        self.params.ploc_max_radius = r

    def set_adaptive(self, on):
        self.params.ploc_adaptive = int(bool(on))


class HLBVHBuilder(BVHBuilder):
    """HLBVHBuilder (rt/bvh/HLBVHBuilder.h, HLBVHBuilder.cpp:412-426): LBVH below
    clusters of the top `hlbvh_sah_bits` Morton bits, binned SAH above."""

    method = "hlbvh"

    def set_morton_code_sah_bits(self, bits):
        if not (6 <= bits <= 15) or bits > self.params.morton_bits:
            raise ValueError("Morton code SAH bits must be in range [6,15] and not above the Morton bits")
        self.params.hlbvh_sah_bits = bits


class ATRBuilder(BVHBuilder):
    """ATRBuilder (rt/bvh/ATRBuilder.h, ATRBuilder.cpp:238-290): LBVH plus
    agglomerative treelet restructuring.  Out-of-range values raise instead of
    being ignored with a warning."""

    method = "atr"

    def set_treelet_size(self, t):
        if not (4 <= t <= 30):
            raise ValueError("treelet size must be in range of [4,30]")
        self.params.atr_treelet_size = t

    def set_iterations(self, k):
        if not (0 <= k <= 25):
            raise ValueError("iterations must be in range of [0,25]")
        self.params.atr_iterations = k


class TRBuilder(BVHBuilder):
    """TRBuilder (rt/bvh/TRBuilder.h, TRBuilder.cpp:281-332): LBVH plus optimal
    restructuring of treelets of 7 leaves (Karras and Aila 2013).  The treelet
    size is fixed at 7, as TRBuilder::setTreeletSize accepts only 7."""

    method = "tr"

    def set_treelet_size(self, t):
        if t != 7:
            raise ValueError("treelet size must be in range of [7,7]")

    def set_iterations(self, k):
        if not (0 <= k <= 25):
            raise ValueError("iterations must be in range of [0,25]")
        self.params.tr_iterations = k


class InsertionBuilder(BVHBuilder):
    """InsertionBuilder (rt/bvh/InsertionBuilder.h, InsertionBuilder.cpp:442-494):
    LBVH plus parallel insertion-based optimisation.  Only the LBVH start is
    available (the reference's SBVH start runs a CPU SBVH builder first)."""

    method = "insertion"

    def set_mod(self, mod):
        if not (1 <= mod <= 64):
            raise ValueError("mod must be in range [1,64]")
        self.params.insertion_mod = mod

    def set_sbvh(self, sbvh):
        if sbvh:
            raise NotImplementedError("the SBVH start of the insertion builder is not available")


class RayBuffer:
    """Rays (float4 origin/tmin, float4 direction/tmax), results int4, stats int2.

    ``launch``: trace tunables ("trace_fetch", "trace_lds_stack",
    "trace_blocks_per_cu", "trace_postpone", "trace_woop_ahead") that HipTracer applies to the context around every
    trace of this ray set, e.g. from Autotuner.tune_trace: coherent primary
    rays and incoherent secondaries favour different refill thresholds.  The
    results do not depend on them."""

    def __init__(self, rays, device=0, closest_hit=True):
        r = torch.as_tensor(rays)
        if r.ndim != 2 or r.shape[1] != 8:
            raise ValueError("rays must be (R, 8) float32")
        self.rays = r.to(device=torch.device("cuda", device), dtype=torch.float32).contiguous()
        n = self.rays.shape[0]
        self.results = torch.zeros((n, 4), dtype=torch.int32, device=self.rays.device)
        self.stats = None
        self.indices = None
        self.closest_hit = closest_hit
        self.launch = {}

    @classmethod
    def wrap(cls, rays, results, closest_hit=True):
        """A ray set over caller-held device tensors (rays (R,8) float32,
        results (R,4) int32), no copy: e.g. one mesh's slice of a batch."""
        rb = cls.__new__(cls)
        rb.rays, rb.results, rb.closest_hit = rays, results, closest_hit
        rb.stats, rb.indices, rb.launch = None, None, {}
        return rb

    @property
    def size(self):
        return int(self.rays.shape[0])


@contextlib.contextmanager
def _launch_scope(ctx, launch):
    """Trace tunables `launch` on the context for one call, then the old values."""
    if not launch:
        yield
        return
    old = {k: ctx.get_tuning(k) for k in launch}
    for k, v in launch.items():
        ctx.set_tuning(k, int(v))
    try:
        yield
    finally:
        for k, v in old.items():
            ctx.set_tuning(k, v)


class HipTracer:
    """HipTracer::setBVH / trace / traceStats / traceSort (rt/tracer/HipTracer.h:16-44)."""

    def __init__(self, ctx=None, bvh=None):
        self.ctx = ctx or Context.default()
        self.bvh = bvh

    def set_bvh(self, bvh):
        self.bvh = bvh

    def _need_bvh(self):
        if self.bvh is None:
            raise RuntimeError("HipTracer: no BVH set")

    def _launch(self, rays):
        return _launch_scope(self.ctx, rays.launch)

    def trace(self, rays):
        self._need_bvh()
        with self._launch(rays):
            call("hx_trace", self.ctx.handle, self.bvh.handle, _ptr(rays.rays), rays.size, int(not rays.closest_hit),
                 _ptr(rays.results), None)
        return rays.results

    def trace_stats(self, rays):
        self._need_bvh()
        if rays.stats is None:
            rays.stats = torch.zeros((rays.size, 2), dtype=torch.int32, device=rays.rays.device)
        with self._launch(rays):
            call("hx_trace", self.ctx.handle, self.bvh.handle, _ptr(rays.rays), rays.size, int(not rays.closest_hit),
                 _ptr(rays.results), _ptr(rays.stats))
        return rays.results, rays.stats

    def trace_sort(self, rays, method="aila", bits=60, ray_length=0.25):
        self._need_bvh()
        if rays.indices is None:
            rays.indices = torch.zeros(rays.size, dtype=torch.int32, device=rays.rays.device)
        with self._launch(rays):
            call("hx_trace_sorted", self.ctx.handle, self.bvh.handle, _ptr(rays.rays), rays.size,
                 int(not rays.closest_hit), RAY_KEY_METHODS[method], int(bits), float(ray_length),
                 _ptr(rays.results), None, _ptr(rays.indices))
        return rays.results

    def trace_sort_noalloc(self, rays, method="aila", bits=60, ray_length=0.25):
        """trace_sort through hx_trace_sorted_noalloc: no device allocation by
        the library (at most HX_NOALLOC_MAX_RAYS rays; calls overlapping on one
        GPU race on its fixed scratch)."""
        self._need_bvh()
        if rays.indices is None:
            rays.indices = torch.zeros(rays.size, dtype=torch.int32, device=rays.rays.device)
        with self._launch(rays):
            call("hx_trace_sorted_noalloc", self.ctx.handle, self.bvh.handle, _ptr(rays.rays), rays.size,
                 int(not rays.closest_hit), RAY_KEY_METHODS[method], int(bits), float(ray_length),
                 _ptr(rays.results), None, _ptr(rays.indices))
        return rays.results

    def trace_sort_timed(self, rays, method="aila", bits=60, ray_length=0.25):
        """HipTracer::traceSort(rays, sortTime, traceTime): (sort seconds, trace
        seconds), from events around the ray keys + sort and around the trace."""
        self._need_bvh()
        if rays.indices is None:
            rays.indices = torch.zeros(rays.size, dtype=torch.int32, device=rays.rays.device)
        ts, tt = ctypes.c_float(), ctypes.c_float()
        with self._launch(rays):
            call("hx_trace_sorted_timed", self.ctx.handle, self.bvh.handle, _ptr(rays.rays), rays.size,
                 int(not rays.closest_hit), RAY_KEY_METHODS[method], int(bits), float(ray_length),
                 _ptr(rays.results), None, _ptr(rays.indices), ctypes.byref(ts), ctypes.byref(tt))
        return ts.value, tt.value


def trace_view(ctx, nodes, woop, tri_indices, rays, any_hit=False, stats=None, indices=None, results=None):
    """TRACE_FUNC over caller-held Bin BVH tensors (int32 (N,16) nodes, float32
    (R,12) Woop rows, int32 triangle indices); rays (n,8) float32 on the device.
    indices: optional int32 ray order (slot i traces ray indices[i])."""
    for name, t in (("nodes", nodes), ("woop", woop), ("tri_indices", tri_indices), ("rays", rays)):
        if not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous CUDA tensor")
    n = int(rays.shape[0])
    if results is None:
        results = torch.zeros((n, 4), dtype=torch.int32, device=rays.device)
    call("hx_trace_view", ctx.handle, _ptr(nodes), int(nodes.shape[0]), _ptr(woop), _ptr(tri_indices),
         int(tri_indices.shape[0]), _ptr(rays), n, int(bool(any_hit)), _ptr(results), _ptr(stats), _ptr(indices))
    return results


def normalize_vertices(ctx, vertices):
    """SceneLoader's load-time normalisation, in place on a (V,3) float32 CUDA
    tensor: v = (v - mn) / largest extent of the vertex box; returns the box."""
    if vertices.dtype != torch.float32 or not vertices.is_cuda or not vertices.is_contiguous():
        raise ValueError("vertices must be a contiguous float32 CUDA tensor")
    box = torch.zeros(6, dtype=torch.float32, device=vertices.device)
    call("hx_normalize_vertices", ctx.handle, _ptr(vertices), int(vertices.shape[0]), _ptr(box))
    return box


def scene_box(ctx, scene):
    out = torch.zeros(6, dtype=torch.float32, device=scene.vertices.device)
    call("hx_scene_box", ctx.handle, _ptr(scene.vertices), _ptr(scene.triangles), scene.n_triangles, _ptr(out))
    return out


def morton_codes(ctx, scene, box, bits):
    box = torch.as_tensor(box, dtype=torch.float32).to(scene.vertices.device).contiguous()
    keys = torch.zeros(scene.n_triangles, dtype=torch.int64, device=scene.vertices.device)
    call("hx_morton_codes", ctx.handle, _ptr(scene.vertices), _ptr(scene.triangles), scene.n_triangles, _ptr(box),
         int(bits), _ptr(keys))
    return keys


def sort_pairs(ctx, keys, values, end_bit):
    """In-place stable radix sort of (int64-stored uint64 key, int32 value)."""
    call("hx_sort_pairs_u64", ctx.handle, _ptr(keys), _ptr(values), int(keys.shape[0]), int(end_bit))


def ray_keys(ctx, rays, method="aila", bits=60, ray_length=0.25):
    keys = torch.zeros(rays.shape[0], dtype=torch.int64, device=rays.device)
    call("hx_ray_keys", ctx.handle, _ptr(rays), int(rays.shape[0]), RAY_KEY_METHODS[method], int(bits),
         float(ray_length), _ptr(keys))
    return keys


class HxBatchMeshView(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_void_p), ("woop", ctypes.c_void_p), ("tri_indices", ctypes.c_void_p),
                ("n_interior", ctypes.c_int64), ("n_leaves", ctypes.c_int64), ("n_references", ctypes.c_int64),
                ("tri_offset", ctypes.c_int64)]


class MeshBatch:
    """Many meshes concatenated: shared vertices, triangles with global vertex
    indices, mesh m = triangles [offsets[m], offsets[m+1])."""

    def __init__(self, vertices, triangles, offsets, device=0):
        self.scene = Scene(vertices, triangles, device)
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if self.offsets.ndim != 1 or self.offsets[0] != 0 or self.offsets[-1] != self.scene.n_triangles:
            raise ValueError("offsets must run from 0 to the number of triangles")

    @classmethod
    def from_meshes(cls, meshes, device=0):
        """meshes: list of (vertices (V,3), triangles (F,3) local indices)."""
        vs, ts, offs, vbase = [], [], [0], 0
        for v, t in meshes:
            vs.append(np.asarray(v, np.float32))
            ts.append(np.asarray(t, np.int32) + vbase)
            vbase += len(v)
            offs.append(offs[-1] + len(t))
        return cls(np.concatenate(vs), np.concatenate(ts), offs, device)

    @property
    def n_meshes(self):
        return int(self.offsets.shape[0] - 1)


class HipBatch:
    """Batched LBVH (hx_build_batch): one Bin BVH per mesh, built, refit and
    traced together.  ``launch``: trace tunables applied around this batch's
    traces (Autotuner.tune_batch_trace), as RayBuffer.launch for one BVH."""

    def __init__(self, ctx, batch, params=None, **kw):
        self.ctx = ctx or Context.default()
        self.batch = batch
        self.params = params or BuildParams(method="lbvh")
        for k, v in kw.items():
            setattr(self.params, k, v)
        h = ctypes.c_void_p()
        s = batch.scene
        call("hx_build_batch", self.ctx.handle, ctypes.byref(self.params.c), _ptr(s.vertices), s.n_vertices,
             _ptr(s.triangles), batch.offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), batch.n_meshes,
             ctypes.byref(h))
        self.handle = h
        self.launch = {}

    def mesh_view(self, m):
        v = HxBatchMeshView()
        call("hx_batch_mesh", self.handle, int(m), ctypes.byref(v))
        return v

    def mesh_host(self, m):
        """Host copies of mesh m's Bin nodes, Woop rows and mesh-local triangle ids."""
        v = self.mesh_view(m)
        nn = v.n_interior + v.n_leaves
        nodes = np.empty((nn, 16), np.int32)
        woop = np.empty((v.n_references, 12), np.float32)
        tri = np.empty(v.n_references, np.int32)
        for arr, ptr in ((nodes, v.nodes), (woop, v.woop), (tri, v.tri_indices)):
            self.ctx.memcpy(arr.ctypes.data, ptr, arr.nbytes)
        return dict(nodes=nodes, woop=woop, tri_idx=tri, n_interior=v.n_interior, n_leaves=v.n_leaves)

    def rebuild(self):
        s = self.batch.scene
        call("hx_batch_rebuild", self.ctx.handle, self.handle, _ptr(s.vertices), s.n_vertices, _ptr(s.triangles),
             s.n_triangles)

    def info(self, key):
        """Build statistic of the whole batch (hx_bvh_info keys, e.g. "references")."""
        v = ctypes.c_int64()
        call("hx_bvh_info", self.handle, key.encode(), ctypes.byref(v))
        return v.value

    def refit(self, vertices=None):
        """Refit every mesh to new vertex positions of the batch's shape."""
        s = self.batch.scene
        if vertices is not None:
            v = torch.as_tensor(vertices).to(device=s.vertices.device, dtype=torch.float32).contiguous()
            if tuple(v.shape) != tuple(s.vertices.shape):
                raise ValueError(f"vertices of shape {tuple(v.shape)} do not match the batch's "
                                 f"{tuple(s.vertices.shape)}")
            s.vertices = v
        call("hx_batch_refit", self.ctx.handle, self.handle, _ptr(s.vertices), s.n_vertices, _ptr(s.triangles),
             s.n_triangles)

    def trace(self, rays, ray_mesh, any_hit=False, results=None, stats=None):
        """rays: RayBuffer or (R,8) tensor on the device; ray_mesh: int32 (R,) on the device;
        stats: optional int32 (R,2) for (nodes visited, triangles tested) per ray."""
        r = rays.rays if isinstance(rays, RayBuffer) else rays
        n = int(r.shape[0])
        if results is None:
            results = torch.zeros((n, 4), dtype=torch.int32, device=r.device)
        with _launch_scope(self.ctx, self.launch):
            call("hx_batch_trace", self.ctx.handle, self.handle, _ptr(r), _ptr(ray_mesh), n, int(bool(any_hit)),
                 _ptr(results), _ptr(stats))
        return results

    def backward(self, rays, ray_mesh, results, dl_dt=None, dl_du=None, dl_dv=None, out=None):
        s = self.batch.scene
        n = int(rays.shape[0])
        if out is None:
            out = torch.zeros_like(s.vertices)
        call("hx_batch_backward", self.ctx.handle, self.handle, _ptr(s.vertices), _ptr(s.triangles), _ptr(rays),
             _ptr(ray_mesh), _ptr(results), n, _ptr(dl_dt), _ptr(dl_du), _ptr(dl_dv), _ptr(out))
        return out

    def recompute_hits(self, rays, ray_mesh, results):
        s = self.batch.scene
        n = int(rays.shape[0])
        t, u, v = (torch.empty(n, dtype=torch.float32, device=rays.device) for _ in range(3))
        call("hx_batch_recompute_hits", self.ctx.handle, self.handle, _ptr(s.vertices), _ptr(s.triangles),
             _ptr(rays), _ptr(ray_mesh), _ptr(results), n, _ptr(t), _ptr(u), _ptr(v))
        return t, u, v

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                call("hx_batch_destroy", h)
            except Exception:
                pass
            self.handle = None
