"""hippie_amd -- MI355X (gfx950) BVH construction and ray traversal behind a C ABI.

The compute path is libhx.so (HIP kernels + extern "C" boundary, include/hx.h);
this package is the host-side mirror of hippie's builder/tracer interface.
"""
from . import _lib  # noqa: F401
from . import formats, raygen  # noqa: F401
from .core import (ATRBuilder, BuildParams, BVHBuilder, Context, HipBatch, HipBVH, HipTracer, HLBVHBuilder,  # noqa: F401
                   InsertionBuilder, LBVHBuilder, MeshBatch, PLOCBuilder, RayBuffer, Scene, TRBuilder, morton_codes,
                   normalize_vertices, ray_keys, scene_box, sort_pairs, trace_view)

__all__ = [
    "ATRBuilder", "BuildParams", "BVHBuilder", "Context", "HipBatch", "HipBVH", "HipTracer", "HLBVHBuilder", "InsertionBuilder",
    "LBVHBuilder",
    "MeshBatch", "PLOCBuilder", "RayBuffer", "Scene", "TRBuilder", "morton_codes", "normalize_vertices", "ray_keys", "scene_box", "sort_pairs",
    "trace_view",
]
