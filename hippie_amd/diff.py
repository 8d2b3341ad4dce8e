"""Differentiable closest hit: refit -> trace -> adjoint, as a torch.autograd.Function.

Forward refits the BVH (fixed topology) to the current vertex positions and
traces closest hits (the trace picks each ray's triangle); the returned t, u, v
are then recomputed from that triangle by the FP64 solve the adjoint
differentiates, so forward and backward describe the same function.  Backward
runs libhx's adjoint kernel, which recomputes each hit from its stored triangle
id and scatter-adds dL/dvertex.  The hit triangle is a discrete choice: its
gradient is not propagated (rays that would change triangle under a
perturbation are where the loss is not differentiable).
"""
import torch

from ._lib import call
from .core import Context, HipTracer, LBVHBuilder, RayBuffer, Scene, _ptr


def backward(ctx, vertices, triangles, rays, results, dl_dt=None, dl_du=None, dl_dv=None, out=None):
    """dL/dvertices (float32 (V,3)) for closest-hit results; accumulates into `out` if given."""
    if out is None:
        out = torch.zeros_like(vertices)
    n = int(rays.shape[0])

    def opt(x):
        return None if x is None else _ptr(x.contiguous())

    keep = [x.contiguous() if x is not None else None for x in (dl_dt, dl_du, dl_dv)]
    call("hx_backward", ctx.handle, _ptr(vertices), _ptr(triangles), _ptr(rays), _ptr(results), n,
         opt(keep[0]), opt(keep[1]), opt(keep[2]), _ptr(out))
    return out


class _ClosestHit(torch.autograd.Function):
    @staticmethod
    def forward(fctx, vertices, tracer, rays):
        tr = tracer
        v = vertices.detach().contiguous()
        if tuple(v.shape) != tuple(tr.scene.vertices.shape):
            raise ValueError(f"vertices of shape {tuple(v.shape)} do not match the traced mesh's "
                             f"{tuple(tr.scene.vertices.shape)}")
        call("hx_refit", tr.ctx.handle, tr.bvh.handle, _ptr(v), int(v.shape[0]), _ptr(tr.triangles),
             int(tr.triangles.shape[0]))
        call("hx_trace", tr.ctx.handle, tr.bvh.handle, _ptr(rays.rays), rays.size, 0, _ptr(rays.results), None)
        res = rays.results.clone()
        n = rays.size
        t = torch.empty(n, dtype=torch.float32, device=v.device)
        u = torch.empty_like(t)
        w = torch.empty_like(t)
        # (t, u, v) from the hit triangle by the function the adjoint differentiates
        call("hx_recompute_hits", tr.ctx.handle, _ptr(v), _ptr(tr.triangles), _ptr(rays.rays), _ptr(res), n,
             _ptr(t), _ptr(u), _ptr(w))
        ids = res[:, 0].clone()
        fctx.save_for_backward(v, rays.rays, res)
        fctx.tracer = tr
        fctx.mark_non_differentiable(ids)
        return t, u, w, ids

    @staticmethod
    def backward(fctx, gt, gu, gv, gid):
        v, rays, res = fctx.saved_tensors
        tr = fctx.tracer
        grad = backward(tr.ctx, v, tr.triangles, rays, res,
                        None if gt is None else gt.float(),
                        None if gu is None else gu.float(),
                        None if gv is None else gv.float())
        return grad, None, None


class DifferentiableTracer:
    """Build once, then per step: refit to the current vertices, trace, and
    backpropagate into the vertices.

        dt = DifferentiableTracer(vertices0, triangles)
        t, u, v, ids = dt(vertices, rays)        # vertices.requires_grad_()
        loss = ((t - t_target)[mask] ** 2).sum(); loss.backward()
    """

    def __init__(self, vertices, triangles, ctx=None, **build_params):
        self.ctx = ctx or Context.default()
        self.scene = Scene(vertices.detach(), triangles)
        self.triangles = self.scene.triangles
        self.bvh = LBVHBuilder(self.ctx, **build_params).build(self.scene)

    def __call__(self, vertices, rays):
        rb = rays if isinstance(rays, RayBuffer) else RayBuffer(rays)
        if vertices.dtype != torch.float32 or not vertices.is_cuda:
            raise ValueError("vertices must be a float32 CUDA tensor")
        return _ClosestHit.apply(vertices, self, rb)

    def tracer(self):
        return HipTracer(self.ctx, self.bvh)


__all__ = ["backward", "DifferentiableTracer"]
