// hx_diff.hip -- adjoint of the closest hit with respect to triangle vertices.
//
// No reference counterpart (the reference has no gradients).  The forward
// trace stores the original triangle id and (t, u, v) per ray; backward
// recomputes the hit from that id instead of saving traversal state:
//
//   o + t d = v2 + u (v0 - v2) + v (v1 - v2),  M = [e1 e2 -d],  x = (u,v,t) = M^-1 (o - v2)
//
// Differentiating, dx = -M^-1 (u dv0 + v dv1 + (1-u-v) dv2), so with the
// upstream gradient g = (dL/du, dL/dv, dL/dt) and lambda = -M^-T g:
//
//   dL/dv0 = u lambda,  dL/dv1 = v lambda,  dL/dv2 = (1-u-v) lambda.
//
// M^-1 has rows (e2 x c, c x e1, e1 x e2) / det with c = -d, det = e1.(e2 x c).
// The solve runs in FP64 (cheap on MI355X, and it keeps the adjoint exact to
// FP32 rounding of the result even for nearly grazing rays); the vertex
// gradients are scatter-added with hardware FP32 atomics.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kDiffBlock = 256;

struct D3 {
    double x, y, z;
};
HX_DEV D3 ld_vert(const float* v, int i) {
    return D3{(double)v[3 * (int64_t)i], (double)v[3 * (int64_t)i + 1], (double)v[3 * (int64_t)i + 2]};
}
HX_DEV D3 sub(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
HX_DEV D3 cross(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
HX_DEV double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// The hit of ray r on triangle id, in FP64: x = (u, v, t) = M^-1 (o - v2) with
// M = [e1 e2 -d]; also returns the rows of M^-1 (scaled by det) and 1/det.
struct Hit {
    int i0, i1, i2;
    double u, v, t, inv;
    D3 bc, ca, ab;
};

HX_DEV bool solve_hit(const BackwardArgs& a, int64_t r, int id, Hit& h) {
    const int32_t* tr = a.tris + 3 * (int64_t)id;
    h.i0 = tr[0];
    h.i1 = tr[1];
    h.i2 = tr[2];
    const D3 v0 = ld_vert(a.verts, h.i0), v1 = ld_vert(a.verts, h.i1), v2 = ld_vert(a.verts, h.i2);
    const float4 o = a.rays[2 * r], d = a.rays[2 * r + 1];
    const D3 e1 = sub(v0, v2), e2 = sub(v1, v2);
    const D3 c{-(double)d.x, -(double)d.y, -(double)d.z};
    const D3 b{(double)o.x - v2.x, (double)o.y - v2.y, (double)o.z - v2.z};
    h.bc = cross(e2, c);
    h.ca = cross(c, e1);
    h.ab = cross(e1, e2);
    const double det = dot(e1, h.bc);
    if (det == 0.0) return false;
    h.inv = 1.0 / det;
    h.u = dot(h.bc, b) * h.inv;
    h.v = dot(h.ca, b) * h.inv;
    h.t = dot(h.ab, b) * h.inv;
    return true;
}

// Triangle of a hit: results hold mesh-local ids in a batched trace.
HX_DEV int hit_triangle(const BackwardArgs& a, int64_t r, int id) {
    return a.ray_mesh ? a.mesh_off[a.ray_mesh[r]] + id : id;
}

__global__ __launch_bounds__(kDiffBlock) void k_backward(BackwardArgs a) {
    const int64_t r = blockIdx.x * (int64_t)kDiffBlock + threadIdx.x;
    if (r >= a.n_rays) return;
    const int id = a.results[r].x;
    if (id < 0) return;
    const double gt = a.dl_dt ? (double)a.dl_dt[r] : 0.0;
    const double gu = a.dl_du ? (double)a.dl_du[r] : 0.0;
    const double gv = a.dl_dv ? (double)a.dl_dv[r] : 0.0;
    if (gt == 0.0 && gu == 0.0 && gv == 0.0) return;
    Hit h;
    if (!solve_hit(a, r, hit_triangle(a, r, id), h)) return;
    const double w = 1.0 - h.u - h.v;
    const D3 lam{-(h.bc.x * gu + h.ca.x * gv + h.ab.x * gt) * h.inv, -(h.bc.y * gu + h.ca.y * gv + h.ab.y * gt) * h.inv,
                 -(h.bc.z * gu + h.ca.z * gv + h.ab.z * gt) * h.inv};
    float* g = a.dl_dverts;
    const double beta[3] = {h.u, h.v, w};
    const int vid[3] = {h.i0, h.i1, h.i2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicAdd(g + 3 * (int64_t)vid[k] + 0, (float)(beta[k] * lam.x));
        atomicAdd(g + 3 * (int64_t)vid[k] + 1, (float)(beta[k] * lam.y));
        atomicAdd(g + 3 * (int64_t)vid[k] + 2, (float)(beta[k] * lam.z));
    }
}

// Forward companion: (t, u, v) of each hit recomputed from its triangle by the
// function the adjoint differentiates, rounded once to FP32.  Misses keep t
// from the trace (tmax) and get u = v = 0.
__global__ __launch_bounds__(kDiffBlock) void k_recompute_hits(BackwardArgs a, float* __restrict__ t_out,
                                                               float* __restrict__ u_out, float* __restrict__ v_out) {
    const int64_t r = blockIdx.x * (int64_t)kDiffBlock + threadIdx.x;
    if (r >= a.n_rays) return;
    const int4 res = a.results[r];
    float t = __int_as_float(res.y), u = 0.0f, v = 0.0f;
    Hit h;
    if (res.x >= 0 && solve_hit(a, r, hit_triangle(a, r, res.x), h)) {
        t = (float)h.t;
        u = (float)h.u;
        v = (float)h.v;
    }
    t_out[r] = t;
    u_out[r] = u;
    v_out[r] = v;
}

}  // namespace

int launch_recompute_hits(Context* c, const BackwardArgs& a, float* t_out, float* u_out, float* v_out) {
    if (a.n_rays <= 0) return HX_OK;
    LaunchScope ls(c, "recompute_hits");
    hipLaunchKernelGGL(k_recompute_hits, dim3((unsigned)((a.n_rays + kDiffBlock - 1) / kDiffBlock)), dim3(kDiffBlock),
                       0, c->stream, a, t_out, u_out, v_out);
    return check_launch("recompute_hits");
}

int launch_backward(Context* c, const BackwardArgs& a) {
    if (a.n_rays <= 0) return HX_OK;
    LaunchScope ls(c, "backward");
    hipLaunchKernelGGL(k_backward, dim3((unsigned)((a.n_rays + kDiffBlock - 1) / kDiffBlock)), dim3(kDiffBlock), 0,
                       c->stream, a);
    return check_launch("backward");
}

}  // namespace hx
