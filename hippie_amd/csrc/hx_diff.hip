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
// gradients are merged per workgroup in LDS and scatter-added with hardware
// FP32 atomics.
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

// Merge of the vertex contributions before the global atomics, in two levels:
// neighbouring lanes that hit the same triangle are summed in registers (a
// segmented DPP scan), then the rays of one workgroup that hit the same or
// adjacent triangles (coherent camera rays hit each triangle several times)
// add into one LDS slot per vertex, and each distinct vertex costs three
// global atomics instead of three per ray.  Open addressing over kAggSlots slots; an insert that finds no slot
// within kAggProbes probes adds straight to global memory, so incoherent rays
// cost a few LDS probes more and nothing else.
constexpr int kAggSlots = 1024;
constexpr int kDiffRaysPerThread = 4;

// The three vertex contributions of one ray (or of a run of rays).
struct Contrib {
    float g[9];
    HX_DEV void add(const Contrib& o) {
#pragma unroll
        for (int k = 0; k < 9; ++k) g[k] += o.g[k];
    }
};
constexpr int kAggProbes = 8;
constexpr int32_t kAggEmpty = -1;

HX_DEV void add_vertex(int32_t* keys, float* vx, float* vy, float* vz, float* g, int vid, float x, float y, float z) {
    uint32_t h = ((uint32_t)vid * 2654435761u) >> (32 - 10);
#pragma unroll 1
    for (int probe = 0; probe < kAggProbes; ++probe) {
        const int32_t old = atomicCAS(&keys[h], kAggEmpty, vid);
        if (old == kAggEmpty || old == vid) {
            atomicAdd(&vx[h], x);
            atomicAdd(&vy[h], y);
            atomicAdd(&vz[h], z);
            return;
        }
        h = (h + 1) & (kAggSlots - 1);
    }
    atomicAdd(g + 3 * (int64_t)vid + 0, x);
    atomicAdd(g + 3 * (int64_t)vid + 1, y);
    atomicAdd(g + 3 * (int64_t)vid + 2, z);
}

__global__ __launch_bounds__(kDiffBlock) void k_backward(BackwardArgs a) {
    static_assert(kAggSlots == 1 << 10, "the hash takes 10 bits");
    __shared__ int32_t s_key[kAggSlots];
    __shared__ float s_x[kAggSlots], s_y[kAggSlots], s_z[kAggSlots];
    for (int i = threadIdx.x; i < kAggSlots; i += kDiffBlock) {
        s_key[i] = kAggEmpty;
        s_x[i] = 0.0f;
        s_y[i] = 0.0f;
        s_z[i] = 0.0f;
    }
    __syncthreads();
    float* g = a.dl_dverts;
    const int lane = lane_id();
    // Consecutive rays of kDiffRaysPerThread block-wide slices share the table.
    for (int slice = 0; slice < kDiffRaysPerThread; ++slice) {
        const int64_t r = (blockIdx.x * (int64_t)kDiffRaysPerThread + slice) * kDiffBlock + threadIdx.x;
        const int id = r < a.n_rays ? a.results[r].x : -1;
        double gt = 0.0, gu = 0.0, gv = 0.0;
        if (id >= 0) {
            gt = a.dl_dt ? (double)a.dl_dt[r] : 0.0;
            gu = a.dl_du ? (double)a.dl_du[r] : 0.0;
            gv = a.dl_dv ? (double)a.dl_dv[r] : 0.0;
        }
        // Per ray: its triangle (-1 without a contribution) and the three vertex
        // contributions beta_k * lambda.
        Hit h;
        int tri = -1;
        Contrib cv{};
        if ((gt != 0.0 || gu != 0.0 || gv != 0.0)) {
            tri = hit_triangle(a, r, id);
            if (solve_hit(a, r, tri, h)) {
                const double w = 1.0 - h.u - h.v;
                const D3 lam{-(h.bc.x * gu + h.ca.x * gv + h.ab.x * gt) * h.inv,
                             -(h.bc.y * gu + h.ca.y * gv + h.ab.y * gt) * h.inv,
                             -(h.bc.z * gu + h.ca.z * gv + h.ab.z * gt) * h.inv};
                const double beta[3] = {h.u, h.v, w};
    #pragma unroll
                for (int k = 0; k < 3; ++k) {
                    cv.g[3 * k + 0] = (float)(beta[k] * lam.x);
                    cv.g[3 * k + 1] = (float)(beta[k] * lam.y);
                    cv.g[3 * k + 2] = (float)(beta[k] * lam.z);
                }
            } else {
                tri = -1;
            }
        }
        // Neighbouring lanes that hit the same triangle (runs inside each row of
        // 16 lanes) are summed by a segmented DPP scan; the run's last lane holds
        // the sum and adds it for all of them.
        const int prev = __shfl_up(tri, 1, kWave);
        const uint64_t heads = ballot64((lane & 15) == 0 || tri != prev);
        const uint64_t upto = lane == kWave - 1 ? ~0ull : ((2ull << lane) - 1ull);
        const int dist = lane - (63 - __clzll(heads & upto));
        Contrib o = dpp_move<kDppRowShr1, 0xF>(cv);
        if (dist >= 1) cv.add(o);
        o = dpp_move<kDppRowShr2, 0xF>(cv);
        if (dist >= 2) cv.add(o);
        o = dpp_move<kDppRowShr4, 0xF>(cv);
        if (dist >= 4) cv.add(o);
        o = dpp_move<kDppRowShr8, 0xF>(cv);
        if (dist >= 8) cv.add(o);
        const bool last = lane == kWave - 1 || ((heads >> (lane + 1)) & 1ull);
        if (last && tri >= 0) {
    #pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int vid = k == 0 ? h.i0 : k == 1 ? h.i1 : h.i2;
                add_vertex(s_key, s_x, s_y, s_z, g, vid, cv.g[3 * k], cv.g[3 * k + 1], cv.g[3 * k + 2]);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAggSlots; i += kDiffBlock) {
        const int vid = s_key[i];
        if (vid == kAggEmpty) continue;
        atomicAdd(g + 3 * (int64_t)vid + 0, s_x[i]);
        atomicAdd(g + 3 * (int64_t)vid + 1, s_y[i]);
        atomicAdd(g + 3 * (int64_t)vid + 2, s_z[i]);
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
    const int64_t per_block = (int64_t)kDiffBlock * kDiffRaysPerThread;
    hipLaunchKernelGGL(k_backward, dim3((unsigned)((a.n_rays + per_block - 1) / per_block)), dim3(kDiffBlock), 0,
                       c->stream, a);
    return check_launch("backward");
}

}  // namespace hx
