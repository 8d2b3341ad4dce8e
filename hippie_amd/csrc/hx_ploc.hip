// hx_ploc.hip -- PLOC clustering on gfx950 (parallel locally-ordered
// clustering over the Morton-sorted leaves), and the relabelling that turns
// its tree into one whose leaves are numbered in depth-first order, so the
// shared climb and scan-based collapse of hx_build.hip apply unchanged.
//
// Reference (src/hippie/rt/bvh/): PLOCBuilder.cpp:21-208 (clustering loop and
// build), PLOCBuilderKernels.cu:11-113 (generateNeighboursCached), :116-205
// (generateNeighbours), :207-286 (merge), :288-389 (prefix scan + compact).
//
// Per round with m clusters (cluster order = Morton order of the leaves):
//   k_ploc_neighbours  nearest neighbour of every cluster in [i-r, i+r] by the
//                      surface area of the union box; the leftmost minimum
//                      wins (strict '>' over a left-to-right scan);
//   k_ploc_tile_counts merges (mutual pair, lower index) and survivors
//                      (everything but the higher index of a mutual pair) per
//                      4096-cluster tile;
//   k_ploc_scan        exclusive scan of the tile counts, totals for the host;
//   k_ploc_merge       new node = (m-2) - (merge rank in cluster order): the
//                      reference takes this rank from a warp atomicAdd, here it
//                      is a scan so the labels are deterministic; survivors are
//                      written compacted, with their boxes, in cluster order.
// The cluster boxes travel with the clusters (coalesced arrays in cluster
// order), so no round gathers boxes through node indices.
//
// Relabelling (k_tree_down, one launch per round in reverse round order, so
// every parent is placed before its children): offsets of the two children
// from the subtree sizes the merges recorded; a leaf child is renumbered to
// n-1+(its depth-first rank) and its triangle id is written at that rank.
#include <algorithm>
#include <vector>

#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kPlocBatch = 16;  // rounds launched between host reads
constexpr int kMaxLdsRadius = 1024;   // window of 256 + 2r boxes, 6 floats each, in LDS

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

struct Sum2 {
    HX_DEV int2 operator()(int2 a, int2 b) const { return make_int2(a.x + b.x, a.y + b.y); }
};

HX_DEV float3 ld3(const float* p) { return make_float3(p[0], p[1], p[2]); }

// Leaves: box of the sorted triangle p (setupLeaves, LBVHBuilderKernels.cu
// :162-195) as cluster p, node n-1+p.  Also clears what the climb expects
// cleared (counters, cover, frontier count).
__global__ __launch_bounds__(kBlock) void k_ploc_leaves(PlocArgs a) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    const int n = (int)a.n;
    if (p >= n) return;
    Box3 b = box_empty();
    if (a.ref_min) {
        const float4 rmn = a.ref_min[a.sorted_tri[p]], rmx = a.ref_max[a.sorted_tri[p]];
        b.mn = make_float3(rmn.x, rmn.y, rmn.z);
        b.mx = make_float3(rmx.x, rmx.y, rmx.z);
    } else {
        const int32_t* tr = a.tris + 3 * (int64_t)a.sorted_tri[p];
        box_grow(b, ld3(a.verts + 3 * (int64_t)tr[0]));
        box_grow(b, ld3(a.verts + 3 * (int64_t)tr[1]));
        box_grow(b, ld3(a.verts + 3 * (int64_t)tr[2]));
    }
    a.cl[0][p] = n - 1 + p;
    if (a.cm[0]) a.cm[0][p] = a.mesh_of[p];
    a.ca[0][p] = make_float4(b.mn.x, b.mn.y, b.mn.z, b.mx.x);
    a.cb[0][p] = make_float2(b.mx.y, b.mx.z);
    a.end_at[p] = -1;
    if (p < n - 1) a.counters[p] = 0;
    if (p == 0) *a.frontier_count = 0;
}

// area(union) in the reference's order: the neighbour's box grown by the
// cluster's min corner, then by its max corner (AABB::grow, util/AABB.h:26-31).
HX_DEV float union_area(float jx0, float jy0, float jz0, float jx1, float jy1, float jz1, const Box3& own) {
    float mnx = mins(mins(jx0, own.mn.x), own.mx.x), mny = mins(mins(jy0, own.mn.y), own.mx.y);
    float mnz = mins(mins(jz0, own.mn.z), own.mx.z);
    float mxx = maxs(maxs(jx1, own.mn.x), own.mx.x), mxy = maxs(maxs(jy1, own.mn.y), own.mx.y);
    float mxz = maxs(maxs(jz1, own.mn.z), own.mx.z);
    return box_area(make_float3(mnx, mny, mnz), make_float3(mxx, mxy, mxz));
}

HX_DEV Box3 cluster_box(float4 a, float2 b) {
    Box3 r;
    r.mn = make_float3(a.x, a.y, a.z);
    r.mx = make_float3(a.w, b.x, b.y);
    return r;
}

// Nearest cluster of i within [i - r, i + r] (same mesh if cmesh), the
// lowest index on ties, from global memory.
HX_DEV int nearest_cluster(const float4* __restrict__ ca, const float2* __restrict__ cb, int m, int r, int i,
                           const int32_t* __restrict__ cmesh) {
    const Box3 own = cluster_box(ca[i], cb[i]);
    int best = -1;
    float bestD = kMaxFloat;
    const int jlo = max(0, i - r), jhi = min(m - 1, i + r);
    const int mine = cmesh ? cmesh[i] : 0;
    for (int j = jlo; j <= jhi; ++j) {
        if (j == i) continue;
        if (cmesh && cmesh[j] != mine) continue;
        const float4 x = ca[j];
        const float2 y = cb[j];
        const float d = union_area(x.x, x.y, x.z, x.w, y.x, y.y, own);
        if (bestD > d) {
            bestD = d;
            best = j;
        }
    }
    return best;
}

// LDS window: boxes of clusters [lo, lo + W) as six rows of stride `stride`
// (a seventh row of mesh ids in batched builds).  With cmesh, a cluster only
// looks at clusters of its own mesh: the meshes' cluster runs are contiguous,
// so this is the single-mesh window clipped to the mesh.
// The round's cluster count m is read from its record (rounds run ahead of the
// host); a finished build (m <= 1) makes the round a no-op.  Batched builds
// pass the cluster mesh ids: once every mesh is one cluster (m <= n_meshes)
// the round pairs clusters (2k, 2k+1) instead (see launch_ploc).
template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_ploc_neighbours(const float4* __restrict__ ca,
                                                            const float2* __restrict__ cb,
                                                            const int4* __restrict__ round, int r,
                                                            int32_t* __restrict__ nb,
                                                            const int32_t* __restrict__ cmesh_in, int n_meshes) {
    extern __shared__ float s_win[];
    const int m = round->x;
    const int b0 = blockIdx.x * kBlock;
    const int i = b0 + threadIdx.x;
    if (m <= 1 || b0 >= m) return;
    if (cmesh_in && m <= n_meshes) {
        if (i < m) nb[i] = (i ^ 1) < m ? (i ^ 1) : -1;
        return;
    }
    const int32_t* cmesh = cmesh_in;
    const int lo = max(0, b0 - r);
    if (LDS) {
        const int hi = min(m, b0 + kBlock + r);
        const int stride = kBlock + 2 * r;
        for (int k = threadIdx.x; k < hi - lo; k += kBlock) {
            const float4 x = ca[lo + k];
            const float2 y = cb[lo + k];
            s_win[0 * stride + k] = x.x;
            s_win[1 * stride + k] = x.y;
            s_win[2 * stride + k] = x.z;
            s_win[3 * stride + k] = x.w;
            s_win[4 * stride + k] = y.x;
            s_win[5 * stride + k] = y.y;
            if (cmesh) s_win[6 * stride + k] = __int_as_float(cmesh[lo + k]);
        }
        __syncthreads();
        if (i >= m) return;
        const int me = i - lo;
        Box3 own;
        own.mn = make_float3(s_win[me], s_win[stride + me], s_win[2 * stride + me]);
        own.mx = make_float3(s_win[3 * stride + me], s_win[4 * stride + me], s_win[5 * stride + me]);
        int best = -1;
        float bestD = kMaxFloat;
        const int jlo = max(0, i - r), jhi = min(m - 1, i + r);
        const int mine = cmesh ? __float_as_int(s_win[6 * stride + me]) : 0;
        for (int j = jlo; j <= jhi; ++j) {
            if (j == i) continue;
            const int q = j - lo;
            if (cmesh && __float_as_int(s_win[6 * stride + q]) != mine) continue;
            const float d = union_area(s_win[q], s_win[stride + q], s_win[2 * stride + q], s_win[3 * stride + q],
                                       s_win[4 * stride + q], s_win[5 * stride + q], own);
            if (bestD > d) {
                bestD = d;
                best = j;
            }
        }
        nb[i] = best;
    } else {
        if (i >= m) return;
        nb[i] = nearest_cluster(ca, cb, m, r, i, cmesh);
    }
}

// (merging, surviving) of cluster i (PLOCBuilderKernels.cu:232-240): -- a mutual pair merges
// into the lower index; the higher index drops out of the compacted list.
HX_DEV int2 ploc_flags(const int32_t* __restrict__ nb, int m, int64_t i) {
    if (i >= m) return make_int2(0, 0);
    const int j = nb[i];
    const bool mutual = j >= 0 && nb[j] == (int)i;
    return make_int2(mutual && i < j ? 1 : 0, (mutual && i > j) ? 0 : 1);
}

__global__ __launch_bounds__(kBlock) void k_ploc_tile_counts(const int32_t* __restrict__ nb,
                                                             const int4* __restrict__ round,
                                                             int2* __restrict__ partial) {
    const int m = round->x;
    const int64_t base = blockIdx.x * (int64_t)kTile;
    if (m <= 1 || base >= m) return;
    int2 c = make_int2(0, 0);
    for (int k = 0; k < kItems; ++k) {
        const int2 f = ploc_flags(nb, m, base + k * kBlock + threadIdx.x);
        c.x += f.x;
        c.y += f.y;
    }
    c.x = wave_sum_i32(c.x);
    c.y = wave_sum_i32(c.y);
    __shared__ int2 ws[kWavesPerBlock];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int2 s = ws[0];
        for (int w = 1; w < kWavesPerBlock; ++w) s = make_int2(s.x + ws[w].x, s.y + ws[w].y);
        partial[blockIdx.x] = s;
    }
}

// Single block: exclusive scan of the tile counts in place; the round's
// (merged, surviving) totals go to its record and the survivors become the
// next round's cluster count.
__global__ __launch_bounds__(kBlock) void k_ploc_scan(int2* __restrict__ partial, int4* __restrict__ round) {
    __shared__ int2 wt[kWavesPerBlock];
    const int m = round->x;
    if (m <= 1) {
        if (threadIdx.x == 0) {
            round->y = 0;
            round->z = m;
            round[1] = make_int4(m, 0, 0, 0);
        }
        return;
    }
    const int ntiles = (m + kTile - 1) / kTile;
    int2 carry = make_int2(0, 0);
    for (int base = 0; base < ntiles; base += kBlock) {
        const int t = base + threadIdx.x;
        const int2 v = t < ntiles ? partial[t] : make_int2(0, 0);
        int2 tot;
        const int2 inc = block_inclusive_scan<kWavesPerBlock>(v, Sum2(), wt, tot);
        if (t < ntiles) partial[t] = make_int2(carry.x + inc.x - v.x, carry.y + inc.y - v.y);
        carry = make_int2(carry.x + tot.x, carry.y + tot.y);
    }
    if (threadIdx.x == 0) {
        round->y = carry.x;
        round->z = carry.y;
        round[1] = make_int4(carry.y, 0, 0, 0);
    }
}

// Cluster i of a round with flags f (merging, surviving) and ranks among the
// merging and surviving clusters: a merging cluster creates node
// node_offset - mrank from itself and its neighbour; survivors move to slot
// srank of the other cluster buffer.
HX_DEV void merge_cluster(const PlocArgs& a, int src, int i, int2 f, int mrank, int srank, int node_offset) {
    const int n = (int)a.n;
    const int32_t* __restrict__ cl = a.cl[src];
    const float4* __restrict__ ca = a.ca[src];
    const float2* __restrict__ cb = a.cb[src];
    int32_t* __restrict__ cl_out = a.cl[src ^ 1];
    float4* __restrict__ ca_out = a.ca[src ^ 1];
    float2* __restrict__ cb_out = a.cb[src ^ 1];
    if (f.x) {
        const int j = a.nb[i];
        const int ln = cl[i], rn = cl[j];
        const int node = node_offset - mrank;
        const Box3 lb = cluster_box(ca[i], cb[i]);
        const Box3 rb = cluster_box(ca[j], cb[j]);
        a.parent[ln] = node;
        a.parent[rn] = node;
        a.left[node] = ln;
        a.right[node] = rn;
        const int sl = ln >= n - 1 ? 1 : a.last[ln];
        const int sr = rn >= n - 1 ? 1 : a.last[rn];
        a.last[node] = sl + sr;
        cl_out[srank] = node;
        if (a.cm[0]) a.cm[src ^ 1][srank] = a.cm[src][i];
        ca_out[srank] = make_float4(mins(lb.mn.x, rb.mn.x), mins(lb.mn.y, rb.mn.y), mins(lb.mn.z, rb.mn.z),
                                    maxs(lb.mx.x, rb.mx.x));
        cb_out[srank] = make_float2(maxs(lb.mx.y, rb.mx.y), maxs(lb.mx.z, rb.mx.z));
    } else if (f.y) {
        cl_out[srank] = cl[i];
        if (a.cm[0]) a.cm[src ^ 1][srank] = a.cm[src][i];
        ca_out[srank] = ca[i];
        cb_out[srank] = cb[i];
    }
}

// merge (PLOCBuilderKernels.cu:207-286) + compact (:288-389) in one pass.
// The new node's box is the select-based union of the two cluster boxes
// (the reference uses fminf/fmaxf, which differ from selects only in the sign
// of a zero); its subtree size goes to last[node] for the relabelling.
__global__ __launch_bounds__(kBlock) void k_ploc_merge(PlocArgs a, int src, const int4* __restrict__ round,
                                                       const int2* __restrict__ partial) {
    __shared__ int2 wt[kWavesPerBlock];
    const int m = round->x;
    if (m <= 1 || blockIdx.x * (int64_t)kTile >= m) return;
    const int node_offset = m - 2;
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int2 carry = partial[blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kBlock + threadIdx.x;
        const int2 f = ploc_flags(a.nb, m, i);
        int2 tot;
        const int2 inc = block_inclusive_scan<kWavesPerBlock>(f, Sum2(), wt, tot);
        if (i < m) merge_cluster(a, src, (int)i, f, carry.x + inc.x - f.x, carry.y + inc.y - f.y, node_offset);
        carry = make_int2(carry.x + tot.x, carry.y + tot.y);
    }
}

__global__ void k_ploc_round_start(int4* round, int m) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *round = make_int4(m, 0, 0, 0);
}

// The last rounds (at most kTailClusters clusters) in one workgroup: the same
// neighbour search (over boxes staged in LDS), flags, ranks and merge as the
// grid-wide kernels, with a barrier between the phases instead of a launch
// each.  Each thread ranks a
// contiguous chunk of clusters.  rec[k] = (clusters, merged, surviving) of
// tail round k; *nrec = rounds run (a round that merges nothing ends it).
constexpr int kTailBlock = 1024;
constexpr int kTailClusters = 2048;

__global__ __launch_bounds__(kTailBlock) void k_ploc_tail(PlocArgs a, int src, int m, int r, int4* __restrict__ rec,
                                                          int* __restrict__ nrec) {
    __shared__ int2 wt[kTailBlock / kWave];
    // The round's cluster boxes (six rows) and mesh ids, staged once per round.
    __shared__ float s_box[6][kTailClusters];
    __shared__ int s_mesh[kTailClusters];
    const bool batched = a.cm[0] != nullptr && a.n_meshes > 1;
    int k = 0;
    while (m > 1) {
        if (a.adaptive) r = min(r + 1, a.max_radius);
        const bool pairs = batched && m <= a.n_meshes;
        const bool by_mesh = batched && !pairs;
        for (int i = threadIdx.x; i < m; i += kTailBlock) {
            const float4 x = a.ca[src][i];
            const float2 y = a.cb[src][i];
            s_box[0][i] = x.x;
            s_box[1][i] = x.y;
            s_box[2][i] = x.z;
            s_box[3][i] = x.w;
            s_box[4][i] = y.x;
            s_box[5][i] = y.y;
            if (by_mesh) s_mesh[i] = a.cm[src][i];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += kTailBlock) {
            if (pairs) {
                a.nb[i] = (i ^ 1) < m ? (i ^ 1) : -1;
                continue;
            }
            Box3 own;
            own.mn = make_float3(s_box[0][i], s_box[1][i], s_box[2][i]);
            own.mx = make_float3(s_box[3][i], s_box[4][i], s_box[5][i]);
            int best = -1;
            float bestD = kMaxFloat;
            const int jlo = max(0, i - r), jhi = min(m - 1, i + r);
            for (int j = jlo; j <= jhi; ++j) {
                if (j == i || (by_mesh && s_mesh[j] != s_mesh[i])) continue;
                const float d = union_area(s_box[0][j], s_box[1][j], s_box[2][j], s_box[3][j], s_box[4][j],
                                           s_box[5][j], own);
                if (bestD > d) {
                    bestD = d;
                    best = j;
                }
            }
            a.nb[i] = best;
        }
        __syncthreads();
        const int chunk = (m + kTailBlock - 1) / kTailBlock;
        const int lo = threadIdx.x * chunk, hi = min(m, lo + chunk);
        int2 cnt = make_int2(0, 0);
        for (int i = lo; i < hi; ++i) {
            const int2 f = ploc_flags(a.nb, m, i);
            cnt = make_int2(cnt.x + f.x, cnt.y + f.y);
        }
        int2 tot;
        const int2 inc = block_inclusive_scan<kTailBlock / kWave>(cnt, Sum2(), wt, tot);
        int mrank = inc.x - cnt.x, srank = inc.y - cnt.y;
        if (tot.x > 0) {
            for (int i = lo; i < hi; ++i) {
                const int2 f = ploc_flags(a.nb, m, i);
                merge_cluster(a, src, i, f, mrank, srank, m - 2);
                mrank += f.x;
                srank += f.y;
            }
        }
        if (threadIdx.x == 0) rec[k] = make_int4(m, tot.x, tot.y, 0);
        ++k;
        __syncthreads();
        if (tot.x == 0) break;
        m = tot.y;
        src ^= 1;
    }
    if (threadIdx.x == 0) *nrec = k;
}

// Depth-first placement of the children of the nodes list[0, count) (or, if
// list is NULL, of lo + [0, count)), all independent: every parent was placed
// by an earlier launch.  last[v] holds v's subtree size on entry and its last
// leaf rank on exit; first[v] was written by v's parent (the root starts at 0).
HX_DEV void tree_down_node(const PlocArgs& a, int v) {
    const int n = (int)a.n;
    const int off = v == 0 ? 0 : a.first[v];
    const int size = a.last[v];
    const int l = a.left[v], r = a.right[v];
    const int sl = l >= n - 1 ? 1 : a.last[l];
    if (v == 0) {
        a.first[0] = 0;
        a.parent[0] = -1;
    }
    a.last[v] = off + size - 1;
    a.split[v] = off + sl - 1;
    const int ch[2] = {l, r};
    const int co[2] = {off, off + sl};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = ch[k], o = co[k];
        // A consistent tree keeps every index in range; a broken one must not
        // write outside the arrays.
        if ((unsigned)c >= (unsigned)(2 * n - 1) || (unsigned)o >= (unsigned)n) continue;
        if (c >= n - 1) {
            const int leaf = n - 1 + o;
            a.tri_out[o] = a.sorted_tri[c - (n - 1)];
            a.parent[leaf] = v;
            a.first[leaf] = o;
            a.last[leaf] = o;
            a.left[leaf] = o;
            a.right[leaf] = o + 1;
            if (k == 0)
                a.left[v] = leaf;
            else
                a.right[v] = leaf;
        } else {
            a.first[c] = o;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_tree_down(PlocArgs a, const int32_t* __restrict__ list, int lo,
                                                      int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    tree_down_node(a, list ? list[k] : lo + k);
}

// Several placement steps in one workgroup, a barrier between steps (the
// barrier orders the global writes of a step before the next step's reads
// within the workgroup).
constexpr int kStepBlock = 1024;
constexpr int kSmallStep = 4 * kStepBlock;  // steps up to this many nodes share the workgroup

__global__ __launch_bounds__(kStepBlock) void k_tree_down_steps(PlocArgs a, const int32_t* __restrict__ list,
                                                                const int2* __restrict__ segs, int nsegs) {
    for (int k = 0; k < nsegs; ++k) {
        const int2 sg = segs[k];
        for (int j = threadIdx.x; j < sg.y; j += kStepBlock) tree_down_node(a, list ? list[sg.x + j] : sg.x + j);
        __syncthreads();
    }
}

}  // namespace

int launch_tree_down_steps(Context* c, const PlocArgs& a, const int32_t* list, const std::vector<int2>& segs) {
    size_t small = 0;
    while (small < segs.size() && segs[small].y <= kSmallStep) ++small;
    if (small > 1) {
        HX_TRY(c->segments.reserve(sizeof(int2) * small));
        c->segments_host.assign(segs.begin(), segs.begin() + small);
        HX_HIP(hipMemcpyAsync(c->segments.ptr, c->segments_host.data(), sizeof(int2) * small, hipMemcpyHostToDevice,
                              c->stream));
        hipLaunchKernelGGL(k_tree_down_steps, dim3(1), dim3(kStepBlock), 0, c->stream, a, list,
                           c->segments.as<int2>(), (int)small);
        HX_TRY(check_launch("tree_down_steps"));
    } else {
        small = 0;
    }
    for (size_t k = small; k < segs.size(); ++k) {
        if (segs[k].y <= 0) continue;
        hipLaunchKernelGGL(k_tree_down, dim3(grid_for(segs[k].y)), dim3(kBlock), 0, c->stream, a,
                           list ? list + segs[k].x : nullptr, list ? 0 : segs[k].x, segs[k].y);
        HX_TRY(check_launch("tree_down"));
    }
    return HX_OK;
}

int launch_ploc(Context* c, const PlocArgs& a, int* rounds_out) {
    const int n = (int)a.n;
    {
        LaunchScope ls(c, "ploc_leaves");
        hipLaunchKernelGGL(k_ploc_leaves, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
        HX_TRY(check_launch("ploc_leaves"));
    }
    const int64_t max_tiles = (n + kTile - 1) / kTile;
    HX_TRY(c->scan_partials.reserve(sizeof(int2) * (size_t)max_tiles + sizeof(int4) * (kPlocBatch + kTailClusters + 2) +
                                    64));
    int2* partial = c->scan_partials.as<int2>();
    int4* rounds = reinterpret_cast<int4*>(partial + ((max_tiles + 1) & ~(int64_t)1));
    int4* tail_rec = rounds + kPlocBatch + 1;
    int* tail_count = reinterpret_cast<int*>(tail_rec + kTailClusters);
    std::vector<int2> ranges;   // [lo, hi] of the nodes each round created
    // Batched builds merge inside each mesh until every mesh is one cluster
    // (m == n_meshes), then join the mesh roots pairwise in mesh order
    // (separators, never emitted), so mesh m's leaves keep positions
    // [off[m], off[m+1]) after the depth-first relabelling.
    const bool batched = a.cm[0] != nullptr && a.n_meshes > 1;
    // Rounds run on the device in batches of kPlocBatch: each reads its cluster
    // count from the previous round's record, and the host reads the records
    // once per batch (grids are sized by the count known at the batch start).
    int m = n, src = 0, r = a.radius;
    std::vector<int4> h_rounds(kPlocBatch + 1);
    hipLaunchKernelGGL(k_ploc_round_start, dim3(1), dim3(64), 0, c->stream, rounds, m);
    HX_TRY(check_launch("ploc_round_start"));
    while (m > 1) {
        if (m <= kTailClusters) {
            LaunchScope ls(c, "ploc_tail");
            hipLaunchKernelGGL(k_ploc_tail, dim3(1), dim3(kTailBlock), 0, c->stream, a, src, m, r, tail_rec,
                               tail_count);
            HX_TRY(check_launch("ploc_tail"));
            int cnt = 0;
            HX_HIP(hipMemcpyAsync(&cnt, tail_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HX_HIP(hipStreamSynchronize(c->stream));
            std::vector<int4> rec((size_t)std::max(cnt, 0));
            if (cnt > 0) {
                HX_HIP(hipMemcpyAsync(rec.data(), tail_rec, sizeof(int4) * rec.size(), hipMemcpyDeviceToHost,
                                      c->stream));
                HX_HIP(hipStreamSynchronize(c->stream));
            }
            for (const int4& t : rec) {
                if (t.y <= 0 || t.y + t.z != t.x) {
                    set_error("PLOC round with %d clusters merged %d pairs (non-finite boxes?)", t.x, t.y);
                    return HX_ERR_INVALID_ARG;
                }
                ranges.push_back(make_int2(t.x - 2 - t.y + 1, t.x - 2));
            }
            break;
        }
        const int ntiles = (m + kTile - 1) / kTile;
        for (int b = 0; b < kPlocBatch; ++b) {
            if (a.adaptive) r = std::min(r + 1, a.max_radius);
            {
                LaunchScope ls(c, "ploc_neighbours");
                const size_t lds = sizeof(float) * (batched ? 7 : 6) * (size_t)(kBlock + 2 * r);
                const int32_t* cmesh = batched ? a.cm[src] : nullptr;
                if (r <= kMaxLdsRadius && lds <= 65536) {
                    hipLaunchKernelGGL(k_ploc_neighbours<true>, dim3(grid_for(m)), dim3(kBlock), lds, c->stream,
                                       a.ca[src], a.cb[src], rounds + b, r, a.nb, cmesh, a.n_meshes);
                } else {
                    hipLaunchKernelGGL(k_ploc_neighbours<false>, dim3(grid_for(m)), dim3(kBlock), 0, c->stream,
                                       a.ca[src], a.cb[src], rounds + b, r, a.nb, cmesh, a.n_meshes);
                }
                HX_TRY(check_launch("ploc_neighbours"));
            }
            {
                LaunchScope ls(c, "ploc_merge");
                hipLaunchKernelGGL(k_ploc_tile_counts, dim3(ntiles), dim3(kBlock), 0, c->stream, a.nb, rounds + b,
                                   partial);
                HX_TRY(check_launch("ploc_tile_counts"));
                hipLaunchKernelGGL(k_ploc_scan, dim3(1), dim3(kBlock), 0, c->stream, partial, rounds + b);
                HX_TRY(check_launch("ploc_scan"));
                hipLaunchKernelGGL(k_ploc_merge, dim3(ntiles), dim3(kBlock), 0, c->stream, a, src, rounds + b,
                                   partial);
                HX_TRY(check_launch("ploc_merge"));
            }
            src ^= 1;
        }
        HX_HIP(hipMemcpyAsync(h_rounds.data(), rounds, sizeof(int4) * h_rounds.size(), hipMemcpyDeviceToHost,
                              c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
        for (int b = 0; b < kPlocBatch; ++b) {
            const int4 t = h_rounds[b];
            if (t.x <= 1) break;
            if (t.y <= 0 || t.y + t.z != t.x) {
                set_error("PLOC round with %d clusters merged %d pairs (non-finite boxes?)", t.x, t.y);
                return HX_ERR_INVALID_ARG;
            }
            ranges.push_back(make_int2(t.x - 2 - t.y + 1, t.x - 2));
        }
        m = h_rounds[kPlocBatch].x;
        if (m > 1) {
            hipLaunchKernelGGL(k_ploc_round_start, dim3(1), dim3(64), 0, c->stream, rounds, m);
            HX_TRY(check_launch("ploc_round_start"));
        }
    }
    {
        LaunchScope ls(c, "ploc_relabel");
        std::vector<int2> segs;  // rounds in reverse order: parents first
        for (size_t k = ranges.size(); k-- > 0;) segs.push_back(make_int2(ranges[k].x, ranges[k].y - ranges[k].x + 1));
        HX_TRY(launch_tree_down_steps(c, a, nullptr, segs));
    }
    if (rounds_out) *rounds_out = (int)ranges.size();
    return HX_OK;
}

}  // namespace hx
