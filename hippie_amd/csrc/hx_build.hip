// hx_build.hip -- gfx950 kernels of the LBVH build: scene box, Morton codes,
// Karras hierarchy, one fused bottom-up pass (refit + Woop + SAH collapse
// decision) and a scan-based deterministic collapse to the 64-byte Bin layout.
//
// Reference (src/hippie/): rt/bvh/LBVHBuilderKernels.cu, BVHCollapserKernels.cu,
// HipBVHKernels.cu, rt/HipUtil.h, rt/bvh/HipBVHUtil.h.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

HX_DEV float3 ld3(const float* p) { return make_float3(p[0], p[1], p[2]); }

HX_DEV void tri_vertices(const float* verts, const int32_t* tris, int64_t t, float3& a, float3& b, float3& c) {
    const int32_t* tr = tris + 3 * t;
    int i0 = tr[0], i1 = tr[1], i2 = tr[2];
    a = ld3(verts + 3 * (int64_t)i0);
    b = ld3(verts + 3 * (int64_t)i1);
    c = ld3(verts + 3 * (int64_t)i2);
}

HX_DEV Box3 tri_box(float3 a, float3 b, float3 c) {
    Box3 bx = box_empty();
    box_grow(bx, a);
    box_grow(bx, b);
    box_grow(bx, c);
    return bx;
}

// ---------------------------------------------------------------------------
// Scene box: exact min/max reduction (order-independent up to the sign of a
// zero, which does not change any Morton code or area).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_scene_box_partial(const float* __restrict__ verts,
                                                               const int32_t* __restrict__ tris, int64_t n,
                                                               float* __restrict__ partial) {
    float r[6] = {kMaxFloat, kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat, -kMaxFloat};
    for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
        float3 a, b, c;
        if (tris) {
            tri_vertices(verts, tris, t, a, b, c);
        } else {  // the vertex box: element t is vertex t
            a = b = c = ld3(verts + 3 * t);
        }
        r[0] = fminf(r[0], fminf(fminf(a.x, b.x), c.x));
        r[1] = fminf(r[1], fminf(fminf(a.y, b.y), c.y));
        r[2] = fminf(r[2], fminf(fminf(a.z, b.z), c.z));
        r[3] = fmaxf(r[3], fmaxf(fmaxf(a.x, b.x), c.x));
        r[4] = fmaxf(r[4], fmaxf(fmaxf(a.y, b.y), c.y));
        r[5] = fmaxf(r[5], fmaxf(fmaxf(a.z, b.z), c.z));
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], __shfl_xor(r[k], m, kWave));
#pragma unroll
        for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], __shfl_xor(r[k], m, kWave));
    }
    __shared__ float s[kWavesPerBlock][6];
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int k = 0; k < 6; ++k) s[w][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s[0][k];
        for (int i = 1; i < kWavesPerBlock; ++i) v = (k < 3) ? fminf(v, s[i][k]) : fmaxf(v, s[i][k]);
        partial[blockIdx.x * 6 + k] = v;
    }
}

// One block: every thread folds a strided share of the partials, then the
// block reduces (min/max are exact, so the order does not matter).
__global__ __launch_bounds__(kBlock) void k_scene_box_final(const float* __restrict__ partial, int nblocks,
                                                            float* __restrict__ box) {
    float r[6] = {kMaxFloat, kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat, -kMaxFloat};
    for (int i = threadIdx.x; i < nblocks; i += kBlock) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], partial[i * 6 + k]);
#pragma unroll
        for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], partial[i * 6 + k]);
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], __shfl_xor(r[k], m, kWave));
#pragma unroll
        for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], __shfl_xor(r[k], m, kWave));
    }
    __shared__ float s[kWavesPerBlock][6];
    if (lane_id() == 0)
        for (int k = 0; k < 6; ++k) s[threadIdx.x / kWave][k] = r[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float v = s[0][k];
        for (int w = 1; w < kWavesPerBlock; ++w) v = (k < 3) ? fminf(v, s[w][k]) : fmaxf(v, s[w][k]);
        box[k] = v;
    }
}

// Scene normalisation (SceneLoader.cpp:134-141): v = (v - mn) * scale with
// scale = 1 / the largest extent of the vertex box (AssimpImporter.cpp:99-113
// grows it over every vertex), subtraction and product rounded separately.
__global__ __launch_bounds__(kBlock) void k_normalize(float* __restrict__ verts, int64_t nv,
                                                     const float* __restrict__ box) {
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (i >= nv) return;
    const float dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
    const float dxy = dx < dy ? dy : dx;  // qMax(a, b) = a < b ? b : a
    const float scale = 1.0f / (dxy < dz ? dz : dxy);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float d = verts[3 * i + k] - box[k];
        verts[3 * i + k] = d * scale;
    }
}

// ---------------------------------------------------------------------------
// Morton codes: setupBoxes + computeMortonCodes60 fused (LBVHBuilderKernels.cu
// :65-98,131-160).  The per-triangle box is recomputed, never stored.
// ---------------------------------------------------------------------------

// mesh_of (batched builds): each triangle is quantised in its own mesh's box.
template <bool K64>
__global__ __launch_bounds__(kBlock) void k_morton(const float* __restrict__ verts, const int32_t* __restrict__ tris,
                                                   int64_t n, const float* __restrict__ boxes, int shift,
                                                   void* __restrict__ keys_out, const int32_t* __restrict__ mesh_of,
                                                   const float4* __restrict__ ref_min,
                                                   const float4* __restrict__ ref_max,
                                                   float4* __restrict__ pack_out) {
    const int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (t >= n) return;
    const float* sb = mesh_of ? boxes + 6 * (int64_t)mesh_of[t] : boxes;
    const float3 smn = make_float3(sb[0], sb[1], sb[2]);
    const float3 scale = make_float3(1.0f / (sb[3] - sb[0]), 1.0f / (sb[4] - sb[1]), 1.0f / (sb[5] - sb[2]));
    Box3 tb;
    if (ref_min) {
        const float4 rmn = ref_min[t], rmx = ref_max[t];
        tb.mn = make_float3(rmn.x, rmn.y, rmn.z);
        tb.mx = make_float3(rmx.x, rmx.y, rmx.z);
    } else {
        float3 a, b, c;
        tri_vertices(verts, tris, t, a, b, c);
        tb = tri_box(a, b, c);
        if (pack_out) {
            pack_out[3 * t + 0] = make_float4(a.x, a.y, a.z, b.x);
            pack_out[3 * t + 1] = make_float4(b.y, b.z, c.x, c.y);
            pack_out[3 * t + 2] = make_float4(c.z, 0.0f, 0.0f, 0.0f);
        }
    }
    // computeMortonCodes60 regrows the stored box from its two corners.
    Box3 bb = box_empty();
    box_grow(bb, tb.mn);
    box_grow(bb, tb.mx);
    const float cx = (bb.mn.x + bb.mx.x) * 0.5f;
    const float cy = (bb.mn.y + bb.mx.y) * 0.5f;
    const float cz = (bb.mn.z + bb.mx.z) * 0.5f;
    const uint64_t code = morton60((cx - smn.x) * scale.x, (cy - smn.y) * scale.y, (cz - smn.z) * scale.z) >> shift;
    if (K64)
        static_cast<uint64_t*>(keys_out)[t] = code;
    else
        static_cast<uint32_t*>(keys_out)[t] = (uint32_t)code;
}

// ---------------------------------------------------------------------------
// Karras 2012 hierarchy (LBVHBuilderKernels.cu:11-63,226-252; delta from
// rt/bvh/HipBVHUtil.h:17-33).  32-bit keys use clz32 with the tie value
// shifted by the same 32 as every other delta, which keeps the tree identical.
// Also initialises the climb counters and the collapse cover array, and writes
// the leaf entries setupLeaves gives left/right (:185-187).
// ---------------------------------------------------------------------------

// Batched builds (mesh_of != NULL) order by the composite key (mesh, Morton):
// positions in different meshes get deltas below 32 (the mesh id bits), and
// inside a mesh every delta is shifted by 32 with the tie broken on the
// mesh-local index, so each mesh's subtree is exactly its own Karras tree.
template <bool K64>
HX_DEV int kdelta(int i, int j, int n, const void* keys, const int32_t* mesh_of, const int32_t* off) {
    if (j < 0 || j >= n) return -1;
    int ti = i, tj = j, shift = 0;
    if (mesh_of) {
        const int mi = mesh_of[i], mj = mesh_of[j];
        if (mi != mj) return __clz((uint32_t)(mi ^ mj));
        ti = i - off[mi];
        tj = j - off[mi];
        shift = 32;
    }
    if (K64) {
        const uint64_t* k = static_cast<const uint64_t*>(keys);
        const uint64_t a = k[i], b = k[j];
        return shift + (a != b ? __clzll(a ^ b) : __clzll((uint64_t)(uint32_t)(ti ^ tj)) + 64);
    } else {
        const uint32_t* k = static_cast<const uint32_t*>(keys);
        const uint32_t a = k[i], b = k[j];
        return shift + (a != b ? __clz(a ^ b) : __clz((uint32_t)(ti ^ tj)) + 64);
    }
}

template <bool K64>
__global__ __launch_bounds__(kBlock) void k_construct(const void* __restrict__ keys, int n, int32_t* __restrict__ left,
                                                      int32_t* __restrict__ right, int32_t* __restrict__ parent,
                                                      int32_t* __restrict__ first, int32_t* __restrict__ last,
                                                      int32_t* __restrict__ counters, int32_t* __restrict__ end_at,
                                                      int32_t* __restrict__ zero_word,
                                                      const int32_t* __restrict__ mesh_of,
                                                      const int32_t* __restrict__ off) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    end_at[i] = -1;
    left[n - 1 + i] = i;
    right[n - 1 + i] = i + 1;
    first[n - 1 + i] = i;
    last[n - 1 + i] = i;
    if (i == 0) {
        parent[0] = -1;
        *zero_word = 0;
    }
    if (i >= n - 1) return;
    counters[i] = 0;
    const int dn = kdelta<K64>(i, i + 1, n, keys, mesh_of, off);
    const int dp = kdelta<K64>(i, i - 1, n, keys, mesh_of, off);
    const int d = (dn > dp) - (dn < dp);
    const int dmin = d > 0 ? dp : dn;
    int lmax = 2;
    while (kdelta<K64>(i, i + lmax * d, n, keys, mesh_of, off) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta<K64>(i, i + (l + t) * d, n, keys, mesh_of, off) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta<K64>(i, j, n, keys, mesh_of, off);
    int s = 0, k = 2, t;
    do {
        t = (l + k - 1) / k;
        k <<= 1;
        if (kdelta<K64>(i, i + (s + t) * d, n, keys, mesh_of, off) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    int lc = gamma, rc = gamma + 1;
    if (min(i, j) == gamma) lc += n - 1;
    if (max(i, j) == gamma + 1) rc += n - 1;
    left[i] = lc;
    right[i] = rc;
    first[i] = min(i, j);
    last[i] = max(i, j);
    parent[lc] = i;
    parent[rc] = i;
}

// ---------------------------------------------------------------------------
// Woop transform: rt/bvh/HipBVHKernels.cu:12-33 with util/Math.h:318-369
// (cofactor inverse, d summing all 16 cofactor products, r * rcp(d) * 4).
// ---------------------------------------------------------------------------

HX_DEV float det3(float a00, float a01, float a02, float a10, float a11, float a12, float a20, float a21,
                  float a22) {
    return a00 * a11 * a22 - a00 * a12 * a21 + a10 * a21 * a02 - a10 * a22 * a01 + a20 * a01 * a12 -
           a20 * a02 * a11;
}

HX_DEV void woop_rows(float3 v0, float3 v1, float3 v2, float4& w0, float4& w1, float4& w2) {
    const float ax = v0.x - v2.x, ay = v0.y - v2.y, az = v0.z - v2.z;
    const float bx = v1.x - v2.x, by = v1.y - v2.y, bz = v1.z - v2.z;
    const float nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    // M[row][col]: columns (v0-v2, v1-v2, n, (v2, 1)).
    const float M[4][4] = {{ax, bx, nx, v2.x}, {ay, by, ny, v2.y}, {az, bz, nz, v2.z}, {0.0f, 0.0f, 0.0f, 1.0f}};
    float R[4][4];
    float d = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float s[3][3];
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) s[k][l] = M[(k < j) ? k : k + 1][(l < i) ? l : l + 1];
            const float sgn = ((i + j) & 1) ? -1.0f : 1.0f;
            const float dd = det3(s[0][0], s[0][1], s[0][2], s[1][0], s[1][1], s[1][2], s[2][0], s[2][1], s[2][2]) * sgn;
            R[i][j] = dd;
            d += dd * M[j][i];
        }
    }
    const float rd = (d != 0.0f) ? 1.0f / d : 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) R[a][b] = (R[a][b] * rd) * 4.0f;
    w0 = make_float4(R[2][0], R[2][1], R[2][2], -R[2][3]);
    w1 = make_float4(R[0][0], R[0][1], R[0][2], R[0][3]);
    w2 = make_float4(R[1][0], R[1][1], R[1][2], R[1][3]);
}

// ---------------------------------------------------------------------------
// Fused bottom-up pass: leaf box (setupLeaves), Woop rows (woopifyTriangles),
// leaf SAH cost, then the termination-counter climb in which the second
// arriver at each internal node unions the child boxes (refit,
// LBVHBuilderKernels.cu:254-305) and takes the collapse decision
// (computeNodeStatesAdaptive / computeNodeStates, BVHCollapserKernels.cu
// :58-184).  A state-0 node other than the root records its range end at its
// range start (atomicMax), the input of the scan-based collapse below.
//
// Two levels, because each XCD of gfx950 has its own L2: an agent-scope
// acquire/release per climb step writes back and invalidates that L2.
//   level 1 (k_climb_local): one workgroup per 512 consecutive sorted leaves
//     climbs every node whose range lies inside its leaves, with boxes, costs
//     and counters in LDS and workgroup-scope atomics.  A node whose parent
//     straddles the block boundary is appended to the frontier list.
//   level 2 (k_climb_top): the ~2n/512 straddling nodes, one thread per
//     frontier entry; the exchanged boxes and costs are written and read with
//     agent-coherent (sc1) relaxed atomic accesses, and each writer waits for
//     its stores to complete before it bumps the parent's counter, so no L2
//     writeback or invalidate is needed.
// ---------------------------------------------------------------------------

constexpr int kClimbThreads = 512;  // one leaf per thread
constexpr int kClimbLeaves = 512;

struct Combined {
    float3 mn, mx;
    float cost;
    int state;
};

// Union of the two child boxes in the reference's grow order, and the
// collapse decision for node [first, last].
HX_DEV Combined combine(const ClimbArgs& a, int first, int last, float3 lmn, float3 lmx, float lcost, float3 rmn,
                        float3 rmx, float rcost) {
    Combined c;
    Box3 bx = box_empty();
    box_grow(bx, lmn);
    box_grow(bx, rmn);
    c.mn = bx.mn;
    box_grow(bx, lmx);
    box_grow(bx, rmx);
    c.mx = bx.mx;
    const int size = last - first + 1;
    c.cost = 0.0f;
    if (a.adaptive) {
        Box3 bb = box_empty();
        box_grow(bb, c.mn);
        box_grow(bb, c.mx);
        const float area = box_area(bb.mn, bb.mx);
        const float cost = a.ct * area + lcost + rcost;
        const float leafCost = a.ci * area * (float)size;
        c.state = (leafCost < cost) ? 0 : 1;
        c.cost = c.state == 0 ? leafCost : cost;
    } else {
        c.state = size <= a.max_leaf_size ? 0 : 1;
    }
    return c;
}

HX_DEV float3 xyz(float4 v) { return make_float3(v.x, v.y, v.z); }

// A node that must stay an interior whatever its SAH state: the root (of each
// mesh, in a batched build) and, in a batched build, the separator nodes
// above the mesh roots (ranges spanning meshes), which are never emitted.
HX_DEV bool seg_top(const int32_t* mesh_of, const int32_t* off, int n, int first, int last) {
    if (!mesh_of) return first == 0 && last == n - 1;
    const int m = mesh_of[first];
    return mesh_of[last] != m || (first == off[m] && last == off[m + 1] - 1);
}

// Surface area in the treelet kernels' order (TRBuilderUtil.h:112-130:
// x*y + x*z + y*z, no validity test), used by the ATR preparation.
HX_DEV float atr_area(float3 mn, float3 mx) {
    const float x = mx.x - mn.x, y = mx.y - mn.y, z = mx.z - mn.z;
    return 2.0f * (x * y + x * z + y * z);
}

// Bin rows 0-2 of a node from its two child boxes (HipBVHKernels.cu layout:
// x and y extents of each child, then both z extents).
HX_DEV void bin_rows(float4 amn, float4 amx, float4 bmn, float4 bmx, int4& r0, int4& r1, int4& r2) {
    r0 = make_int4(__float_as_int(amn.x), __float_as_int(amx.x), __float_as_int(amn.y), __float_as_int(amx.y));
    r1 = make_int4(__float_as_int(bmn.x), __float_as_int(bmx.x), __float_as_int(bmn.y), __float_as_int(bmx.y));
    r2 = make_int4(__float_as_int(amn.z), __float_as_int(amx.z), __float_as_int(bmn.z), __float_as_int(bmx.z));
}

// Refit of a Karras-tree Bin BVH: rows 0-2 of the Bin node of binary node
// `node`, if it has one -- an interior's from its children's boxes, a leaf's
// from its own box twice (the emit's formulas; row 3 does not change).
HX_DEV void refit_bin_rows(const ClimbArgs& a, int node, float3 mn, float3 mx, float3 lmn, float3 lmx, float3 rmn,
                           float3 rmx) {
    const int b = a.bin_of[node];
    if (b == -1) return;
    int4 r0, r1, r2;
    if (b >= 0)
        bin_rows(make_float4(lmn.x, lmn.y, lmn.z, 0.0f), make_float4(lmx.x, lmx.y, lmx.z, 0.0f),
                 make_float4(rmn.x, rmn.y, rmn.z, 0.0f), make_float4(rmx.x, rmx.y, rmx.z, 0.0f), r0, r1, r2);
    else
        bin_rows(make_float4(mn.x, mn.y, mn.z, 0.0f), make_float4(mx.x, mx.y, mx.z, 0.0f),
                 make_float4(mn.x, mn.y, mn.z, 0.0f), make_float4(mx.x, mx.y, mx.z, 0.0f), r0, r1, r2);
    int4* o = a.bin_nodes + 4 * (int64_t)(b >= 0 ? b : -2 - b);
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
}

// SAH cost of a one-reference leaf with box [mn, mx] (0 without adaptive
// collapse).
HX_DEV float leaf_sah_cost(const ClimbArgs& a, float3 mn, float3 mx) {
    if (!a.adaptive) return 0.0f;
    Box3 bb = box_empty();
    box_grow(bb, mn);
    box_grow(bb, mx);
    return a.ci * box_area(bb.mn, bb.mx);
}

// Leaf p of the sorted order: its box (the reference box when presplit),
// Woop rows and SAH leaf cost; the Woop rows are stored, and with STORE the
// box and cost to the leaf's node entries.
template <bool STORE = true>
HX_DEV Box3 climb_leaf(const ClimbArgs& a, int p, float& cost) {
    const int n = (int)a.n;
    const int id = a.sorted_tri[p];
    const int tri = a.ref_tri ? a.ref_tri[id] : id;
    float3 v0, v1, v2;
    if (a.tri_pack) {
        const float4 q0 = a.tri_pack[3 * (int64_t)tri], q1 = a.tri_pack[3 * (int64_t)tri + 1];
        const float4 q2 = a.tri_pack[3 * (int64_t)tri + 2];
        v0 = make_float3(q0.x, q0.y, q0.z);
        v1 = make_float3(q0.w, q1.x, q1.y);
        v2 = make_float3(q1.z, q1.w, q2.x);
    } else {
        tri_vertices(a.verts, a.tris, tri, v0, v1, v2);
    }
    Box3 tb;
    if (a.ref_min) {
        const float4 rmn = a.ref_min[id], rmx = a.ref_max[id];
        tb.mn = make_float3(rmn.x, rmn.y, rmn.z);
        tb.mx = make_float3(rmx.x, rmx.y, rmx.z);
    } else {
        tb = tri_box(v0, v1, v2);
    }
    const int leaf = n - 1 + p;
    if (STORE) {
        a.nmin[leaf] = make_float4(tb.mn.x, tb.mn.y, tb.mn.z, 0.0f);
        a.nmax[leaf] = make_float4(tb.mx.x, tb.mx.y, tb.mx.z, 0.0f);
    }
    float4 w0, w1, w2;
    woop_rows(v0, v1, v2, w0, w1, w2);
    a.woop[3 * (int64_t)p + 0] = w0;
    a.woop[3 * (int64_t)p + 1] = w1;
    a.woop[3 * (int64_t)p + 2] = w2;
    cost = leaf_sah_cost(a, tb.mn, tb.mx);
    if (STORE && a.adaptive) a.cost[leaf] = cost;
    return tb;
}

// GENERAL: the tree is not a Karras tree (PLOC, ATRBVH), so an internal
// node's LDS slot is its split position a.split[node] (the last leaf of its
// left child), which lies in [first, last-1]; a Karras node's index already
// does.  ATR: also produce the ATRBVH inputs -- per node the treelet-order
// area, the binary SAH cost (leaf ci*A, interior ct*A + children;
// ATRBuilderKernels.cu:271-309), the subtree size and the height.
template <bool GENERAL, bool ATR>
__global__ __launch_bounds__(kClimbThreads) void k_climb_local(ClimbArgs a, int32_t* __restrict__ frontier,
                                                              int32_t* __restrict__ frontier_count) {
    constexpr int T = kClimbLeaves;
    // Slots: internal node i -> key(i) - base (in [0, T-1)), leaf p -> T + p - base.
    __shared__ float s_box[6][2 * T];
    __shared__ float s_cost[2 * T];
    __shared__ int s_cnt[T];
    __shared__ float s_ac[ATR ? 2 * T : 1];
    __shared__ int s_h[ATR ? 2 * T : 1];
    // Karras trees: the topology of the internal nodes [base, base + T) --
    // every node this block can finish -- staged once, coalesced, so the
    // climb's dependent reads hit LDS instead of HBM.
    constexpr int TT = GENERAL ? 1 : T;
    __shared__ int s_first[TT], s_last[TT], s_left[TT], s_right[TT], s_par[TT];
    __shared__ int s_state[TT];
    // Nodes whose parent straddles the block (at most one per leaf) and the
    // collapse input end_at for the block's leaf positions: gathered in LDS and
    // written once per block instead of one global atomic per entry.
    __shared__ int s_front[T];
    __shared__ int s_end[T];
    __shared__ int s_nfront, s_fbase;
    const int n = (int)a.n;
    const int base = blockIdx.x * T;
    const int last_leaf = min(base + T, n) - 1;
    if (threadIdx.x == 0) s_nfront = 0;
    for (int i = threadIdx.x; i < T; i += kClimbThreads) {
        s_cnt[i] = 0;
        s_end[i] = -1;
        if (!GENERAL && base + i < n - 1) {
            s_first[i] = a.first[base + i];
            s_last[i] = a.last[base + i];
            s_left[i] = a.left[base + i];
            s_right[i] = a.right[base + i];
            s_par[i] = a.parent[base + i];
        }
    }
    int leaf_parent[T / kClimbThreads];

#pragma unroll
    for (int k = 0; k < T / kClimbThreads; ++k) {
        const int p = base + k * kClimbThreads + threadIdx.x;
        if (p > last_leaf) break;
        leaf_parent[k] = a.parent[n - 1 + p];
        float cost;
        const Box3 tb = climb_leaf(a, p, cost);
        const int leaf = n - 1 + p;
        const int sl = T + p - base;
        s_box[0][sl] = tb.mn.x; s_box[1][sl] = tb.mn.y; s_box[2][sl] = tb.mn.z;
        s_box[3][sl] = tb.mx.x; s_box[4][sl] = tb.mx.y; s_box[5][sl] = tb.mx.z;
        s_cost[sl] = cost;
        if (ATR) {
            const float ar = atr_area(tb.mn, tb.mx);
            a.atr_area[leaf] = ar;
            a.atr_cost[leaf] = a.ci * ar;
            s_ac[sl] = a.ci * ar;
            s_h[sl] = 0;
        }
    }
    __syncthreads();

#pragma unroll
    for (int k = 0; k < T / kClimbThreads; ++k) {
        const int p = base + k * kClimbThreads + threadIdx.x;
        if (p > last_leaf) break;
        int child = n - 1 + p;
        int node = leaf_parent[k];
        while (node >= 0) {
            const int li = node - base;
            // A Karras node finishing in this block has its index in the block.
            if (!GENERAL && (unsigned)li >= (unsigned)T) {
                s_front[atomicAdd(&s_nfront, 1)] = child;
                break;
            }
            const int first = GENERAL ? a.first[node] : s_first[li];
            const int last = GENERAL ? a.last[node] : s_last[li];
            if (first < base || last > last_leaf) {
                s_front[atomicAdd(&s_nfront, 1)] = child;
                break;
            }
            const int sn = GENERAL ? a.split[node] - base : li;
            if (__hip_atomic_fetch_add(&s_cnt[sn], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
            const int l = GENERAL ? a.left[node] : s_left[li];
            const int r = GENERAL ? a.right[node] : s_right[li];
            const int ls = l >= n - 1 ? T + (l - (n - 1) - base) : (GENERAL ? a.split[l] : l) - base;
            const int rs = r >= n - 1 ? T + (r - (n - 1) - base) : (GENERAL ? a.split[r] : r) - base;
            const Combined c = combine(
                a, first, last, make_float3(s_box[0][ls], s_box[1][ls], s_box[2][ls]),
                make_float3(s_box[3][ls], s_box[4][ls], s_box[5][ls]), s_cost[ls],
                make_float3(s_box[0][rs], s_box[1][rs], s_box[2][rs]),
                make_float3(s_box[3][rs], s_box[4][rs], s_box[5][rs]), s_cost[rs]);
            s_box[0][sn] = c.mn.x; s_box[1][sn] = c.mn.y; s_box[2][sn] = c.mn.z;
            s_box[3][sn] = c.mx.x; s_box[4][sn] = c.mx.y; s_box[5][sn] = c.mx.z;
            s_cost[sn] = c.cost;
            if (ATR) {
                s_ac[sn] = a.ct * atr_area(c.mn, c.mx) + s_ac[ls] + s_ac[rs];
                s_h[sn] = 1 + max(s_h[ls], s_h[rs]);
            }
            if (GENERAL) {
                a.nmin[node] = make_float4(c.mn.x, c.mn.y, c.mn.z, 0.0f);
                a.nmax[node] = make_float4(c.mx.x, c.mx.y, c.mx.z, 0.0f);
                if (!a.refit_only) {
                    if (a.adaptive) a.cost[node] = c.cost;
                    a.state[node] = c.state;
                }
            } else {
                s_state[li] = c.state;
            }
            if (!a.refit_only && c.state == 0 && !seg_top(a.mesh_of, a.mesh_off, n, first, last))
                atomicMax(&s_end[first - base], last);
            child = node;
            node = GENERAL ? a.parent[node] : s_par[li];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_fbase = atomicAdd(frontier_count, s_nfront);
    __syncthreads();
    for (int i = threadIdx.x; i < s_nfront; i += kClimbThreads) frontier[s_fbase + i] = s_front[i];
    if (!a.refit_only)
        for (int i = threadIdx.x; i < T && base + i <= last_leaf; i += kClimbThreads) a.end_at[base + i] = s_end[i];
    if (GENERAL) return;
    // Karras trees: the nodes this block finished (both children arrived,
    // s_cnt == 2) are the block's index range; write them out coalesced.
    if (a.bin_nodes)
        for (int i = threadIdx.x; i < T && base + i <= last_leaf; i += kClimbThreads) {
            const int sl = T + i;
            const float3 mn = make_float3(s_box[0][sl], s_box[1][sl], s_box[2][sl]);
            const float3 mx = make_float3(s_box[3][sl], s_box[4][sl], s_box[5][sl]);
            refit_bin_rows(a, n - 1 + base + i, mn, mx, mn, mx, mn, mx);
        }
    for (int i = threadIdx.x; i < T; i += kClimbThreads) {
        const int node = base + i;
        if (node >= n - 1 || s_cnt[i] != 2) continue;
        const float3 mn = make_float3(s_box[0][i], s_box[1][i], s_box[2][i]);
        const float3 mx = make_float3(s_box[3][i], s_box[4][i], s_box[5][i]);
        a.nmin[node] = make_float4(mn.x, mn.y, mn.z, 0.0f);
        a.nmax[node] = make_float4(mx.x, mx.y, mx.z, 0.0f);
        if (a.bin_nodes) {
            // Both children finished in this block: their boxes are in LDS.
            const int l = s_left[i], r = s_right[i];
            const int ls = l >= n - 1 ? T + (l - (n - 1) - base) : l - base;
            const int rs = r >= n - 1 ? T + (r - (n - 1) - base) : r - base;
            refit_bin_rows(a, node, mn, mx, make_float3(s_box[0][ls], s_box[1][ls], s_box[2][ls]),
                           make_float3(s_box[3][ls], s_box[4][ls], s_box[5][ls]),
                           make_float3(s_box[0][rs], s_box[1][rs], s_box[2][rs]),
                           make_float3(s_box[3][rs], s_box[4][rs], s_box[5][rs]));
        }
        if (ATR) {
            a.atr_area[node] = atr_area(mn, mx);
            a.atr_cost[node] = s_ac[i];
            a.atr_height[node] = s_h[i];
            a.atr_refs[node] = s_last[i] - s_first[i] + 1;
        }
        if (!a.refit_only) {
            if (a.adaptive) a.cost[node] = s_cost[i];
            a.state[node] = s_state[i];
        }
    }
}

// Agent-coherent (sc1) relaxed accesses of a float4 as two 64-bit words.
HX_DEV float4 ld_agent(float4* p) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    const uint64_t lo = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float((uint32_t)lo), __uint_as_float((uint32_t)(lo >> 32)),
                       __uint_as_float((uint32_t)hi), __uint_as_float((uint32_t)(hi >> 32)));
}
HX_DEV void st_agent(float4* p, float4 v) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    const uint64_t lo = (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32);
    const uint64_t hi = (uint64_t)__float_as_uint(v.z) | ((uint64_t)__float_as_uint(v.w) << 32);
    __hip_atomic_store(q, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
HX_DEV float ld_agent(float* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
HX_DEV void st_agent(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
HX_DEV int ld_agent(int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
HX_DEV void st_agent(int32_t* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Stores issued so far complete before the next memory operation (the
// compiler fences keep them above the wait).
HX_DEV void complete_stores() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

template <bool ATR>
__global__ __launch_bounds__(kBlock) void k_climb_top(ClimbArgs a, const int32_t* __restrict__ frontier,
                                                     const int32_t* __restrict__ frontier_count) {
    const int count = *frontier_count;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < count; e += gridDim.x * kBlock) {
        int node = a.parent[frontier[e]];
        while (node >= 0) {
            if (__hip_atomic_fetch_add(&a.counters[node], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
            const int first = a.first[node], last = a.last[node];
            const int l = a.left[node], r = a.right[node];
            const float4 lmn = ld_agent(a.nmin + l), lmx = ld_agent(a.nmax + l);
            const float4 rmn = ld_agent(a.nmin + r), rmx = ld_agent(a.nmax + r);
            const float lc = a.adaptive ? ld_agent(a.cost + l) : 0.0f;
            const float rc = a.adaptive ? ld_agent(a.cost + r) : 0.0f;
            const Combined c = combine(a, first, last, xyz(lmn), xyz(lmx), lc, xyz(rmn), xyz(rmx), rc);
            st_agent(a.nmin + node, make_float4(c.mn.x, c.mn.y, c.mn.z, 0.0f));
            st_agent(a.nmax + node, make_float4(c.mx.x, c.mx.y, c.mx.z, 0.0f));
            if (a.bin_nodes) refit_bin_rows(a, node, c.mn, c.mx, xyz(lmn), xyz(lmx), xyz(rmn), xyz(rmx));
            if (ATR) {
                const int n = (int)a.n;
                const float ar = atr_area(c.mn, c.mx);
                const float ac = a.ct * ar + ld_agent(a.atr_cost + l) + ld_agent(a.atr_cost + r);
                const int hl = l >= n - 1 ? 0 : __hip_atomic_load(a.atr_height + l, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
                const int hr = r >= n - 1 ? 0 : __hip_atomic_load(a.atr_height + r, __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
                a.atr_area[node] = ar;
                st_agent(a.atr_cost + node, ac);
                __hip_atomic_store(a.atr_height + node, 1 + max(hl, hr), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                a.atr_refs[node] = last - first + 1;
            }
            if (!a.refit_only) {
                if (a.adaptive) st_agent(a.cost + node, c.cost);
                a.state[node] = c.state;
                if (c.state == 0 && !seg_top(a.mesh_of, a.mesh_off, (int)a.n, first, last))
                    atomicMax(&a.end_at[first], last);
            }
            // Complete this node's stores before the parent's counter can be
            // observed (the compiler fence keeps the stores above the wait).
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            node = a.parent[node];
        }
    }
}

// ---------------------------------------------------------------------------
// LBVH hierarchy built during the climb (Apetrei's bottom-up formulation of
// the Karras binary radix tree; the reference builds it top-down in
// construct, LBVHBuilderKernels.cu:152-252, then climbs).  With
// delta(q) = common-prefix length of sorted keys q and q+1 (index tie-break
// and mesh bits as in kdelta; -1 outside [0, n-2]), node [l, r] is the left
// child of the node split at r when delta(r) > delta(l-1), and the right
// child of the node split at l-1 otherwise (the two never tie for distinct
// augmented keys).  The tree is the Karras tree; internal nodes are numbered
// by split position s (left child ends at s), which lies inside their range,
// so the two-level scheme of the climb applies unchanged:
//   k_lbvh_local: one workgroup per 512 leaves; deltas of its range in LDS;
//     each arrival at an in-block parent leaves its id and range end in LDS,
//     the second arrival combines.  A child whose parent straddles the block
//     (split outside the block, or only one child arrived) joins the frontier.
//   k_lbvh_top: the frontier climbs with the same rule on global memory, each
//     arrival writing its side (child id, range end) with sc1 stores that
//     complete before the parent's counter is bumped.
// ---------------------------------------------------------------------------

// One leaf per thread, 512 leaves per block: measured at 16.7M triangles,
// 256- and 1024-leaf blocks were both ~18 % slower (a larger frontier for the
// top level, or longer serial chains per block).
constexpr int kLbvhThreads = 512;
constexpr int kLbvhLeaves = 512;

template <bool K64>
__global__ __launch_bounds__(kLbvhThreads) void k_lbvh_local(ClimbArgs a) {
    constexpr int T = kLbvhLeaves;
    // Slots: internal node s -> s - base, leaf p -> T + p - base.  Every link
    // kept in LDS stays inside the block -- range ends, child slots, parent
    // splits, frontier slots -- so it is a 16-bit offset from base; deltas
    // (-1..159, stored plus one) and states fit 8 bits; leaf costs are
    // recomputed from the staged box: 38 KB per block, four blocks per CU.
    __shared__ float s_box[6][2 * T];
    __shared__ float s_cost[T];         // internal slots
    __shared__ int16_t s_par[2 * T];    // parent split - base, -1 if outside [base, last_leaf)
    __shared__ int s_cnt[T];
    __shared__ int16_t s_first[T], s_last[T], s_left[T], s_right[T];  // offsets; children as slots
    __shared__ int8_t s_state[T];
    __shared__ uint8_t s_delta[T + 1];  // delta(base - 1 + k) + 1
    __shared__ int16_t s_front[T];
    __shared__ int s_end[T];
    __shared__ int s_nfront, s_fbase;
    const int n = (int)a.n;
    const int base = blockIdx.x * T;
    const int last_leaf = min(base + T, n) - 1;
    if (threadIdx.x == 0) s_nfront = 0;
    for (int i = threadIdx.x; i < T; i += kLbvhThreads) {
        s_cnt[i] = 0;
        s_end[i] = -1;
        s_left[i] = -1;
        s_right[i] = -1;
        s_par[i] = -1;
        s_par[T + i] = -1;
    }
    for (int k = threadIdx.x; k <= T; k += kLbvhThreads) {
        const int q = base - 1 + k;
        s_delta[k] = (uint8_t)((q < 0 || q >= n - 1) ? 0 : 1 + kdelta<K64>(q, q + 1, n, a.keys, a.mesh_of, a.mesh_off));
    }
#pragma unroll
    for (int k = 0; k < T / kLbvhThreads; ++k) {
        const int p = base + k * kLbvhThreads + threadIdx.x;
        if (p > last_leaf) break;
        float unused;
        const Box3 tb = climb_leaf<false>(a, p, unused);
        const int sl = T + p - base;
        s_box[0][sl] = tb.mn.x; s_box[1][sl] = tb.mn.y; s_box[2][sl] = tb.mn.z;
        s_box[3][sl] = tb.mx.x; s_box[4][sl] = tb.mx.y; s_box[5][sl] = tb.mx.z;
    }
    __syncthreads();

    auto node_of = [&](int sl) { return sl < T ? base + sl : n - 1 + base + (sl - T); };
    auto box_mn = [&](int sl) { return make_float3(s_box[0][sl], s_box[1][sl], s_box[2][sl]); };
    auto box_mx = [&](int sl) { return make_float3(s_box[3][sl], s_box[4][sl], s_box[5][sl]); };
    auto cost_at = [&](int sl) { return sl < T ? s_cost[sl] : leaf_sah_cost(a, box_mn(sl), box_mx(sl)); };
#pragma unroll
    for (int k = 0; k < T / kLbvhThreads; ++k) {
        const int p = base + k * kLbvhThreads + threadIdx.x;
        if (p > last_leaf) break;
        int ns = T + p - base, l = p, r = p;
        while (true) {
            if (l == 0 && r == n - 1) {
                *a.root_id = node_of(ns);
                break;
            }
            const bool as_left = s_delta[r - base + 1] > s_delta[l - base];
            const int sp = as_left ? r : l - 1;
            if (sp < base || sp >= last_leaf) {
                s_front[atomicAdd(&s_nfront, 1)] = (int16_t)ns;
                break;
            }
            const int li = sp - base;
            s_par[ns] = (int16_t)li;
            if (as_left) {
                s_left[li] = (int16_t)ns;
                s_first[li] = (int16_t)(l - base);
            } else {
                s_right[li] = (int16_t)ns;
                s_last[li] = (int16_t)(r - base);
            }
            if (__hip_atomic_fetch_add(&s_cnt[li], 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) break;
            int ls, rs;
            if (as_left) {
                r = base + s_last[li];
                ls = ns;
                rs = s_right[li];
            } else {
                l = base + s_first[li];
                ls = s_left[li];
                rs = ns;
            }
            const Combined c = combine(a, l, r, box_mn(ls), box_mx(ls), cost_at(ls), box_mn(rs), box_mx(rs), cost_at(rs));
            s_box[0][li] = c.mn.x; s_box[1][li] = c.mn.y; s_box[2][li] = c.mn.z;
            s_box[3][li] = c.mx.x; s_box[4][li] = c.mx.y; s_box[5][li] = c.mx.z;
            s_cost[li] = c.cost;
            s_state[li] = (int8_t)c.state;
            if (c.state == 0 && !seg_top(a.mesh_of, a.mesh_off, n, l, r)) atomicMax(&s_end[l - base], r);
            ns = li;
        }
    }
    __syncthreads();
    // In-block parents with one arrival: the other child lies (partly)
    // outside the block, so the arrived child continues at the top level.
    for (int i = threadIdx.x; i < T; i += kLbvhThreads)
        if (s_cnt[i] == 1) s_front[atomicAdd(&s_nfront, 1)] = s_left[i] >= 0 ? s_left[i] : s_right[i];
    __syncthreads();
    if (threadIdx.x == 0) s_fbase = atomicAdd(a.frontier_count, s_nfront);
    __syncthreads();
    for (int i = threadIdx.x; i < s_nfront; i += kLbvhThreads) a.frontier[s_fbase + i] = node_of(s_front[i]);
    // Boxes and costs go out only where the top level reads them (nodes whose
    // parent it completes); the emit reads the Bin rows each state-1 or root
    // node gets from its children's boxes here.  Parents outside the block are
    // written -1 here and set by the top level, which climbs from every
    // frontier node.
    auto store_box = [&](int node, int sl) {
        const int po = s_par[sl];
        const bool parent_here = po >= 0 && s_cnt[po] == 2;
        if (a.all_boxes || !parent_here) {
            a.nmin[node] = make_float4(s_box[0][sl], s_box[1][sl], s_box[2][sl], 0.0f);
            a.nmax[node] = make_float4(s_box[3][sl], s_box[4][sl], s_box[5][sl], 0.0f);
            if (a.adaptive) a.cost[node] = cost_at(sl);
        }
    };
    auto parent_of = [&](int sl) { return s_par[sl] >= 0 ? base + s_par[sl] : -1; };
    auto box_at = [&](int sl, float4& mn, float4& mx) {
        mn = make_float4(s_box[0][sl], s_box[1][sl], s_box[2][sl], 0.0f);
        mx = make_float4(s_box[3][sl], s_box[4][sl], s_box[5][sl], 0.0f);
    };
    for (int i = threadIdx.x; i < T && base + i <= last_leaf; i += kLbvhThreads) {
        a.end_at[base + i] = s_end[i];
        // Every leaf's parent: the refit climb (k_climb_local) starts from it.
        a.parent[n - 1 + base + i] = parent_of(T + i);
        store_box(n - 1 + base + i, T + i);
    }
    // Internal nodes of the block's split range: the finished ones written
    // out coalesced; the others are completed by the top level, whose
    // arrival words start cleared.
    for (int i = threadIdx.x; i < T; i += kLbvhThreads) {
        const int node = base + i;
        if (node >= n - 1) break;
        if (s_cnt[i] != 2) {
            a.link[node] = 0;
            continue;
        }
        const int first = base + s_first[i], last = base + s_last[i];
        a.first[node] = first;
        a.last[node] = last;
        a.left[node] = node_of(s_left[i]);
        a.right[node] = node_of(s_right[i]);
        a.parent[node] = parent_of(i);
        store_box(node, i);
        a.state[node] = s_state[i];
        if (s_state[i] == 1 || seg_top(a.mesh_of, a.mesh_off, n, first, last)) {
            float4 amn, amx, bmn, bmx;
            box_at(s_left[i], amn, amx);
            box_at(s_right[i], bmn, bmx);
            int4 r0, r1, r2;
            bin_rows(amn, amx, bmn, bmx, r0, r1, r2);
            int4* pr = reinterpret_cast<int4*>(a.proto + 3 * (int64_t)node);
            pr[0] = r0;
            pr[1] = r1;
            pr[2] = r2;
        }
    }
}

// Each arrival at a parent ORs one 64-bit word: presence (bit 63), its node
// id (bits 31-61) and its far range end (bits 0-30: first of the parent if
// it is the left child, last if the right), so the second arrival learns
// the sibling and the parent's range from the atomic's result.  A node's box
// travels in registers up the chain and through nmin/nmax (sc1 stores that
// complete before the arrival) to the sibling.
template <bool K64>
__global__ __launch_bounds__(kBlock) void k_lbvh_top(ClimbArgs a) {
    const int count = *a.frontier_count;
    const int n = (int)a.n;
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < count; e += gridDim.x * kBlock) {
        int node = a.frontier[e];
        int l, r;
        if (node >= n - 1) {
            l = r = node - (n - 1);
        } else {
            l = a.first[node];
            r = a.last[node];
        }
        int dl = l == 0 ? -1 : kdelta<K64>(l - 1, l, n, a.keys, a.mesh_of, a.mesh_off);
        int dr = r == n - 1 ? -1 : kdelta<K64>(r, r + 1, n, a.keys, a.mesh_of, a.mesh_off);
        float3 mn = xyz(a.nmin[node]), mx = xyz(a.nmax[node]);
        float cost = a.adaptive ? a.cost[node] : 0.0f;
        while (true) {
            if (l == 0 && r == n - 1) {
                a.parent[node] = -1;
                *a.root_id = node;
                break;
            }
            const bool as_left = dr > dl;
            const int sp = as_left ? r : l - 1;
            complete_stores();
            const uint64_t mine = (1ull << 63) | ((uint64_t)(uint32_t)node << 31) | (uint32_t)(as_left ? l : r);
            const uint64_t old = __hip_atomic_fetch_or(&a.link[sp], mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            a.parent[node] = sp;
            if (as_left) {
                a.left[sp] = node;
                a.first[sp] = l;
            } else {
                a.right[sp] = node;
                a.last[sp] = r;
            }
            if (old == 0) break;
            const int sib = (int)((old >> 31) & 0x7fffffffu);
            const int far = (int)(old & 0x7fffffffu);
            const float4 smn = ld_agent(a.nmin + sib), smx = ld_agent(a.nmax + sib);
            const float scost = a.adaptive ? ld_agent(a.cost + sib) : 0.0f;
            Combined c;
            if (as_left) {
                r = far;
                dr = r == n - 1 ? -1 : kdelta<K64>(r, r + 1, n, a.keys, a.mesh_of, a.mesh_off);
                c = combine(a, l, r, mn, mx, cost, xyz(smn), xyz(smx), scost);
            } else {
                l = far;
                dl = l == 0 ? -1 : kdelta<K64>(l - 1, l, n, a.keys, a.mesh_of, a.mesh_off);
                c = combine(a, l, r, xyz(smn), xyz(smx), scost, mn, mx, cost);
            }
            st_agent(a.nmin + sp, make_float4(c.mn.x, c.mn.y, c.mn.z, 0.0f));
            st_agent(a.nmax + sp, make_float4(c.mx.x, c.mx.y, c.mx.z, 0.0f));
            if (a.adaptive) st_agent(a.cost + sp, c.cost);
            a.state[sp] = c.state;
            const bool top = seg_top(a.mesh_of, a.mesh_off, n, l, r);
            if (c.state == 0 && !top) atomicMax(&a.end_at[l], r);
            if (c.state == 1 || top) {
                const float4 omn = make_float4(mn.x, mn.y, mn.z, 0.0f), omx = make_float4(mx.x, mx.y, mx.z, 0.0f);
                int4 r0, r1, r2;
                if (as_left)
                    bin_rows(omn, omx, smn, smx, r0, r1, r2);
                else
                    bin_rows(smn, smx, omn, omx, r0, r1, r2);
                int4* pr = reinterpret_cast<int4*>(a.proto + 3 * (int64_t)sp);
                pr[0] = r0;
                pr[1] = r1;
                pr[2] = r2;
            }
            mn = c.mn;
            mx = c.mx;
            cost = c.cost;
            node = sp;
        }
    }
}

// ---------------------------------------------------------------------------
// Deterministic collapse.
//
// The reference marks the topmost SAH-leaf below the root per reference
// (computeLeafIndices), invalidates everything under it, lays out nodes by a
// BFS with warp atomics and reorders triangles with atomics.  Here the
// collapsed leaves are found with scans instead: the ranges of state-0 nodes
// form a laminar family, so with cover[p] = max over a <= p of end_at[a]
//   - a final leaf starts at p iff p == 0 or cover[p-1] < p;
//   - internal node i is a final interior iff i is the root or
//     (state[i] == 1 and cover[first_i] < last_i).
// Interior slots number the final interiors in slot-position order (the node
// id in Karras and split-numbered trees, the split position in general
// trees), each mesh's root rotated to its first slot; leaf slots follow in
// start order, and a leaf's triangles are the contiguous sorted range
// [first, last], so the triangle index array is the sorted triangle ids
// themselves.
// ---------------------------------------------------------------------------

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

struct MaxOp {
    HX_DEV int operator()(int a, int b) const { return a > b ? a : b; }
};
struct Sum2Op {
    HX_DEV int2 operator()(int2 a, int2 b) const { return make_int2(a.x + b.x, a.y + b.y); }
};

struct FlagCtx {
    int n;
    const int32_t* cover;
    const int32_t* first;
    const int32_t* last;
    const int32_t* state;
    const int32_t* mesh_of;  // batched builds only
    const int32_t* off;
    const int32_t* node_at;  // general trees: node at each split position (else NULL)
};

// The internal node whose slot position is p (the node at split position p
// in a general tree; node p itself in Karras and split-numbered trees).
HX_DEV int node_of_pos(const FlagCtx& f, int p) { return f.node_at ? f.node_at[p] : p; }

HX_DEV int leaf_start(const FlagCtx& f, int p) {
    return (p < f.n) && (p == 0 || f.cover[p - 1] < p);
}

HX_DEV int interior_flag(const FlagCtx& f, int i) {
    if (i >= f.n - 1) return 0;
    const int fi = f.first[i], la = f.last[i];
    if (!f.mesh_of) {
        if (fi == 0 && la == f.n - 1) return 1;  // root
    } else {
        const int m = f.mesh_of[fi];
        if (f.mesh_of[la] != m) return 0;                       // separator
        if (fi == f.off[m] && la == f.off[m + 1] - 1) return 1;  // mesh root
    }
    if (f.state[i] != 1) return 0;
    return f.cover[fi] < la;
}

// Single-pass versions of the two scans (decoupled look-back over tiles;
// each workgroup takes the next tile number from a counter, so every tile it
// waits on has started and publishes without waiting).  Status words: flag
// in the top two bits (aggregate or inclusive prefix), value below.
constexpr uint64_t kLb2Aggregate = 1ull << 62;
constexpr uint64_t kLb2Prefix = 2ull << 62;
constexpr uint64_t kLb2Half = (1ull << 31) - 1ull;

// Look-back by one whole wave: lane k reads the status of tile j - k, so 64
// predecessors arrive per round trip; the lanes up to the nearest published
// prefix are combined, or, if a tile among them has published nothing yet,
// the lanes before it, and the walk resumes there.  Tiles before 0 read as
// an identity prefix.  P: status word type, decoding and combine.
template <class P>
HX_DEV typename P::V lookback_wave(const typename P::W* status, int tile) {
    using V = typename P::V;
    using W = typename P::W;
    const int lane = lane_id();
    V acc = P::identity();
    int j = tile - 1;
    while (true) {
        const int q = j - lane;
        const W w = q >= 0 ? P::load(status + q) : P::kPrefixIdentity;
        const int flag = P::flag(w);
        const uint64_t pre = ballot64(flag == 2);
        const uint64_t none = ballot64(flag == 0);
        const int stop = pre ? __ffsll((unsigned long long)pre) - 1 : kWave;
        const uint64_t upto = stop < kWave ? (2ull << stop) - 1ull : ~0ull;
        const uint64_t waiting = none & upto;
        const int take = waiting ? __ffsll((unsigned long long)waiting) - 1 : (stop < kWave ? stop + 1 : kWave);
        acc = P::combine(acc, P::wave_reduce(lane < take ? P::value(w) : P::identity()));
        if (!waiting && stop < kWave) return acc;
        j -= take;
        if (waiting) __builtin_amdgcn_s_sleep(1);
    }
}

struct SlotsLb {  // (interior count, leaf count), 31 bits each
    using V = int2;
    using W = uint64_t;
    static constexpr W kPrefixIdentity = kLb2Prefix;
    HX_DEV static W load(const W* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    HX_DEV static int flag(W w) { return (int)(w >> 62); }
    HX_DEV static V value(W w) { return make_int2((int)((w >> 31) & kLb2Half), (int)(w & kLb2Half)); }
    HX_DEV static V identity() { return make_int2(0, 0); }
    HX_DEV static V combine(V a, V b) { return make_int2(a.x + b.x, a.y + b.y); }
    HX_DEV static V wave_reduce(V v) { return make_int2(wave_sum_i32(v.x), wave_sum_i32(v.y)); }
};

// cover[p] = max over a <= p of end_at[a], in place: tile maxima, their
// exclusive scan, then the tile scans.
__global__ __launch_bounds__(kBlock) void k_tile_max(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ partial) {
    const int64_t base = blockIdx.x * (int64_t)kScanTile;
    int m = -1;
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t p = base + k * kBlock + threadIdx.x;
        if (p < n) m = max(m, in[p]);
    }
#pragma unroll
    for (int s = 1; s < kWave; s <<= 1) m = max(m, __shfl_xor(m, s, kWave));
    __shared__ int wm[kWavesPerBlock];
    if (lane_id() == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int r = wm[0];
        for (int i = 1; i < kWavesPerBlock; ++i) r = max(r, wm[i]);
        partial[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_partials_max(int32_t* __restrict__ partial, int64_t ntiles) {
    __shared__ int wt[kWavesPerBlock];
    __shared__ int incs[kBlock];
    int carry = -1;
    for (int64_t base = 0; base < ntiles; base += kBlock) {
        const int64_t t = base + threadIdx.x;
        const int v = t < ntiles ? partial[t] : -1;
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(v, MaxOp(), wt, tot);
        incs[threadIdx.x] = inc;
        __syncthreads();
        const int e = threadIdx.x == 0 ? carry : max(carry, incs[threadIdx.x - 1]);
        if (t < ntiles) partial[t] = e;
        carry = max(carry, tot);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_tile_cover(int32_t* __restrict__ cover, int64_t n,
                                                       const int32_t* __restrict__ partial) {
    __shared__ int wt[kWavesPerBlock];
    const int64_t base = blockIdx.x * (int64_t)kScanTile;
    int carry = partial[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t p = base + k * kBlock + threadIdx.x;
        const int v = p < n ? cover[p] : -1;
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(v, MaxOp(), wt, tot);
        if (p < n) cover[p] = max(carry, inc);
        carry = max(carry, tot);
    }
}

// Interior slots and leaf numbers in one pass: int_slot[node] is the number
// of final interiors at earlier positions, complemented (~) when the node is
// not a final interior; leaf_num[p] the number of leaves starting before p.
// The last tile writes the totals to counts[0..1].  Both flags are bits, so
// each wave counts an item's flags with two ballots; the kScanItems x
// kWavesPerBlock wave counts (one per wave of the block, in position order)
// are scanned by the first wave, which also runs the look-back.
static_assert(kScanItems * kWavesPerBlock == kWave, "one wave count per lane of the first wave");
__global__ __launch_bounds__(kBlock) void k_slots_lb(FlagCtx f, int32_t* __restrict__ int_slot,
                                                     int32_t* __restrict__ leaf_num, int32_t* __restrict__ counts,
                                                     uint64_t* __restrict__ status, uint32_t* __restrict__ counter) {
    __shared__ int2 s_pre[kScanItems * kWavesPerBlock];  // wave counts, then their exclusive prefixes
    __shared__ int s_tile;
    __shared__ int2 s_excl;
    if (threadIdx.x == 0) s_tile = (int)atomicAdd(counter, 1u);
    __syncthreads();
    const int tile = s_tile;
    const int64_t base = tile * (int64_t)kScanTile;
    const int ntiles = (f.n + kScanTile - 1) / kScanTile;
    const int w = threadIdx.x / kWave;
    const int lane = lane_id();
    uint32_t bits = 0;  // bit 2k: interior flag, bit 2k+1: leaf start of item k
    if (!f.mesh_of && !f.node_at) {
        // One tree, node i at split position i: interior_flag and leaf_start
        // with every item's loads issued before any is used (two dependent
        // round trips per thread instead of three per item).
        int fi[kScanItems], la[kScanItems], st[kScanItems], cp[kScanItems], cf[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t p = base + k * kBlock + threadIdx.x;
            const bool in = p < f.n - 1;
            fi[k] = in ? f.first[p] : 0;
            la[k] = in ? f.last[p] : 0;
            st[k] = in ? f.state[p] : 0;
            cp[k] = p > 0 && p < f.n ? f.cover[p - 1] : 0;
        }
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) cf[k] = st[k] == 1 ? f.cover[fi[k]] : 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t p = base + k * kBlock + threadIdx.x;
            if (p < f.n) {
                const bool root = fi[k] == 0 && la[k] == f.n - 1;
                const int fl = p < f.n - 1 && (root || (st[k] == 1 && cf[k] < la[k]));
                const int ls = p == 0 || cp[k] < p;
                bits |= (uint32_t)fl << (2 * k) | (uint32_t)ls << (2 * k + 1);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            const int64_t p = base + k * kBlock + threadIdx.x;
            if (p < f.n) {
                const int fi = p < f.n - 1 ? interior_flag(f, node_of_pos(f, (int)p)) : 0;
                const int ls = leaf_start(f, (int)p);
                bits |= (uint32_t)fi << (2 * k) | (uint32_t)ls << (2 * k + 1);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t mi = ballot64((bits >> (2 * k)) & 1u), ml = ballot64((bits >> (2 * k + 1)) & 1u);
        if (lane == 0) s_pre[k * kWavesPerBlock + w] = make_int2(__popcll(mi), __popcll(ml));
    }
    __syncthreads();
    if (w == 0) {
        const int2 c = s_pre[lane];
        const int2 inc = wave_inclusive_scan(c, Sum2Op());
        s_pre[lane] = make_int2(inc.x - c.x, inc.y - c.y);
        const int2 agg = make_int2(__builtin_amdgcn_readlane(inc.x, kWave - 1), __builtin_amdgcn_readlane(inc.y, kWave - 1));
        const uint64_t aw = ((uint64_t)agg.x << 31) | (uint64_t)agg.y;
        if (lane == 0)
            __hip_atomic_store(status + tile, (tile == 0 ? kLb2Prefix : kLb2Aggregate) | aw, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const int2 excl = tile == 0 ? make_int2(0, 0) : lookback_wave<SlotsLb>(status, tile);
        if (lane == 0) {
            const uint64_t iw = ((uint64_t)(excl.x + agg.x) << 31) | (uint64_t)(excl.y + agg.y);
            if (tile != 0) __hip_atomic_store(status + tile, kLb2Prefix | iw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tile == ntiles - 1) {
                counts[0] = excl.x + agg.x;
                counts[1] = excl.y + agg.y;
            }
            s_excl = excl;
        }
    }
    __syncthreads();
    const int2 excl = s_excl;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t p = base + k * kBlock + threadIdx.x;
        const uint32_t vi = (bits >> (2 * k)) & 1u, vl = (bits >> (2 * k + 1)) & 1u;
        const uint64_t mi = ballot64(vi), ml = ballot64(vl);
        const int2 pre = s_pre[k * kWavesPerBlock + w];
        if (p < f.n) {
            const int slot = excl.x + pre.x + count_below(mi);
            if (p < f.n - 1) int_slot[node_of_pos(f, (int)p)] = vi ? slot : ~slot;
            leaf_num[p] = excl.y + pre.y + count_below(ml);
        }
    }
}

HX_DEV void write_node(int4* nodes, int slot, float4 amn, float4 amx, float4 bmn, float4 bmx, int c0, int c1, int size,
                       int parent) {
    int4* o = nodes + 4 * (int64_t)slot;
    o[0] = make_int4(__float_as_int(amn.x), __float_as_int(amx.x), __float_as_int(amn.y), __float_as_int(amx.y));
    o[1] = make_int4(__float_as_int(bmn.x), __float_as_int(bmx.x), __float_as_int(bmn.y), __float_as_int(bmx.y));
    o[2] = make_int4(__float_as_int(amn.z), __float_as_int(amx.z), __float_as_int(bmn.z), __float_as_int(bmx.z));
    o[3] = make_int4(c0, c1, size, parent);
}

// Per-mesh layout (one mesh unless batched): mesh m's nodes start at
// tab[m].x; its interiors keep the global interior order with the mesh root
// rotated to local slot 0; its leaves follow (tab[m].w interiors) in start
// order; triangle ranges and child/parent links are mesh-local.
struct MeshSlots {
    int4 tab;   // node base, first interior slot, first leaf number, interiors
    int root;   // interior slot of the mesh root
    int toff;   // first triangle position of the mesh
    HX_DEV int interior(int s) const { return s == root ? 0 : s - tab.y + (s < root ? 1 : 0); }
    HX_DEV int leaf(int num) const { return tab.w + num - tab.z; }
};

HX_DEV MeshSlots mesh_slots(const EmitArgs& a, const FlagCtx& f, int first) {
    const int m = f.mesh_of ? f.mesh_of[first] : 0;
    MeshSlots ms;
    ms.tab = a.mesh_tab[m];
    ms.root = a.root_slot[m];
    ms.toff = f.mesh_of ? f.off[m] : 0;
    return ms;
}

// PROTO: the rows come from the climb's per-node record (one coalesced read)
// instead of the children's boxes.
template <bool PROTO>
__global__ __launch_bounds__(kBlock) void k_emit(EmitArgs a, FlagCtx f) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int n = (int)a.n;
    if (i >= n - 1) return;
    const int s = a.int_slot[i];
    if (s < 0) return;  // not a final interior
    const int first = a.first[i], last = a.last[i];
    const MeshSlots ms = mesh_slots(a, f, first);
    const int ls = ms.interior(s);
    const int ch[2] = {a.left[i], a.right[i]};
    int4 r0, r1, r2;
    float4 cmn[2], cmx[2];
    if (PROTO) {
        const int4* pr = reinterpret_cast<const int4*>(a.proto + 3 * (int64_t)i);
        r0 = pr[0];
        r1 = pr[1];
        r2 = pr[2];
        cmn[0] = make_float4(__int_as_float(r0.x), __int_as_float(r0.z), __int_as_float(r2.x), 0.0f);
        cmx[0] = make_float4(__int_as_float(r0.y), __int_as_float(r0.w), __int_as_float(r2.y), 0.0f);
        cmn[1] = make_float4(__int_as_float(r1.x), __int_as_float(r1.z), __int_as_float(r2.z), 0.0f);
        cmx[1] = make_float4(__int_as_float(r1.y), __int_as_float(r1.w), __int_as_float(r2.w), 0.0f);
    } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            cmn[k] = a.nmin[ch[k]];
            cmx[k] = a.nmax[ch[k]];
        }
        bin_rows(cmn[0], cmx[0], cmn[1], cmx[1], r0, r1, r2);
    }
    int cidx[2];
    // With PROTO the tree is split-numbered: node i's children cover
    // [first, i] and [i+1, last], and a child is a final interior exactly when
    // its int_slot is non-negative, so no child range or state is gathered.
    int cslot[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) cslot[k] = PROTO && ch[k] < n - 1 ? a.int_slot[ch[k]] : -1;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = ch[k];
        if (PROTO ? cslot[k] >= 0 : (c < n - 1 && a.state[c] == 1)) {
            cidx[k] = ms.interior(PROTO ? cslot[k] : a.int_slot[c]);
        } else {
            int cf, cl;
            if (PROTO) {
                cf = k == 0 ? first : i + 1;
                cl = k == 0 ? i : last;
            } else {
                cf = c >= n - 1 ? c - (n - 1) : a.first[c];
                cl = c >= n - 1 ? cf : a.last[c];
            }
            const int L = ms.leaf(a.leaf_num[cf]);
            cidx[k] = ~L;
            write_node(a.nodes, ms.tab.x + L, cmn[k], cmx[k], cmn[k], cmx[k], cf - ms.toff, cl + 1 - ms.toff,
                       ~(cl - cf + 1), ls);
            if (a.bin_of) a.bin_of[c] = -2 - (ms.tab.x + L);
        }
    }
    const int par = (s == ms.root) ? -1 : ms.interior(a.int_slot[a.parent[i]]);
    int4* o = a.nodes + 4 * (int64_t)(ms.tab.x + ls);
    o[0] = r0;
    o[1] = r1;
    o[2] = r2;
    o[3] = make_int4(cidx[0], cidx[1], last - first + 1, par);
    if (a.bin_of) a.bin_of[i] = ms.tab.x + ls;
}



// Mesh tables from the slot scans: each mesh root's interior slot (batched
// builds), then per mesh the node base, first interior slot, first leaf
// number and interior count.
__global__ __launch_bounds__(kBlock) void k_mesh_roots(EmitArgs a, FlagCtx f) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= f.n - 1) return;
    const int fi = f.first[i], la = f.last[i];
    const int m = f.mesh_of[fi];
    if (f.mesh_of[la] == m && fi == f.off[m] && la == f.off[m + 1] - 1) a.root_slot[m] = a.int_slot[i];
}

__global__ __launch_bounds__(kBlock) void k_mesh_tab(EmitArgs a, FlagCtx f) {
    const int m = blockIdx.x * kBlock + threadIdx.x;
    if (m >= a.n_meshes) return;
    // Interior slots before position p (position n-1 holds no internal node).
    auto slots_before = [&](int p) {
        if (p >= f.n - 1) return a.counts[0];
        const int v = a.int_slot[node_of_pos(f, p)];
        return v >= 0 ? v : ~v;
    };
    int ib = 0, lb = 0, ie = a.counts[0];
    if (!f.mesh_of) a.root_slot[0] = a.int_slot[a.root_id ? *a.root_id : 0];
    if (f.mesh_of) {
        ib = slots_before(f.off[m]);
        lb = a.leaf_num[f.off[m]];
        if (m + 1 < a.n_meshes) ie = slots_before(f.off[m + 1]);
    }
    a.mesh_tab[m] = make_int4(ib + lb, ib, lb, ie - ib);
}

// n == 1: interior root whose two children are the single leaf (the
// reference indexes termCounters[-1] here, BVHCollapserKernels.cu:29-32).
__global__ void k_single_ref(const float* __restrict__ verts, const int32_t* __restrict__ tris,
                             const int32_t* __restrict__ sorted_tri, float4* nmin, float4* nmax, float4* woop,
                             int4* nodes, int32_t* counts) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float3 v0, v1, v2;
    tri_vertices(verts, tris, sorted_tri[0], v0, v1, v2);
    const Box3 tb = tri_box(v0, v1, v2);
    const float4 mn = make_float4(tb.mn.x, tb.mn.y, tb.mn.z, 0.0f);
    const float4 mx = make_float4(tb.mx.x, tb.mx.y, tb.mx.z, 0.0f);
    nmin[0] = mn;
    nmax[0] = mx;
    float4 w0, w1, w2;
    woop_rows(v0, v1, v2, w0, w1, w2);
    woop[0] = w0;
    woop[1] = w1;
    woop[2] = w2;
    write_node(nodes, 0, mn, mx, mn, mx, ~1, ~1, 1, -1);
    write_node(nodes, 1, mn, mx, mn, mx, 0, 1, ~1, 0);
    counts[0] = 1;
    counts[1] = 1;
}

// ---------------------------------------------------------------------------
// SAH cost (HipBVHKernels.cu:184-238): FP32 per-node terms, FP64 sums in a
// fixed tree order (the reference adds with a float atomicAdd).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_cost_partial(const int4* __restrict__ nodes, const int32_t* __restrict__ counts,
                                                         const float* __restrict__ sb, float ct, float ci,
                                                         double* __restrict__ partial) {
    const int total = counts[0] + counts[1];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const float sceneArea = box_area(make_float3(sb[0], sb[1], sb[2]), make_float3(sb[3], sb[4], sb[5]));
    double term = 0.0;
    if (i < total) {
        const int4* nd = nodes + 4 * (int64_t)i;
        const int4 q0 = nd[0], q1 = nd[1], q2 = nd[2], q3 = nd[3];
        Box3 b = box_empty();
        box_grow(b, make_float3(__int_as_float(q0.x), __int_as_float(q0.z), __int_as_float(q2.x)));
        box_grow(b, make_float3(__int_as_float(q0.y), __int_as_float(q0.w), __int_as_float(q2.y)));
        box_grow(b, make_float3(__int_as_float(q1.x), __int_as_float(q1.z), __int_as_float(q2.z)));
        box_grow(b, make_float3(__int_as_float(q1.y), __int_as_float(q1.w), __int_as_float(q2.w)));
        const float P = box_area(b.mn, b.mx) / sceneArea;
        const float t = q3.z < 0 ? ci * P * (float)(~q3.z) : ct * P;
        term = (double)t;
    }
    __shared__ double s[kBlock];
    s[threadIdx.x] = term;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// One workgroup: strided per-thread sums, then a fixed-order tree.
__global__ __launch_bounds__(kBlock) void k_cost_final(const double* __restrict__ partial, int nblocks,
                                                      double* __restrict__ out) {
    __shared__ double s[kBlock];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += kBlock) acc += partial[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Leaf statistics (computeSumOfLeafSizes / computeLeafSizeHistogram,
// HipBVHKernels.cu:124-182): per leaf node its size (~word 14); out[0..32]
// counts sizes 0..32, out[33] sums all sizes.  Block counts in LDS, then one
// atomic per non-zero entry.
__global__ __launch_bounds__(kBlock) void k_leaf_stats(const int4* __restrict__ nodes,
                                                      const int32_t* __restrict__ counts,
                                                      int32_t* __restrict__ out) {
    __shared__ int s_hist[kMaxLeafSizeStat + 2];
    for (int k = threadIdx.x; k < kMaxLeafSizeStat + 2; k += kBlock) s_hist[k] = 0;
    __syncthreads();
    const int ni = counts[0], nl = counts[1];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < nl) {
        const int size = ~nodes[4 * (int64_t)(ni + i) + 3].z;
        if (size >= 0 && size <= kMaxLeafSizeStat) atomicAdd(&s_hist[size], 1);
        atomicAdd(&s_hist[kMaxLeafSizeStat + 1], size);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kMaxLeafSizeStat + 2; k += kBlock)
        if (s_hist[k]) atomicAdd(&out[k], s_hist[k]);
}

// HipBVH::validate (HipBVH.cpp:157-310) on the device, one thread per Bin
// node: an interior's children must name it as parent and their sizes sum to
// its size; a leaf's range must be non-empty and inside the references and
// its box must overlap each of its triangles' boxes.  Every child reference
// and every leaf triangle is counted (node_hits, tri_hits); k_validate_counts
// then requires one reference per node but the root and one or more per
// triangle.  Every index is range-checked before use.  Failures OR bits into
// *flags: 1 child index, 2 parent, 4 size, 8 leaf range, 16 leaf box,
// 32 triangle index, 64 node references, 128 triangle references.
__global__ __launch_bounds__(kBlock) void k_validate_nodes(const int4* __restrict__ nodes, const int32_t* __restrict__ counts,
                                                          const int32_t* __restrict__ tri_idx, int64_t n_refs,
                                                          const float* __restrict__ verts, const int32_t* __restrict__ tris,
                                                          int64_t n_tri, int32_t* __restrict__ node_hits,
                                                          int32_t* __restrict__ tri_hits, int32_t* __restrict__ flags) {
    const int ni = counts[0], nn = counts[0] + counts[1];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nn) return;
    const int4* nd = nodes + 4 * (int64_t)i;
    const int4 q3 = nd[3];
    int f = 0;
    if (i < ni) {
        int sum = 0;
        const int ch[2] = {q3.x, q3.y};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            // A single reference is one leaf named by both of the root's links.
            if (k == 1 && ch[0] == ch[1] && n_refs == 1) continue;
            const int c = ch[k] < 0 ? ~ch[k] : ch[k];
            if (c <= 0 || c >= nn) {
                f |= 1;
                continue;
            }
            const int4 cq = nodes[4 * (int64_t)c + 3];
            if (cq.w != i) f |= 2;
            sum += cq.z < 0 ? ~cq.z : cq.z;
            atomicAdd(&node_hits[c], 1);
        }
        if (sum != q3.z) f |= 4;
        if (i == 0 && q3.w != -1) f |= 2;
    } else {
        const int b = q3.x, e = q3.y;
        if (!(b < e) || b < 0 || e > n_refs || ~q3.z != e - b) {
            f |= 8;
        } else {
            const float4 r0 = reinterpret_cast<const float4*>(nd)[0], r2 = reinterpret_cast<const float4*>(nd)[2];
            const float mnx = r0.x, mxx = r0.y, mny = r0.z, mxy = r0.w, mnz = r2.x, mxz = r2.y;
            for (int p = b; p < e; ++p) {
                const int t = tri_idx[p];
                if (t < 0 || t >= n_tri) {
                    f |= 32;
                    continue;
                }
                atomicAdd(&tri_hits[t], 1);
                float3 v0, v1, v2;
                tri_vertices(verts, tris, t, v0, v1, v2);
                const Box3 tb = tri_box(v0, v1, v2);
                if (tb.mx.x < mnx || tb.mn.x > mxx || tb.mx.y < mny || tb.mn.y > mxy || tb.mx.z < mnz ||
                    tb.mn.z > mxz)
                    f |= 16;
            }
        }
    }
    if (f) atomicOr(flags, f);
}

__global__ __launch_bounds__(kBlock) void k_validate_counts(const int32_t* __restrict__ counts,
                                                           const int32_t* __restrict__ node_hits,
                                                           const int32_t* __restrict__ tri_hits, int64_t n_tri,
                                                           int32_t* __restrict__ flags) {
    const int nn = counts[0] + counts[1];
    const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    int f = 0;
    if (i < nn && node_hits[i] != (i == 0 ? 0 : 1)) f |= 64;
    if (i < n_tri && tri_hits[i] < 1) f |= 128;
    if (f) atomicOr(flags, f);
}

__global__ __launch_bounds__(kBlock) void k_node_at(const int32_t* __restrict__ split, int m,
                                                   int32_t* __restrict__ node_at) {
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v < m) node_at[split[v]] = v;
}

__global__ void k_refit_reset(int32_t* __restrict__ counters, int64_t n_internal, int32_t* __restrict__ frontier_count) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n_internal) counters[i] = 0;
    if (i == 0) *frontier_count = 0;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers.
// ---------------------------------------------------------------------------

int launch_fill_i32(Context* c, int32_t* p, int32_t v, int64_t n) {
    if (n <= 0) return HX_OK;
    LaunchScope ls(c, "fill_i32");
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, p, v, n);
    return check_launch("fill_i32");
}

int launch_scene_box(Context* c, const float* verts, const int32_t* tris, int64_t n, float* d_box6) {
    const int nblocks = (int)std::min<int64_t>(std::max<int64_t>((n + kBlock * 8 - 1) / (kBlock * 8), 1), 2048);
    HX_TRY(c->scan_partials.reserve(sizeof(float) * 6 * nblocks));
    {
        LaunchScope ls(c, "scene_box");
        hipLaunchKernelGGL(k_scene_box_partial, dim3(nblocks), dim3(kBlock), 0, c->stream, verts, tris, n,
                           c->scan_partials.as<float>());
        HX_TRY(check_launch("scene_box_partial"));
        hipLaunchKernelGGL(k_scene_box_final, dim3(1), dim3(kBlock), 0, c->stream, c->scan_partials.as<float>(), nblocks,
                           d_box6);
    }
    return check_launch("scene_box_final");
}

int launch_normalize(Context* c, float* verts, int64_t nv, float* d_box6) {
    HX_TRY(launch_scene_box(c, verts, nullptr, nv, d_box6));
    LaunchScope ls(c, "normalize");
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(nv)), dim3(kBlock), 0, c->stream, verts, nv, d_box6);
    return check_launch("normalize");
}

int launch_morton(Context* c, const float* verts, const int32_t* tris, int64_t n, const float* d_box6, int bits,
                  void* d_keys, bool keys64, const int32_t* mesh_of, const float4* ref_min, const float4* ref_max,
                  float4* pack_out) {
    LaunchScope ls(c, "morton");
    const int shift = 60 - bits;
    if (keys64)
        hipLaunchKernelGGL(k_morton<true>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, verts, tris, n, d_box6, shift,
                           d_keys, mesh_of, ref_min, ref_max, pack_out);
    else
        hipLaunchKernelGGL(k_morton<false>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, verts, tris, n, d_box6,
                           shift, d_keys, mesh_of, ref_min, ref_max, pack_out);
    return check_launch("morton");
}

int launch_construct(Context* c, const void* d_keys, bool keys64, int64_t n, int32_t* left, int32_t* right,
                     int32_t* parent, int32_t* first, int32_t* last, int32_t* counters, int32_t* end_at,
                     int32_t* zero_word, const int32_t* mesh_of, const int32_t* off) {
    LaunchScope ls(c, "construct");
    if (keys64)
        hipLaunchKernelGGL(k_construct<true>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, d_keys, (int)n, left, right,
                           parent, first, last, counters, end_at, zero_word, mesh_of, off);
    else
        hipLaunchKernelGGL(k_construct<false>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, d_keys, (int)n, left,
                           right, parent, first, last, counters, end_at, zero_word, mesh_of, off);
    return check_launch("construct");
}

int launch_climb(Context* c, const ClimbArgs& a) {
    LaunchScope ls(c, "climb");
    const unsigned blocks = (unsigned)((a.n + kClimbLeaves - 1) / kClimbLeaves);
    const bool atr = a.atr_cost != nullptr;
    if (a.split)
        hipLaunchKernelGGL((k_climb_local<true, false>), dim3(blocks), dim3(kClimbThreads), 0, c->stream, a,
                           a.frontier, a.frontier_count);
    else if (atr)
        hipLaunchKernelGGL((k_climb_local<false, true>), dim3(blocks), dim3(kClimbThreads), 0, c->stream, a,
                           a.frontier, a.frontier_count);
    else
        hipLaunchKernelGGL((k_climb_local<false, false>), dim3(blocks), dim3(kClimbThreads), 0, c->stream, a,
                           a.frontier, a.frontier_count);
    HX_TRY(check_launch("climb_local"));
    // One frontier chain per thread: the chains are latency-bound walks, so
    // the grid fills every wave slot (8 blocks of 256 per CU).
    const unsigned top = (unsigned)std::max(1, std::min<int>(c->cu_count * 8, (int)((a.n + kBlock - 1) / kBlock)));
    if (atr)
        hipLaunchKernelGGL(k_climb_top<true>, dim3(top), dim3(kBlock), 0, c->stream, a, a.frontier, a.frontier_count);
    else
        hipLaunchKernelGGL(k_climb_top<false>, dim3(top), dim3(kBlock), 0, c->stream, a, a.frontier,
                           a.frontier_count);
    return check_launch("climb_top");
}

int launch_lbvh_climb(Context* c, const ClimbArgs& a) {
    LaunchScope ls(c, "climb");
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, a.frontier_count, 0, (int64_t)1);
    HX_TRY(check_launch("frontier_reset"));
    const unsigned blocks = (unsigned)((a.n + kLbvhLeaves - 1) / kLbvhLeaves);
    if (a.keys64)
        hipLaunchKernelGGL(k_lbvh_local<true>, dim3(blocks), dim3(kLbvhThreads), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_lbvh_local<false>, dim3(blocks), dim3(kLbvhThreads), 0, c->stream, a);
    HX_TRY(check_launch("lbvh_local"));
    // Every wave slot, as for the general climb (measured 1.483 -> 1.455 ms
    // for the two climb kernels at 16.7M triangles against 4 blocks per CU).
    const unsigned top = (unsigned)std::max(1, std::min<int>(c->cu_count * 8, (int)((a.n + kBlock - 1) / kBlock)));
    if (a.keys64)
        hipLaunchKernelGGL(k_lbvh_top<true>, dim3(top), dim3(kBlock), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_lbvh_top<false>, dim3(top), dim3(kBlock), 0, c->stream, a);
    return check_launch("lbvh_top");
}

int launch_collapse_emit(Context* c, const EmitArgs& a) {
    const int64_t n = a.n;
    const int64_t ntiles = (n + kScanTile - 1) / kScanTile;
    HX_TRY(c->scan_partials.reserve(16 * (size_t)ntiles + 64));
    const bool general = a.split != nullptr;
    if (general) {
        hipLaunchKernelGGL(k_node_at, dim3(grid_for(n - 1)), dim3(kBlock), 0, c->stream, a.split, (int)(n - 1),
                           a.node_at);
        HX_TRY(check_launch("node_at"));
    }
    FlagCtx f{(int)n, a.cover, a.first, a.last, a.state, a.mesh_of, a.mesh_off, general ? a.node_at : nullptr};
    // The slots' look-back status words (8 B per tile) and tile counter,
    // after the cover's tile maxima.
    int32_t* pm = c->scan_partials.as<int32_t>();
    uint64_t* st_slots = reinterpret_cast<uint64_t*>(pm + ((ntiles + 1) & ~(int64_t)1));
    uint32_t* tile_ctr = reinterpret_cast<uint32_t*>(st_slots + ntiles);
    {
        LaunchScope ls(c, "collapse_cover");
        hipLaunchKernelGGL(k_tile_max, dim3((unsigned)ntiles), dim3(kBlock), 0, c->stream, a.cover, n, pm);
        HX_TRY(check_launch("tile_max"));
        hipLaunchKernelGGL(k_scan_partials_max, dim3(1), dim3(kBlock), 0, c->stream, pm, ntiles);
        HX_TRY(check_launch("scan_partials_max"));
        hipLaunchKernelGGL(k_tile_cover, dim3((unsigned)ntiles), dim3(kBlock), 0, c->stream, a.cover, n, pm);
        HX_TRY(check_launch("tile_cover"));
    }
    {
        LaunchScope ls(c, "collapse_slots");
        const int64_t words = 2 * ntiles + 1;
        hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(words)), dim3(kBlock), 0, c->stream,
                           reinterpret_cast<int32_t*>(st_slots), 0, words);
        HX_TRY(check_launch("slots_status"));
        hipLaunchKernelGGL(k_slots_lb, dim3((unsigned)ntiles), dim3(kBlock), 0, c->stream, f, a.int_slot, a.leaf_num,
                           a.counts, st_slots, tile_ctr);
        HX_TRY(check_launch("slots_lb"));
    }
    {
        LaunchScope ls(c, "collapse_emit");
        if (a.mesh_of) {
            hipLaunchKernelGGL(k_mesh_roots, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, f);
            HX_TRY(check_launch("mesh_roots"));
        }
        hipLaunchKernelGGL(k_mesh_tab, dim3(grid_for(a.n_meshes)), dim3(kBlock), 0, c->stream, a, f);
        HX_TRY(check_launch("mesh_tab"));
        if (a.proto)
            hipLaunchKernelGGL(k_emit<true>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, f);
        else
            hipLaunchKernelGGL(k_emit<false>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, f);
    }
    return check_launch("emit");
}

int launch_refit_reset(Context* c, int32_t* counters, int64_t n_internal, int32_t* frontier_count) {
    LaunchScope ls(c, "refit_reset");
    hipLaunchKernelGGL(k_refit_reset, dim3(grid_for(std::max<int64_t>(n_internal, 1))), dim3(kBlock), 0, c->stream,
                       counters, n_internal, frontier_count);
    return check_launch("refit_reset");
}

int launch_emit_only(Context* c, const EmitArgs& a) {
    FlagCtx f{(int)a.n, a.cover, a.first, a.last, a.state, a.mesh_of, a.mesh_off, a.split ? a.node_at : nullptr};
    LaunchScope ls(c, "collapse_emit");
    if (a.proto)
        hipLaunchKernelGGL(k_emit<true>, dim3(grid_for(a.n)), dim3(kBlock), 0, c->stream, a, f);
    else
        hipLaunchKernelGGL(k_emit<false>, dim3(grid_for(a.n)), dim3(kBlock), 0, c->stream, a, f);
    return check_launch("emit");
}

int launch_single_ref(Context* c, const float* verts, const int32_t* tris, const int32_t* sorted_tri, float4* nmin,
                      float4* nmax, float4* woop, int4* nodes, int32_t* counts) {
    LaunchScope ls(c, "single_ref");
    hipLaunchKernelGGL(k_single_ref, dim3(1), dim3(64), 0, c->stream, verts, tris, sorted_tri, nmin, nmax, woop, nodes,
                       counts);
    return check_launch("single_ref");
}

int launch_validate(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_nodes, const int32_t* tri_idx,
                    int64_t n_refs, const float* verts, const int32_t* tris, int64_t n_tri, int32_t* work,
                    int32_t* d_flags) {
    LaunchScope ls(c, "validate");
    int32_t* node_hits = work;
    int32_t* tri_hits = work + max_nodes;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(max_nodes + n_tri)), dim3(kBlock), 0, c->stream, work, 0,
                       max_nodes + n_tri);
    HX_TRY(check_launch("validate_clear"));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, d_flags, 0, (int64_t)1);
    HX_TRY(check_launch("validate_clear_flags"));
    hipLaunchKernelGGL(k_validate_nodes, dim3(grid_for(max_nodes)), dim3(kBlock), 0, c->stream, nodes, d_counts,
                       tri_idx, n_refs, verts, tris, n_tri, node_hits, tri_hits, d_flags);
    HX_TRY(check_launch("validate_nodes"));
    hipLaunchKernelGGL(k_validate_counts, dim3(grid_for(std::max(max_nodes, n_tri))), dim3(kBlock), 0, c->stream,
                       d_counts, node_hits, tri_hits, n_tri, d_flags);
    return check_launch("validate_counts");
}

int launch_leaf_stats(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_leaves, int32_t* d_out) {
    LaunchScope ls(c, "leaf_stats");
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, d_out, 0, (int64_t)(kMaxLeafSizeStat + 2));
    HX_TRY(check_launch("leaf_stats_clear"));
    hipLaunchKernelGGL(k_leaf_stats, dim3(grid_for(std::max<int64_t>(max_leaves, 1))), dim3(kBlock), 0, c->stream,
                       nodes, d_counts, d_out);
    return check_launch("leaf_stats");
}

int launch_cost(Context* c, const int4* nodes, const int32_t* d_counts, int64_t max_nodes, const float* d_box6,
                float ct, float ci, double* d_out) {
    const int nblocks = (int)grid_for(max_nodes);
    HX_TRY(c->scan_partials.reserve(sizeof(double) * nblocks));
    LaunchScope ls(c, "cost");
    hipLaunchKernelGGL(k_cost_partial, dim3(nblocks), dim3(kBlock), 0, c->stream, nodes, d_counts, d_box6, ct, ci,
                       c->scan_partials.as<double>());
    HX_TRY(check_launch("cost_partial"));
    hipLaunchKernelGGL(k_cost_final, dim3(1), dim3(kBlock), 0, c->stream, c->scan_partials.as<double>(), nblocks, d_out);
    return check_launch("cost_final");
}

}  // namespace hx
