// hx_trace.hip -- persistent while-while traversal of the Bin BVH for gfx950,
// plus the Morton ray keys used by the sorted trace.
//
// Reference (src/hippie/): rt/tracer/HipTracerBinKernels.cu:19-286 (trace),
// :572-842 (traceSort), rt/tracer/HipTracerKernels.h:29-199,
// rt/ray/RayBufferKernels.cu:12-56,492-614.
//
// Wave64 design: each wave fetches rays for its terminated lanes with one
// atomicAdd (64-bit ballot + mbcnt), traverses with the Aila-Laine
// while-while loop, postpones one leaf per lane, and refills when fewer than
// kDynamicFetch of its 64 lanes are still busy.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kTraceBlock = 256;
constexpr int kStackSize = 64;                 // HipTracerBinKernels.cu:5
constexpr int32_t kSentinel = 0x76543210;      // HipTracerKernels.h:11

// fminf/fmaxf (v_min_f32 / v_max_f32: a NaN operand yields the other, -0
// orders below +0, as the reference's fminf/fmaxf do) and the integer min/max
// on float bits of HipTracerKernels.h:144-195 (spanBeginKepler / spanEndKepler).
HX_DEV float fmin_s(float a, float b) { return fminf(a, b); }
HX_DEV float fmax_s(float a, float b) { return fmaxf(a, b); }
HX_DEV float span_begin(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    const int c = max(min(__float_as_int(c0), __float_as_int(c1)), __float_as_int(d));
    return __int_as_float(max(max(__float_as_int(fmin_s(a0, a1)), __float_as_int(fmin_s(b0, b1))), c));
}
HX_DEV float span_end(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    const int c = min(max(__float_as_int(c0), __float_as_int(c1)), __float_as_int(d));
    return __int_as_float(min(min(__float_as_int(fmax_s(a0, a1)), __float_as_int(fmax_s(b0, b1))), c));
}

// Bin node reads (k_trace).  A vector load of a node costs the CU's texture
// address and data units four 16-byte-per-lane instructions however many lanes
// share the node, and on coherent (sorted primary) rays those units, not the
// ALUs, set the pace.  When every active lane of the wave visits the same node
// -- the common case near the root for coherent rays -- the node is read with
// scalar loads instead (constant address space: the trace never writes the
// nodes), so no vector-memory instruction is issued for it.
HX_DEV int64_t first_lane_i64(int64_t v) {
    return (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
                     (uint32_t)__builtin_amdgcn_readfirstlane((int)v));
}

// Traversal stack: entries [0, LDS_DEPTH) in LDS, one column per lane (entry e
// of lane t at e * kTraceBlock + t, so a wave's accesses hit distinct banks);
// deeper entries in private (scratch) memory, touched only by deep traversals.
// The LDS column is typed as LDS (address space 3) so that the compiler keeps
// LDS and private accesses apart: through a generic pointer it merges the two
// into one flat load, which waits on both memory counters.
typedef __attribute__((address_space(3))) int32_t lds_i32;

template <int LDS_DEPTH, int SIZE = kStackSize>
struct TraversalStack {
    static constexpr int kDeep = SIZE - LDS_DEPTH > 0 ? SIZE - LDS_DEPTH : 1;
    lds_i32* col;
    int32_t deep[kDeep];
    HX_DEV void put(int e, int32_t v) {
        if (e < LDS_DEPTH) col[e * kTraceBlock] = v;
        else deep[e - LDS_DEPTH] = v;
    }
    HX_DEV int32_t get(int e) const {
        if (e < LDS_DEPTH) return col[e * kTraceBlock];
        return deep[e - LDS_DEPTH];
    }
};

// POSTPONE: leaves a lane may hold back while its wave still traverses (1, as
// HipTracerBinKernels.cu:169-177; 2 is the reference's commented variant,
// :171-173, 255; up to 4 here).  Held leaves wait in leafAddr (newest) ..
// leafAddr4; a lane whose slots are full stops until its wave has a leaf in
// every lane.  The closest hit does not depend on the order leaves are tested.
template <bool STATS, bool ORDERED, int LDS_DEPTH, int POSTPONE, int AHEAD>
__global__ __launch_bounds__(kTraceBlock) void k_trace(TraceArgs a) {
    __shared__ int32_t s_stack[LDS_DEPTH > 0 ? LDS_DEPTH * kTraceBlock : 1];
    TraversalStack<LDS_DEPTH> stack;
    stack.col = (lds_i32*)(s_stack + threadIdx.x);
    int sp = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
    float idx = 0.f, idy = 0.f, idz = 0.f, oodx = 0.f, oody = 0.f, oodz = 0.f;
    float hitT = 0.f, hitU = 0.f, hitV = 0.f;
    int32_t nodeAddr = kSentinel, leafAddr = 0, leafAddr2 = 0, leafAddr3 = 0, leafAddr4 = 0, hitIndex = -1,
            hitId = -1;
    int64_t rayidx = 0;
    int32_t nNodes = 0, nTris = 0;
    // Batched trace: this lane's mesh (node base, first triangle position).
    int64_t nodeBase = 0, triBase = 0;
    int32_t* rayCounter = a.control;

    while (true) {
        // Fetch new rays for the terminated lanes of this wave.
        const bool terminated = nodeAddr == kSentinel;
        const uint64_t tmask = ballot64(terminated);
        const int ntm = __popcll(tmask);
        const int below = count_below(tmask);
        const int leader = __ffsll((unsigned long long)tmask) - 1;
        int base = 0;
        if (terminated && below == 0) base = atomicAdd(rayCounter, ntm);
        base = __shfl(base, leader, kWave);
        if (terminated) {
            rayidx = (int64_t)base + below;
            if (rayidx >= a.n_rays) break;
            if (ORDERED) rayidx = a.order[rayidx];
            if (a.ray_mesh) {
                const int m = a.ray_mesh[rayidx];
                if ((unsigned)m >= (unsigned)a.n_meshes) {  // no such mesh: a miss
                    reinterpret_cast<int2*>(a.results)[2 * rayidx] =
                        make_int2(-1, __float_as_int(a.rays[2 * rayidx + 1].w));
                    if (STATS) a.stats[rayidx] = make_int2(0, 0);
                    continue;
                }
                nodeBase = a.mesh_tab[m].x;
                triBase = a.mesh_off[m];
            }
            const float4 o = a.rays[2 * rayidx + 0];
            const float4 d = a.rays[2 * rayidx + 1];
            ox = o.x; oy = o.y; oz = o.z; tmin = o.w;
            dx = d.x; dy = d.y; dz = d.z; hitT = d.w;
            const float ooeps = 8.2718061255302767e-25f;  // exp2f(-80)
            idx = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
            idy = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
            idz = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
            oodx = ox * idx;
            oody = oy * idy;
            oodz = oz * idz;
            stack.put(0, kSentinel);
            sp = 0;
            leafAddr = 0;
            leafAddr2 = 0;
            leafAddr3 = 0;
            leafAddr4 = 0;
            nodeAddr = 0;
            hitIndex = -1;
            hitId = -1;
            if (STATS) { nNodes = 0; nTris = 0; }
        }

        while (nodeAddr != kSentinel) {
            while ((uint32_t)nodeAddr < (uint32_t)kSentinel) {
                if (STATS) ++nNodes;
                // The box test is written out in both branches so that, for a
                // whole-wave node, it reads the scalar registers directly.
                auto box_test = [&](float4 n0xy, float4 n1xy, float4 nz, float& c0min, float& c0max, float& c1min,
                                    float& c1max) {
                    const float c0lox = fmaf(n0xy.x, idx, -oodx);
                    const float c0hix = fmaf(n0xy.y, idx, -oodx);
                    const float c0loy = fmaf(n0xy.z, idy, -oody);
                    const float c0hiy = fmaf(n0xy.w, idy, -oody);
                    const float c0loz = fmaf(nz.x, idz, -oodz);
                    const float c0hiz = fmaf(nz.y, idz, -oodz);
                    const float c1loz = fmaf(nz.z, idz, -oodz);
                    const float c1hiz = fmaf(nz.w, idz, -oodz);
                    c0min = span_begin(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, tmin);
                    c0max = span_end(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, hitT);
                    const float c1lox = fmaf(n1xy.x, idx, -oodx);
                    const float c1hix = fmaf(n1xy.y, idx, -oodx);
                    const float c1loy = fmaf(n1xy.z, idy, -oody);
                    const float c1hiy = fmaf(n1xy.w, idy, -oody);
                    c1min = span_begin(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, tmin);
                    c1max = span_end(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, hitT);
                };
                float c0min, c0max, c1min, c1max;
                int2 cn;
                {
                    const int64_t na = nodeBase + nodeAddr;
                    const int64_t n0 = first_lane_i64(na);
                    if (__all(na == n0)) {
                        typedef const __attribute__((address_space(4))) float const_f32;
                        typedef const __attribute__((address_space(4))) int32_t const_i32;
                        const_f32* f = (const_f32*)(a.nodes + 4 * n0);
                        const_i32* w = (const_i32*)(a.nodes + 4 * n0);
                        box_test(make_float4(f[0], f[1], f[2], f[3]), make_float4(f[4], f[5], f[6], f[7]),
                                 make_float4(f[8], f[9], f[10], f[11]), c0min, c0max, c1min, c1max);
                        cn = make_int2(w[12], w[13]);
                    } else {
                        const float4* np = reinterpret_cast<const float4*>(a.nodes + 4 * na);
                        const float4 n0xy = np[0];
                        const float4 n1xy = np[1];
                        const float4 nz = np[2];
                        const int4 links = a.nodes[4 * na + 3];
                        // The child links share the node's cache line: keep their
                        // load with the box loads instead of after the box test.
                        asm volatile("" ::"v"(links.x), "v"(links.y));
                        cn = make_int2(links.x, links.y);
                        box_test(n0xy, n1xy, nz, c0min, c0max, c1min, c1max);
                    }
                }
                const bool swp = c1min < c0min;
                const bool t0 = c0max >= c0min;
                const bool t1 = c1max >= c1min;
                int32_t cy = cn.y;
                if (!t0 && !t1) {
                    nodeAddr = stack.get(sp--);
                } else {
                    nodeAddr = t0 ? cn.x : cy;
                    if (t0 && t1) {
                        if (swp) {
                            const int32_t tmp = nodeAddr;
                            nodeAddr = cy;
                            cy = tmp;
                        }
                        if (sp + 1 >= kStackSize) {
                            atomicAdd(a.control + 1, 1);
                            nodeAddr = kSentinel;
                            leafAddr = leafAddr2 = leafAddr3 = leafAddr4 = 0;
                            break;
                        }
                        stack.put(++sp, cy);
                    }
                }
                const int32_t lastSlot = POSTPONE == 1 ? leafAddr : POSTPONE == 2 ? leafAddr2
                                         : POSTPONE == 3 ? leafAddr3 : leafAddr4;
                if (nodeAddr < 0 && lastSlot >= 0) {
                    if (POSTPONE >= 4) leafAddr4 = leafAddr3;
                    if (POSTPONE >= 3) leafAddr3 = leafAddr2;
                    if (POSTPONE >= 2) leafAddr2 = leafAddr;
                    leafAddr = nodeAddr;
                    nodeAddr = stack.get(sp--);
                }
                if (!__any(leafAddr >= 0)) break;
            }

            while (leafAddr < 0) {
                const int4 lb = a.nodes[4 * (nodeBase + ~leafAddr) + 3];
                // The next triangle's first Woop row (AHEAD = 1), or its first
                // two (AHEAD = 2), are read while this one is tested: leaves of
                // long thin triangles test hundreds per ray, and most of those
                // tests pass the plane test and need the second row.
                const float4* wrow = a.woop + 3 * (triBase + lb.x);
                float4 vnext = lb.x < lb.y ? wrow[0] : make_float4(0.f, 0.f, 0.f, 0.f);
                float4 vnext1 = AHEAD >= 2 && lb.x < lb.y ? wrow[1] : make_float4(0.f, 0.f, 0.f, 0.f);
                for (int tri = lb.x; tri < lb.y; ++tri, wrow += 3) {
                    if (STATS) ++nTris;
                    const float4 v00 = vnext;
                    const float4 v11n = vnext1;
                    if (tri + 1 < lb.y) {
                        vnext = wrow[3];
                        if (AHEAD >= 2) vnext1 = wrow[4];
                    }
                    const float Oz = fmaf(-oz, v00.z, fmaf(-oy, v00.y, fmaf(-ox, v00.x, v00.w)));
                    const float den = fmaf(dz, v00.z, fmaf(dy, v00.y, dx * v00.x));
                    const float t = Oz * (1.0f / den);
                    if (t > tmin && t <= hitT) {
                        const float4 v11 = AHEAD >= 2 ? v11n : wrow[1];
                        const float Ox = fmaf(oz, v11.z, fmaf(oy, v11.y, fmaf(ox, v11.x, v11.w)));
                        const float Dx = fmaf(dz, v11.z, fmaf(dy, v11.y, dx * v11.x));
                        const float u = fmaf(t, Dx, Ox);
                        if (u >= 0.0f) {
                            const float4 v22 = wrow[2];
                            const float Oy = fmaf(oz, v22.z, fmaf(oy, v22.y, fmaf(ox, v22.x, v22.w)));
                            const float Dy = fmaf(dz, v22.z, fmaf(dy, v22.y, dx * v22.x));
                            const float v = fmaf(t, Dy, Oy);
                            if (v >= 0.0f && u + v <= 1.0f) {
                                const int32_t id = a.tri_idx[triBase + tri];
                                // Closest hit; equal t resolves to the smaller triangle id.
                                if (t < hitT || id < hitId) {
                                    hitT = t;
                                    hitIndex = tri;
                                    hitId = id;
                                    hitU = u;
                                    hitV = v;
                                    if (a.any_hit) {
                                        nodeAddr = kSentinel;
                                        leafAddr2 = leafAddr3 = leafAddr4 = 0;
                                        break;
                                    }
                                }
                            }
                        }
                    }
                }
                if (POSTPONE >= 2 && leafAddr2 < 0) {  // the other held leaves next
                    leafAddr = leafAddr2;
                    leafAddr2 = POSTPONE >= 3 ? leafAddr3 : 0;
                    leafAddr3 = POSTPONE >= 4 ? leafAddr4 : 0;
                    leafAddr4 = 0;
                } else {
                    leafAddr = nodeAddr;
                    if (nodeAddr < 0) nodeAddr = stack.get(sp--);
                }
            }

            if (__popcll(ballot64(true)) < a.fetch_threshold) break;
        }

        if (nodeAddr == kSentinel) {
            if (hitIndex == -1)
                reinterpret_cast<int2*>(a.results)[2 * rayidx] = make_int2(-1, __float_as_int(hitT));
            else
                a.results[rayidx] =
                    make_int4(hitId, __float_as_int(hitT), __float_as_int(hitU), __float_as_int(hitV));
            if (STATS) a.stats[rayidx] = make_int2(nNodes, nTris);
        }
    }
}

// ---------------------------------------------------------------------------
// Quad / Oct traversal (HipTracerQuadKernels.cu, HipTracerOctKernels.cu; node
// words of HipBVHNode.h:13-114).  One ray per lane, as the Bin kernel: a node
// visit tests its N child boxes (child c: float4 c = min/max x, min/max y;
// float4 N + c = min/max z, link), sorts the hits by entry distance with a
// compare-exchange network in registers (ties by link), continues into the
// nearest and pushes the others farthest first.  The reference gives each
// ray N lanes; on 64-lane waves that leaves 64/N rays per wave to hide each
// node fetch, so here every lane keeps its own ray and the waves stay full.
// Leaves (link ~index) hold [begin, end) at words 4N+2, 4N+3; the triangle
// test, closest-hit tie rule, postponed leaf and wave refill are the Bin
// kernel's.
// ---------------------------------------------------------------------------

template <int N>
struct WideNet;
template <>
struct WideNet<4> {
    static constexpr int kPairs = 5;
    HX_DEV static int2 pair(int i) {
        constexpr int a[5] = {0, 2, 0, 1, 1}, b[5] = {1, 3, 2, 3, 2};
        return make_int2(a[i], b[i]);
    }
};
template <>
struct WideNet<8> {
    // Batcher's odd-even merge sort for 8 inputs.
    static constexpr int kPairs = 19;
    HX_DEV static int2 pair(int i) {
        constexpr int a[19] = {0, 2, 4, 6, 0, 1, 4, 5, 1, 5, 0, 1, 2, 3, 2, 3, 1, 3, 5};
        constexpr int b[19] = {1, 3, 5, 7, 2, 3, 6, 7, 2, 6, 4, 5, 6, 7, 4, 5, 2, 4, 6};
        return make_int2(a[i], b[i]);
    }
};

template <int N>
struct WideStackSize {
    static constexpr int kSize = N == 4 ? 96 : 160;
};

template <int N, bool STATS, bool ORDERED, int LDS_DEPTH>
__global__ __launch_bounds__(kTraceBlock) void k_trace_wide(TraceArgs a) {
    constexpr int kSize = WideStackSize<N>::kSize;
    __shared__ int32_t s_stack[LDS_DEPTH > 0 ? LDS_DEPTH * kTraceBlock : 1];
    TraversalStack<LDS_DEPTH, kSize> stack;
    stack.col = (lds_i32*)(s_stack + threadIdx.x);
    int sp = 0;
    float ox = 0.f, oy = 0.f, oz = 0.f, tmin = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
    float idx = 0.f, idy = 0.f, idz = 0.f, oodx = 0.f, oody = 0.f, oodz = 0.f;
    float hitT = 0.f, hitU = 0.f, hitV = 0.f;
    int32_t nodeAddr = kSentinel, leafAddr = 0, hitIndex = -1, hitId = -1;
    int64_t rayidx = 0;
    int32_t nNodes = 0, nTris = 0;
    int32_t* rayCounter = a.control;
    const float4* nodes4 = reinterpret_cast<const float4*>(a.nodes);

    while (true) {
        const bool terminated = nodeAddr == kSentinel;
        const uint64_t tmask = ballot64(terminated);
        const int ntm = __popcll(tmask);
        const int below = count_below(tmask);
        const int leader = __ffsll((unsigned long long)tmask) - 1;
        int base = 0;
        if (terminated && below == 0) base = atomicAdd(rayCounter, ntm);
        base = __shfl(base, leader, kWave);
        if (terminated) {
            rayidx = (int64_t)base + below;
            if (rayidx >= a.n_rays) break;
            if (ORDERED) rayidx = a.order[rayidx];
            const float4 o = a.rays[2 * rayidx + 0];
            const float4 d = a.rays[2 * rayidx + 1];
            ox = o.x; oy = o.y; oz = o.z; tmin = o.w;
            dx = d.x; dy = d.y; dz = d.z; hitT = d.w;
            const float ooeps = 8.2718061255302767e-25f;  // exp2f(-80)
            idx = 1.0f / (fabsf(dx) > ooeps ? dx : copysignf(ooeps, dx));
            idy = 1.0f / (fabsf(dy) > ooeps ? dy : copysignf(ooeps, dy));
            idz = 1.0f / (fabsf(dz) > ooeps ? dz : copysignf(ooeps, dz));
            oodx = ox * idx;
            oody = oy * idy;
            oodz = oz * idz;
            stack.put(0, kSentinel);
            sp = 0;
            leafAddr = 0;
            nodeAddr = 0;
            hitIndex = -1;
            hitId = -1;
            if (STATS) { nNodes = 0; nTris = 0; }
        }
        while (nodeAddr != kSentinel) {
            while ((uint32_t)nodeAddr < (uint32_t)kSentinel) {
                if (STATS) ++nNodes;
                float dist[N];
                int32_t link[N];
                // Child c's test; written out in both node-read branches so that
                // a whole-wave node is tested straight from scalar registers.
                auto child_test = [&](int c, float4 xy, float4 zl) {
                    const float lox = fmaf(xy.x, idx, -oodx);
                    const float hix = fmaf(xy.y, idx, -oodx);
                    const float loy = fmaf(xy.z, idy, -oody);
                    const float hiy = fmaf(xy.w, idy, -oody);
                    const float loz = fmaf(zl.x, idz, -oodz);
                    const float hiz = fmaf(zl.y, idz, -oodz);
                    const float cmin = span_begin(lox, hix, loy, hiy, loz, hiz, tmin);
                    const float cmax = span_end(lox, hix, loy, hiy, loz, hiz, hitT);
                    const bool hit = cmax >= cmin;
                    dist[c] = hit ? cmin : __int_as_float(0x7f800000);
                    link[c] = hit ? __float_as_int(zl.z) : kSentinel;
                };
                const int32_t n0 = __builtin_amdgcn_readfirstlane(nodeAddr);
                if (__all(nodeAddr == n0)) {
                    typedef const __attribute__((address_space(4))) float const_f32;
                    const_f32* f = (const_f32*)(nodes4 + (int64_t)n0 * 2 * N);
#pragma unroll
                    for (int c = 0; c < N; ++c)
                        child_test(c, make_float4(f[4 * c], f[4 * c + 1], f[4 * c + 2], f[4 * c + 3]),
                                   make_float4(f[4 * (N + c)], f[4 * (N + c) + 1], f[4 * (N + c) + 2],
                                               f[4 * (N + c) + 3]));
                } else {
                    const float4* np = nodes4 + (int64_t)nodeAddr * 2 * N;
                    float4 xys[N], zls[N];
#pragma unroll
                    for (int c = 0; c < N; ++c) {
                        xys[c] = np[c];
                        zls[c] = np[N + c];
                    }
#pragma unroll
                    for (int c = 0; c < N; ++c) child_test(c, xys[c], zls[c]);
                }
#pragma unroll
                for (int q = 0; q < WideNet<N>::kPairs; ++q) {
                    const int2 pr = WideNet<N>::pair(q);
                    const float da = dist[pr.x], db = dist[pr.y];
                    const int32_t la = link[pr.x], lb = link[pr.y];
                    const bool sw = db < da || (db == da && lb < la);
                    dist[pr.x] = sw ? db : da;
                    dist[pr.y] = sw ? da : db;
                    link[pr.x] = sw ? lb : la;
                    link[pr.y] = sw ? la : lb;
                }
                int cnt = 0;
#pragma unroll
                for (int c = 0; c < N; ++c) cnt += link[c] != kSentinel ? 1 : 0;
                if (cnt == 0) {
                    nodeAddr = stack.get(sp--);
                } else {
                    if (sp + cnt >= kSize) {
                        atomicAdd(a.control + 1, 1);
                        nodeAddr = kSentinel;
                        leafAddr = 0;
                        break;
                    }
#pragma unroll
                    for (int c = N - 1; c >= 1; --c)
                        if (c < cnt) stack.put(++sp, link[c]);
                    nodeAddr = link[0];
                }
                if (nodeAddr < 0 && leafAddr >= 0) {
                    leafAddr = nodeAddr;
                    nodeAddr = stack.get(sp--);
                }
                if (!__any(leafAddr >= 0)) break;
            }

            while (leafAddr < 0) {
                const float4 lb = nodes4[(int64_t)(~leafAddr) * 2 * N + N];
                const int end = __float_as_int(lb.w);
                for (int tri = __float_as_int(lb.z); tri < end; ++tri) {
                    if (STATS) ++nTris;
                    const float4 v00 = a.woop[3 * (int64_t)tri + 0];
                    const float Oz = fmaf(-oz, v00.z, fmaf(-oy, v00.y, fmaf(-ox, v00.x, v00.w)));
                    const float den = fmaf(dz, v00.z, fmaf(dy, v00.y, dx * v00.x));
                    const float t = Oz * (1.0f / den);
                    if (t > tmin && t <= hitT) {
                        const float4 v11 = a.woop[3 * (int64_t)tri + 1];
                        const float Ox = fmaf(oz, v11.z, fmaf(oy, v11.y, fmaf(ox, v11.x, v11.w)));
                        const float Dx = fmaf(dz, v11.z, fmaf(dy, v11.y, dx * v11.x));
                        const float u = fmaf(t, Dx, Ox);
                        if (u >= 0.0f) {
                            const float4 v22 = a.woop[3 * (int64_t)tri + 2];
                            const float Oy = fmaf(oz, v22.z, fmaf(oy, v22.y, fmaf(ox, v22.x, v22.w)));
                            const float Dy = fmaf(dz, v22.z, fmaf(dy, v22.y, dx * v22.x));
                            const float v = fmaf(t, Dy, Oy);
                            if (v >= 0.0f && u + v <= 1.0f) {
                                const int32_t id = a.tri_idx[tri];
                                // Closest hit; equal t resolves to the smaller triangle id.
                                if (t < hitT || id < hitId) {
                                    hitT = t;
                                    hitIndex = tri;
                                    hitId = id;
                                    hitU = u;
                                    hitV = v;
                                    if (a.any_hit) {
                                        nodeAddr = kSentinel;
                                        break;
                                    }
                                }
                            }
                        }
                    }
                }
                leafAddr = nodeAddr;
                if (nodeAddr < 0) nodeAddr = stack.get(sp--);
            }

            if (__popcll(ballot64(true)) < a.fetch_threshold) break;
        }

        if (nodeAddr == kSentinel) {
            if (hitIndex == -1)
                reinterpret_cast<int2*>(a.results)[2 * rayidx] = make_int2(-1, __float_as_int(hitT));
            else
                a.results[rayidx] =
                    make_int4(hitId, __float_as_int(hitT), __float_as_int(hitU), __float_as_int(hitV));
            if (STATS) a.stats[rayidx] = make_int2(nNodes, nTris);
        }
    }
}

// ---------------------------------------------------------------------------
// Ray Morton keys for the [0,1]^3 ray box (RayBuffer.cpp:63-66).
// ---------------------------------------------------------------------------

HX_DEV float fmin_n(float a, float b) { return (a != a) ? b : (b != b) ? a : ((b < a) ? b : a); }
HX_DEV float fmax_n(float a, float b) { return (a != a) ? b : (b != b) ? a : ((b > a) ? b : a); }

__global__ void k_ray_keys_aila(const float4* __restrict__ rays, int64_t n, int bits, uint64_t* __restrict__ keys) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 o = rays[2 * i], d = rays[2 * i + 1];
    const float ax = (o.x - 0.0f) / (1.0f - 0.0f), ay = (o.y - 0.0f) / (1.0f - 0.0f), az = (o.z - 0.0f) / (1.0f - 0.0f);
    float ls = 0.0f;
    ls += d.x * d.x;
    ls += d.y * d.y;
    ls += d.z * d.z;
    const float len = sqrtf(ls);
    const float rl = (len != 0.0f) ? 1.0f / len : 0.0f;
    const float bx = (d.x * rl + 1.0f) * 0.5f, by = (d.y * rl + 1.0f) * 0.5f, bz = (d.z * rl + 1.0f) * 0.5f;
    const int iax = f2i_sat(ax * 8191.0f), iay = f2i_sat(ay * 8191.0f), iaz = f2i_sat(az * 8191.0f);
    const int ibx = f2i_sat(bx * 8191.0f), iby = f2i_sat(by * 8191.0f), ibz = f2i_sat(bz * 8191.0f);
    uint64_t code = 0;
    code |= (uint64_t)((iax >> 12) & 1) << 63;
    code |= (uint64_t)((iay >> 12) & 1) << 62;
    code |= (uint64_t)((iaz >> 12) & 1) << 61;
    code |= (uint64_t)((iax >> 11) & 1) << 60;
    code |= (uint64_t)((iay >> 11) & 1) << 59;
    code |= (uint64_t)((iaz >> 11) & 1) << 58;
    code |= (uint64_t)((iax >> 10) & 1) << 57;
    code |= (uint64_t)((iay >> 10) & 1) << 56;
    code |= (uint64_t)((iaz >> 10) & 1) << 55;
#pragma unroll
    for (int k = 9; k >= 1; --k) {
        code |= (uint64_t)((iax >> k) & 1) << (6 * k + 0);
        code |= (uint64_t)((iay >> k) & 1) << (6 * k - 1);
        code |= (uint64_t)((iaz >> k) & 1) << (6 * k - 2);
    }
#pragma unroll
    for (int k = 12; k >= 4; --k) {
        code |= (uint64_t)((ibx >> k) & 1) << (6 * k - 21);
        code |= (uint64_t)((iby >> k) & 1) << (6 * k - 22);
        code |= (uint64_t)((ibz >> k) & 1) << (6 * k - 23);
    }
    keys[i] = code >> (64 - bits);
}

__global__ void k_ray_keys_twopoint(const float4* __restrict__ rays, int64_t n, int bits, float ray_length,
                                    uint64_t* __restrict__ keys) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 ro = rays[2 * i], rd = rays[2 * i + 1];
    const float o[3] = {ro.x, ro.y, ro.z}, d[3] = {rd.x, rd.y, rd.z};
    float tmn[3], tmx[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (d[k] >= 0.0f) {
            tmn[k] = (0.0f - o[k]) / d[k];
            tmx[k] = (1.0f - o[k]) / d[k];
        } else {
            tmn[k] = (1.0f - o[k]) / d[k];
            tmx[k] = (0.0f - o[k]) / d[k];
        }
    }
    const float t0 = fmax_n(tmn[2], fmax_n(tmn[0], tmn[1]));
    const float t1 = fmin_n(tmx[2], fmin_n(tmx[0], tmx[1]));
    const float boxT = (t0 > t1) ? -1.0f : (1.0f - 1.0e-3f) * (t0 >= 0.0f ? t0 : t1);
    float len = ray_length;
    len = fmin_n(len, fmax_n(rd.w, 0.0f));
    len = fmin_n(len, boxT);
    int ia[3], ib[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float av = (o[k] - 0.0f) * (1.0f / (1.0f - 0.0f));
        const float bv = ((o[k] + len * d[k]) - 0.0f) * (1.0f / (1.0f - 0.0f));
        ia[k] = f2i_sat(av * 8191.0f);
        ib[k] = f2i_sat(bv * 8191.0f);
    }
    uint64_t code = 0;
#pragma unroll
    for (int k = 12; k >= 2; --k) {
        code |= (uint64_t)((ia[0] >> k) & 1) << (6 * k - 9);
        code |= (uint64_t)((ia[1] >> k) & 1) << (6 * k - 10);
        code |= (uint64_t)((ia[2] >> k) & 1) << (6 * k - 11);
    }
#pragma unroll
    for (int k = 12; k >= 3; --k) {
        code |= (uint64_t)((ib[0] >> k) & 1) << (6 * k - 12);
        code |= (uint64_t)((ib[1] >> k) & 1) << (6 * k - 13);
        code |= (uint64_t)((ib[2] >> k) & 1) << (6 * k - 14);
    }
    code |= (uint64_t)(ib[0] & 1);
    keys[i] = code >> (64 - bits);
}

__global__ void k_zero_control(int32_t* control) {
    if (threadIdx.x < 2) control[threadIdx.x] = 0;
}

template <bool STATS, bool ORDERED, int LDS_DEPTH, int POSTPONE, int AHEAD>
int launch_trace_t(Context* c, const TraceArgs& a) {
    // Occupancy limit of this instantiation, queried once per process.
    static int occ = 0;
    if (occ == 0) {
        int q = 0;
        HX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, k_trace<STATS, ORDERED, LDS_DEPTH, POSTPONE, AHEAD>,
                                                             kTraceBlock, 0));
        occ = std::max(q, 1);
    }
    const int per_cu = c->tune.trace_blocks_per_cu > 0 ? std::min(c->tune.trace_blocks_per_cu, occ) : occ;
    const int64_t want = (a.n_rays + kTraceBlock - 1) / kTraceBlock;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->cu_count * per_cu, want));
    hipLaunchKernelGGL((k_trace<STATS, ORDERED, LDS_DEPTH, POSTPONE, AHEAD>), dim3(blocks), dim3(kTraceBlock), 0,
                       c->stream, a);
    return check_launch("trace");
}

template <bool STATS, bool ORDERED, int POSTPONE, int AHEAD>
int launch_trace_depth(Context* c, const TraceArgs& a) {
    switch (c->tune.trace_lds_stack) {
        case 0: return launch_trace_t<STATS, ORDERED, 0, POSTPONE, AHEAD>(c, a);
        case 16: return launch_trace_t<STATS, ORDERED, 16, POSTPONE, AHEAD>(c, a);
        case 24: return launch_trace_t<STATS, ORDERED, 24, POSTPONE, AHEAD>(c, a);
        case 32: return launch_trace_t<STATS, ORDERED, 32, POSTPONE, AHEAD>(c, a);
        default: return launch_trace_t<STATS, ORDERED, 64, POSTPONE, AHEAD>(c, a);
    }
}

template <bool STATS, bool ORDERED, int POSTPONE>
int launch_trace_ahead(Context* c, const TraceArgs& a) {
    if (c->tune.trace_woop_ahead >= 2) return launch_trace_depth<STATS, ORDERED, POSTPONE, 2>(c, a);
    return launch_trace_depth<STATS, ORDERED, POSTPONE, 1>(c, a);
}

template <bool STATS, bool ORDERED>
int launch_trace_postpone(Context* c, const TraceArgs& a) {
    // An any-hit trace reports the first hit in traversal order, which holding
    // leaves back changes: it keeps the reference's one postponed leaf.
    if (a.any_hit) return launch_trace_ahead<STATS, ORDERED, 1>(c, a);
    switch (c->tune.trace_postpone) {
        case 2: return launch_trace_ahead<STATS, ORDERED, 2>(c, a);
        case 3: return launch_trace_ahead<STATS, ORDERED, 3>(c, a);
        case 4: return launch_trace_ahead<STATS, ORDERED, 4>(c, a);
        default: return launch_trace_ahead<STATS, ORDERED, 1>(c, a);
    }
}

template <int N, bool STATS, bool ORDERED, int LDS_DEPTH>
int launch_trace_wide_t(Context* c, const TraceArgs& a) {
    static int occ = 0;
    if (occ == 0) {
        int q = 0;
        HX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, k_trace_wide<N, STATS, ORDERED, LDS_DEPTH>, kTraceBlock,
                                                             0));
        occ = std::max(q, 1);
    }
    const int per_cu = c->tune.trace_blocks_per_cu > 0 ? std::min(c->tune.trace_blocks_per_cu, occ) : occ;
    const int64_t want = (a.n_rays + kTraceBlock - 1) / kTraceBlock;
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->cu_count * per_cu, want));
    hipLaunchKernelGGL((k_trace_wide<N, STATS, ORDERED, LDS_DEPTH>), dim3(blocks), dim3(kTraceBlock), 0, c->stream, a);
    return check_launch("trace_wide");
}

template <int N, bool STATS, bool ORDERED>
int launch_trace_wide_depth(Context* c, const TraceArgs& a) {
    switch (c->tune.trace_lds_stack) {
        case 0: return launch_trace_wide_t<N, STATS, ORDERED, 0>(c, a);
        case 16: return launch_trace_wide_t<N, STATS, ORDERED, 16>(c, a);
        case 24: return launch_trace_wide_t<N, STATS, ORDERED, 24>(c, a);
        case 32: return launch_trace_wide_t<N, STATS, ORDERED, 32>(c, a);
        default: return launch_trace_wide_t<N, STATS, ORDERED, 64>(c, a);
    }
}

template <int N>
int launch_trace_wide_n(Context* c, const TraceArgs& a) {
    if (a.stats)
        return a.order ? launch_trace_wide_depth<N, true, true>(c, a) : launch_trace_wide_depth<N, true, false>(c, a);
    return a.order ? launch_trace_wide_depth<N, false, true>(c, a) : launch_trace_wide_depth<N, false, false>(c, a);
}

}  // namespace

int launch_trace_wide(Context* c, int N, const TraceArgs& a) {
    if (a.n_rays <= 0) return HX_OK;
    hipLaunchKernelGGL(k_zero_control, dim3(1), dim3(64), 0, c->stream, a.control);
    HX_TRY(check_launch("zero_control"));
    LaunchScope ls(c, a.order ? "trace_wide_sorted" : "trace_wide");
    if (N == 4) return launch_trace_wide_n<4>(c, a);
    if (N == 8) return launch_trace_wide_n<8>(c, a);
    set_error("wide trace: %d children per node is not supported (4 or 8)", N);
    return HX_ERR_INVALID_ARG;
}

int launch_trace(Context* c, const TraceArgs& a) {
    if (a.n_rays <= 0) return HX_OK;
    hipLaunchKernelGGL(k_zero_control, dim3(1), dim3(64), 0, c->stream, a.control);
    HX_TRY(check_launch("zero_control"));
    LaunchScope ls(c, a.order ? "trace_sorted" : "trace");
    if (a.stats)
        return a.order ? launch_trace_postpone<true, true>(c, a) : launch_trace_postpone<true, false>(c, a);
    return a.order ? launch_trace_postpone<false, true>(c, a) : launch_trace_postpone<false, false>(c, a);
}

int launch_ray_keys(Context* c, const float4* rays, int64_t n, int method, int bits, float ray_length,
                    uint64_t* keys) {
    if (n <= 0) return HX_OK;
    LaunchScope ls(c, "ray_keys");
    const unsigned g = (unsigned)((n + 255) / 256);
    if (method == HX_RAYKEY_AILA)
        hipLaunchKernelGGL(k_ray_keys_aila, dim3(g), dim3(256), 0, c->stream, rays, n, bits, keys);
    else
        hipLaunchKernelGGL(k_ray_keys_twopoint, dim3(g), dim3(256), 0, c->stream, rays, n, bits, ray_length, keys);
    return check_launch("ray_keys");
}

}  // namespace hx
