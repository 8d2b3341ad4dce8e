// hx_insertion.hip -- insertion-based optimisation of a binary BVH on gfx950.
//
// Reference (src/hippie/rt/bvh/): InsertionBuilder.cpp:22-274 (the step loop),
// InsertionBuilderKernels.cu:11-515 (findBestNode, lockNodes, checkLocks,
// reinsert, computeCost), LBVHBuilderKernels.cu:254-305 (refit).
//
// One step:
//   find     every node i = mod*k + rem (i > 0) searches the best new sibling
//            by the reference's branch-and-bound walk, one thread per node;
//   lock     a found move takes the 64-bit maximum of (reduction bits << 32 |
//            i) on every node it would write;
//   check    a move is kept only if it holds all of them, so the kept moves
//            write disjoint nodes and the order they are applied in is free;
//   reinsert the kept moves are applied;
//   refit    boxes by an atomic-counter climb from the leaves; boxes move
//            between workgroups with agent-coherent (sc1) accesses, and each
//            writer waits for its stores before it bumps the parent's counter;
//   cost     the SAH cost as an exact sum: each node's term c * (area / scene
//            area) is truncated to a multiple of 2^-32 and the multiples are
//            added as 64-bit integers, so the total (and every convergence
//            decision taken from it) is independent of the order of the
//            additions.  The reference adds float block sums atomically.
// The host reads the cost and the counts back after each step and applies the
// reference's rules: keep the best tree, stop when the cost changes by at
// most 0.01 (or nothing moved) at mod 1, lower mod when it stalls.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

HX_DEV Box3 node_box(const InsArgs& a, int v) {
    const float4 mn = a.nmin[v], mx = a.nmax[v];
    Box3 b;
    b.mn = make_float3(mn.x, mn.y, mn.z);
    b.mx = make_float3(mx.x, mx.y, mx.z);
    return b;
}

// AABB(a, b) (util/AABB.h:22): an empty box grown by a's corners, then b's.
HX_DEV Box3 box_union(const Box3& a, const Box3& b) {
    Box3 r = box_empty();
    box_grow(r, a.mn);
    box_grow(r, a.mx);
    box_grow(r, b.mn);
    box_grow(r, b.mx);
    return r;
}

HX_DEV float area_of(const Box3& b) { return box_area(b.mn, b.mx); }

HX_DEV int sibling_of(const InsArgs& a, int v, int p) { return a.left[p] == v ? a.right[p] : a.left[p]; }

// The walk's view of the tree, packed per node into one 64-byte record so
// that a step issues two vector loads instead of four or five (boxes, parent
// and child links otherwise live in five arrays, each its own cache line):
//   [0] min corner, parent    [1] max corner, left child
//   [2] left child, right child (internal nodes; -1 for leaves), sibling,
//       1 if the node is its parent's left child (the root: -1, 0)    [3] unused
// With the sibling in the node's own record a step up the tree reads the
// line it already holds instead of the parent's.  Written before every find
// from the arrays the other kernels update.
__global__ __launch_bounds__(kBlock) void k_ins_pack(InsArgs a) {
    const int N = (int)(2 * a.n - 1);
    const int nl = (int)a.n - 1;
    const int v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= N) return;
    const float4 mn = a.nmin[v], mx = a.nmax[v];
    const int l = v < nl ? a.left[v] : -1, r = v < nl ? a.right[v] : -1;
    const int p = a.parent[v];
    int sib = -1, is_left = 0;
    if (p >= 0) {
        const int pl = a.left[p];
        is_left = pl == v ? 1 : 0;
        sib = is_left ? a.right[p] : pl;
    }
    float4* rec = reinterpret_cast<float4*>(a.rec) + 4 * (int64_t)v;
    rec[0] = make_float4(mn.x, mn.y, mn.z, __int_as_float(p));
    rec[1] = make_float4(mx.x, mx.y, mx.z, __int_as_float(l));
    reinterpret_cast<int4*>(rec)[2] = make_int4(l, r, sib, is_left);
}

struct WalkNode {
    Box3 box;
    int parent, left;
};

HX_DEV WalkNode walk_node(const InsArgs& a, int v) {
    const float4* rec = reinterpret_cast<const float4*>(a.rec) + 4 * (int64_t)v;
    const float4 r0 = rec[0], r1 = rec[1];
    WalkNode w;
    w.box.mn = make_float3(r0.x, r0.y, r0.z);
    w.box.mx = make_float3(r1.x, r1.y, r1.z);
    w.parent = __float_as_int(r0.w);
    w.left = __float_as_int(r1.w);
    return w;
}

// (sibling, 1 if v is its parent's left child)
HX_DEV int2 walk_side(const InsArgs& a, int v) {
    return *reinterpret_cast<const int2*>(reinterpret_cast<const int32_t*>(a.rec + 4 * (int64_t)v + 2) + 2);
}

// findBestNode (InsertionBuilderKernels.cu:11-185).  The walk visits the
// sibling's subtree (the sibling itself is the current position and is not a
// candidate), then climbs: at each ancestor it visits the other child's
// subtree and then the ancestor itself (neither the old parent nor the root
// is a candidate).  Going down, `red` gains each visited node's growth; a
// subtree is skipped when even a perfect merge cannot beat the best so far.
// The walk is bounded by 4N + 64 visits (a consistent tree needs fewer than
// 2N + depth); a thread that runs out reports no move.
__global__ __launch_bounds__(kBlock) void k_ins_find(InsArgs a, int mod, int rem, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const int N = (int)(2 * a.n - 1);
    const int in = mod * k + rem;
    if (in <= 0 || in >= N) return;
    const int nl = (int)a.n - 1;
    const WalkNode inNode = walk_node(a, in);
    const int p = inNode.parent;
    const Box3 inBox = inNode.box;
    const float pArea = area_of(walk_node(a, p).box);
    const float bound = pArea - area_of(inBox);
    Box3 highest = box_empty();
    const int sib = walk_side(a, in).x;
    int out = sib, highestIdx = p;
    bool down = true;
    int best = -1;
    float bestRed = 0.0f, red = 0.0f;
    const int64_t limit = 4 * (int64_t)N + 64;
    int64_t it = 0;
    for (; it < limit; ++it) {
        WalkNode on = walk_node(a, out);
        Box3 outBox = on.box;
        Box3 merged = box_union(inBox, outBox);
        if (down) {
            if (out != sib) {
                const float direct = pArea - area_of(merged);
                if (bestRed < direct + red) {
                    bestRed = direct + red;
                    best = out;
                }
            }
            red += area_of(outBox) - area_of(merged);
            if (out >= nl || red + bound <= bestRed) down = false;
            else out = on.left;
        } else {
            int op = on.parent;
            red -= area_of(outBox) - area_of(merged);
            if (op == highestIdx) {
                box_grow(highest, outBox.mn);
                box_grow(highest, outBox.mx);
                out = op;
                const WalkNode up = walk_node(a, out);
                if (out != p && out != 0) {
                    merged = box_union(inBox, highest);
                    const float direct = pArea - area_of(merged);
                    if (bestRed < direct + red) {
                        bestRed = direct + red;
                        best = out;
                    }
                    outBox = up.box;
                    red += area_of(outBox) - area_of(highest);
                }
                op = up.parent;
                if (op < 0) break;
                highestIdx = op;
                down = true;
                out = walk_side(a, out).x;  // out's parent is highestIdx
            } else {
                const int2 side = walk_side(a, out);  // out's parent is op
                if (side.y) {
                    down = true;
                    out = side.x;
                } else {
                    out = op;
                }
            }
        }
    }
    if (it >= limit) best = -1;
    a.red[in] = bestRed;
    a.out[in] = best;
}

// Nodes a move of `in` beside `out` writes (lockNodes, :188-270): in, its
// sibling and parent, the grandparent (or, under the root, the sibling's
// children, which become the root's), out and out's parent.
HX_DEV int lock_set(const InsArgs& a, int in, int out, int set[8]) {
    const int nl = (int)a.n - 1;
    int k = 0;
    const int p = a.parent[in];
    const int sib = sibling_of(a, in, p);
    set[k++] = in;
    set[k++] = sib;
    set[k++] = p;
    if (p > 0) {
        set[k++] = a.parent[p];
    } else if (sib < nl) {
        set[k++] = a.left[sib];
        set[k++] = a.right[sib];
    }
    set[k++] = out;
    if (out > 0) {
        set[k++] = a.parent[out];
    } else {
        set[k++] = a.left[out];
        set[k++] = a.right[out];
    }
    return k;
}

HX_DEV unsigned long long lock_key(const InsArgs& a, int in) {
    return ((unsigned long long)__float_as_uint(a.red[in]) << 32) | (unsigned long long)(unsigned)in;
}

__global__ __launch_bounds__(kBlock) void k_ins_lock(InsArgs a, int mod, int rem, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const int N = (int)(2 * a.n - 1);
    const int in = mod * k + rem;
    if (in <= 0 || in >= N) return;
    const int out = a.out[in];
    if (out < 0) return;
    const unsigned long long key = lock_key(a, in);
    int set[8];
    const int m = lock_set(a, in, out, set);
    for (int j = 0; j < m; ++j) atomicMax(a.locks + set[j], key);
}

// Wave-aggregated add of a per-lane count to a 64-bit counter.
HX_DEV void count_add(unsigned long long* ctr, bool flag) {
    const uint64_t m = ballot64(flag);
    if (m && lane_id() == __builtin_ctzll(m)) atomicAdd(ctr, (unsigned long long)__builtin_popcountll(m));
}

// checkLocks (:273-358): acc[1] counts the moves found.
__global__ __launch_bounds__(kBlock) void k_ins_check(InsArgs a, int mod, int rem, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int N = (int)(2 * a.n - 1);
    const int in = mod * k + rem;
    const bool live = k < count && in > 0 && in < N;
    const int out = live ? a.out[in] : -1;
    if (out >= 0) {
        const unsigned long long key = lock_key(a, in);
        int set[8];
        const int m = lock_set(a, in, out, set);
        bool ok = true;
        for (int j = 0; j < m; ++j) ok = ok && a.locks[set[j]] == key;
        if (!ok) a.out[in] = -1;
    }
    count_add(a.acc + 1, out >= 0);
}

// reinsert (:361-455): unlink in's parent (under the root, the sibling's
// slot is reused and its children move up to the root), then make it the
// parent of out and in.  acc[2] counts the moves applied.
__global__ __launch_bounds__(kBlock) void k_ins_reinsert(InsArgs a, int mod, int rem, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const int N = (int)(2 * a.n - 1);
    const int in = mod * k + rem;
    const bool live = k < count && in > 0 && in < N;
    const int out = live ? a.out[in] : -1;
    if (out >= 0) {
        int p = a.parent[in];
        const int sib = sibling_of(a, in, p);
        if (p != 0) {
            const int pp = a.parent[p];
            if (a.left[pp] == p) a.left[pp] = sib;
            else a.right[pp] = sib;
            a.parent[sib] = pp;
        } else {
            const int sl = a.left[sib], sr = a.right[sib];
            a.left[0] = sl;
            a.right[0] = sr;
            a.parent[sl] = 0;
            a.parent[sr] = 0;
            p = sib;
        }
        if (out != 0) {
            const int op = a.parent[out];
            if (a.left[op] == out) a.left[op] = p;
            else a.right[op] = p;
            a.parent[p] = op;
            a.left[p] = out;
            a.right[p] = in;
            a.parent[out] = p;
            a.parent[in] = p;
        } else {
            const int ol = a.left[0], orr = a.right[0];
            a.left[p] = ol;
            a.right[p] = orr;
            a.parent[ol] = p;
            a.parent[orr] = p;
            a.parent[p] = 0;
            a.parent[in] = 0;
            a.left[0] = in;
            a.right[0] = p;
        }
    }
    count_add(a.acc + 2, out >= 0);
}

HX_DEV float4 ld_agent4(const float4* p) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
    const uint64_t lo = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float((uint32_t)lo), __uint_as_float((uint32_t)(lo >> 32)),
                       __uint_as_float((uint32_t)hi), __uint_as_float((uint32_t)(hi >> 32)));
}
HX_DEV void st_agent4(float4* p, float4 v) {
    uint64_t* q = reinterpret_cast<uint64_t*>(p);
    const uint64_t lo = (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32);
    const uint64_t hi = (uint64_t)__float_as_uint(v.z) | ((uint64_t)__float_as_uint(v.w) << 32);
    __hip_atomic_store(q, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
HX_DEV int ld_agent_i(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
HX_DEV void st_agent_i(int32_t* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// refit (LBVHBuilderKernels.cu:254-305): one thread per leaf climbs; the
// second arrival at a node computes its box (min corners, then max corners,
// left child first).  With SIZES the climb also records every internal
// node's leaf count and height (leaves: 1 and 0) for the relabelling.
template <bool SIZES>
__global__ __launch_bounds__(kBlock) void k_ins_refit(InsArgs a) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    const int nl = (int)a.n - 1;
    if (j > nl) return;
    int node = a.parent[nl + j];
    for (int guard = 0; node >= 0 && guard <= nl; ++guard) {
        if (__hip_atomic_fetch_add(&a.counters[node], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) break;
        const int l = a.left[node], r = a.right[node];
        const float4 lmn = ld_agent4(a.nmin + l), rmn = ld_agent4(a.nmin + r);
        const float4 lmx = ld_agent4(a.nmax + l), rmx = ld_agent4(a.nmax + r);
        Box3 b = box_empty();
        box_grow(b, make_float3(lmn.x, lmn.y, lmn.z));
        box_grow(b, make_float3(rmn.x, rmn.y, rmn.z));
        box_grow(b, make_float3(lmx.x, lmx.y, lmx.z));
        box_grow(b, make_float3(rmx.x, rmx.y, rmx.z));
        st_agent4(a.nmin + node, make_float4(b.mn.x, b.mn.y, b.mn.z, 0.0f));
        st_agent4(a.nmax + node, make_float4(b.mx.x, b.mx.y, b.mx.z, 0.0f));
        if (SIZES) {
            const int rl = l >= nl ? 1 : ld_agent_i(a.refs + l), rr = r >= nl ? 1 : ld_agent_i(a.refs + r);
            const int hl = l >= nl ? 0 : ld_agent_i(a.height + l), hr = r >= nl ? 0 : ld_agent_i(a.height + r);
            st_agent_i(a.refs + node, rl + rr);
            st_agent_i(a.height + node, 1 + max(hl, hr));
        }
        // Complete this node's stores before the parent's counter can be
        // observed (the compiler fence keeps the stores above the wait).
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        node = a.parent[node];
    }
}

// computeCost (:457-515) in 2^-32 fixed point; a term outside [0, 1e9]
// (a NaN from a zero-area scene) counts as 0.
HX_DEV unsigned long long fixed_term(float t) {
    if (!(t >= 0.0f) || !(t <= 1.0e9f)) return 0ull;
    return (unsigned long long)((double)t * 4294967296.0);
}

__global__ __launch_bounds__(kBlock) void k_ins_cost(InsArgs a) {
    const int64_t N = 2 * a.n - 1;
    const int nl = (int)a.n - 1;
    const float scene = area_of(node_box(a, 0));
    unsigned long long s = 0;
    for (int64_t v = blockIdx.x * (int64_t)kBlock + threadIdx.x; v < N; v += (int64_t)gridDim.x * kBlock) {
        const float P = area_of(node_box(a, (int)v)) / scene;
        s += fixed_term((v >= nl ? a.ci : a.ct) * P);
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)s, m, kWave);
        const uint32_t hi = __shfl_xor((uint32_t)(s >> 32), m, kWave);
        s += ((unsigned long long)hi << 32) | lo;
    }
    if (lane_id() == 0 && s) atomicAdd(a.acc, s);
}

__global__ void k_ins_zero(unsigned long long* p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_ins_copy3(const int32_t* __restrict__ a, const int32_t* __restrict__ b,
                                                     const int32_t* __restrict__ c, int32_t* __restrict__ da,
                                                     int32_t* __restrict__ db, int32_t* __restrict__ dc, int count) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < count) {
        da[i] = a[i];
        db[i] = b[i];
        dc[i] = c[i];
    }
}

int copy_tree(Context* c, const int32_t* l, const int32_t* r, const int32_t* p, int32_t* dl, int32_t* dr,
              int32_t* dp, int count) {
    hipLaunchKernelGGL(k_ins_copy3, dim3(grid_for(count)), dim3(kBlock), 0, c->stream, l, r, p, dl, dr, dp, count);
    return check_launch("ins_copy");
}

}  // namespace

int launch_ins_refit(Context* c, const InsArgs& a, bool sizes) {
    const int64_t n = a.n;
    if (n < 2) return HX_OK;
    HX_TRY(launch_fill_i32(c, a.counters, 0, n - 1));
    if (sizes)
        hipLaunchKernelGGL(k_ins_refit<true>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_ins_refit<false>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
    return check_launch("ins_refit");
}

int launch_insertion(Context* c, const InsArgs& a, int64_t* steps_out, int64_t* reinserted_out) {
    const int64_t n = a.n;
    const int N = (int)(2 * n - 1);
    int steps = 0, mod_cur = std::max(a.mod, 1);
    int64_t total = 0;
    double cost = 3.402823466e+38, best = 3.402823466e+38, prev;
    bool best_is_current = true;
    LaunchScope ls(c, "insertion");
    HX_TRY(copy_tree(c, a.left, a.right, a.parent, a.best_left, a.best_right, a.best_parent, N));
    const unsigned cost_grid = std::min<unsigned>(grid_for(N), 1024u);
    while (steps < kInsMaxSteps) {
        prev = cost;
        const int rem = steps % mod_cur;
        const int count = (N - rem + mod_cur - 1) / mod_cur;
        const unsigned g = grid_for(count);
        hipLaunchKernelGGL(k_ins_zero, dim3(1), dim3(64), 0, c->stream, a.acc, 4);
        HX_TRY(check_launch("ins_zero"));
        hipLaunchKernelGGL(k_ins_pack, dim3(grid_for(N)), dim3(kBlock), 0, c->stream, a);
        HX_TRY(check_launch("ins_pack"));
        hipLaunchKernelGGL(k_ins_find, dim3(g), dim3(kBlock), 0, c->stream, a, mod_cur, rem, count);
        HX_TRY(check_launch("ins_find"));
        HX_TRY(launch_fill_i32(c, reinterpret_cast<int32_t*>(a.locks), 0, 2 * (int64_t)N));
        hipLaunchKernelGGL(k_ins_lock, dim3(g), dim3(kBlock), 0, c->stream, a, mod_cur, rem, count);
        HX_TRY(check_launch("ins_lock"));
        hipLaunchKernelGGL(k_ins_check, dim3(g), dim3(kBlock), 0, c->stream, a, mod_cur, rem, count);
        HX_TRY(check_launch("ins_check"));
        hipLaunchKernelGGL(k_ins_reinsert, dim3(g), dim3(kBlock), 0, c->stream, a, mod_cur, rem, count);
        HX_TRY(check_launch("ins_reinsert"));
        HX_TRY(launch_ins_refit(c, a, false));
        hipLaunchKernelGGL(k_ins_cost, dim3(cost_grid), dim3(kBlock), 0, c->stream, a);
        HX_TRY(check_launch("ins_cost"));
        unsigned long long h[4] = {0, 0, 0, 0};
        HX_HIP(hipMemcpyAsync(h, a.acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
        cost = (double)h[0] * (1.0 / 4294967296.0);
        const int64_t inserted = (int64_t)h[2];
        total += inserted;
        best_is_current = false;
        if (best > cost) {
            best = cost;
            HX_TRY(copy_tree(c, a.left, a.right, a.parent, a.best_left, a.best_right, a.best_parent, N));
            best_is_current = true;
        }
        ++steps;
        const bool stalled = std::fabs(prev - cost) <= 0.01;
        if ((stalled || inserted == 0) && mod_cur == 1) break;
        if (stalled) mod_cur = std::max(1, mod_cur - 1);
    }
    if (!best_is_current)
        HX_TRY(copy_tree(c, a.best_left, a.best_right, a.best_parent, a.left, a.right, a.parent, N));
    HX_TRY(launch_ins_refit(c, a, true));
    if (steps_out) *steps_out = steps;
    if (reinserted_out) *reinserted_out = total;
    return HX_OK;
}

}  // namespace hx
