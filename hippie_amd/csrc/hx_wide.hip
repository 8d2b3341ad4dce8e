// hx_wide.hip -- the Quad (128-byte) and Oct (256-byte) layouts on gfx950:
// collapse of a built binary tree into N-wide nodes, and their traversal.
//
// Reference (src/hippie/): rt/bvh/BVH.h:74-120 (recomputeCostsWide), :228-291
// (collapseWide), :401-423; rt/bvh/BVHConverter.cpp:12-78 (breadth-first
// numbering and node records); rt/bvh/HipBVHNode.h:13-114 (word map);
// rt/tracer/HipTracerQuadKernels.cu:13-850 and HipTracerOctKernels.cu:14-902
// (traversal with N lanes per ray); HipBVHKernels.cu:184-238 (cost).  The
// reference collapses on the host; here every stage runs on the device:
//
//  1. k_wide_dp: bottom-up over the binary tree (one thread per leaf, arrival
//     counters), the DP of the best cost of a subtree with j = 1..N-1 slots and
//     the split (k, j-k) of those slots between its children (or "leaf").
//  2. Top-down, one launch pair per wide level: each wide interior expands its
//     children (the first child whose best split fits the remaining slots is
//     replaced by its two children, as collapseWide does), a scan of (children,
//     interior children) per level numbers the next level breadth-first --
//     interiors 0.. in BFS order, leaves after all interiors in BFS order, as
//     BVHConverter's queue does.
//  3. k_wide_emit writes the node records from the binary boxes; a refit
//     reruns it after the binary climb (same entries, new boxes).
//
// The traversal of these layouts is k_trace_wide in hx_trace.hip.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kWideBlock = 256;
constexpr uint16_t kLeafCnt = 0xFFFF;

HX_DEV float area4(float4 mn, float4 mx) { return box_area(make_float3(mn.x, mn.y, mn.z), make_float3(mx.x, mx.y, mx.z)); }
HX_DEV int div_ceil_i(int a, int b) { return (a + b - 1) / b; }
HX_DEV uint16_t pack_cnt(int x, int y) { return (uint16_t)(x | (y << 8)); }
HX_DEV int cnt_x(uint16_t c) { return c == kLeafCnt ? -1 : (int)(c & 0xFF); }
HX_DEV int cnt_y(uint16_t c) { return c == kLeafCnt ? -1 : (int)(c >> 8); }

HX_DEV int root_of(const WideArgs& a) { return a.root_id ? *a.root_id : 0; }
HX_DEV float ld_agent(float* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
HX_DEV void st_agent(float* p, float v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Cost of a subtree (binary node b) with j slots, j = 1..N-1, at cost[(j-1) T + b]
// and the split of those slots at cnt[(j-1) T + b]; the areas are divided by the
// root's as the reference writes it (ci * A / Ar * ceil(size / N) * N).
template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_dp(WideArgs a) {
    const int64_t p = blockIdx.x * (int64_t)kWideBlock + threadIdx.x;
    if (p >= a.n) return;
    const int64_t T = 2 * a.n - 1;
    const int S = N - 1;
    const int root = root_of(a);
    const float Ar = area4(a.nmin[root], a.nmax[root]);
    // The leaf: one reference, ceil(1 / N) * N slots of leaf cost.  Costs
    // pass between threads (and XCDs) through agent-coherent relaxed stores
    // and loads; a thread's stores complete before it bumps the parent's
    // arrival counter, so the second arrival reads both children's costs.
    {
        const int64_t b = a.n - 1 + p;
        const float A = area4(a.nmin[b], a.nmax[b]);
        const float c = a.ci * A / Ar * (float)div_ceil_i(1, N) * (float)N;
        for (int j = 0; j < S; ++j) {
            st_agent(a.cost + j * T + b, c);
            a.cnt[j * T + b] = kLeafCnt;
        }
    }
    if (a.n == 1) return;
    int node = a.parent[a.n - 1 + p];
    while (node >= 0) {
        __builtin_amdgcn_s_waitcnt(0);
        if (__hip_atomic_fetch_add(a.counters + node, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
        const int l = a.left[node], r = a.right[node];
        const float A = area4(a.nmin[node], a.nmax[node]);
        const int size = a.last[node] - a.first[node] + 1;
        const float leaf_cost = size > a.max_leaf ? kMaxFloat : a.ci * A / Ar * (float)div_ceil_i(size, N) * (float)N;
        float cl[N - 1], cr[N - 1];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            cl[j] = ld_agent(a.cost + j * T + l);
            cr[j] = ld_agent(a.cost + j * T + r);
        }
        float subtree = a.ct * A / Ar;
        float cd = kMaxFloat;
        uint16_t bc = kLeafCnt;
#pragma unroll
        for (int k = 1; k < N; ++k) {
            const float c = cl[k - 1] + cr[N - k - 1];
            if (cd > c) {
                cd = c;
                bc = pack_cnt(k, N - k);
            }
        }
        subtree += cd;
        float prev;
        uint16_t prevc;
        if (leaf_cost > subtree) {
            prev = subtree;
            prevc = bc;
        } else {
            prev = leaf_cost;
            prevc = kLeafCnt;
        }
        st_agent(a.cost + node, prev);
        a.cnt[node] = prevc;
#pragma unroll
        for (int j = 2; j < N; ++j) {
            cd = kMaxFloat;
#pragma unroll
            for (int k = 1; k < j; ++k) {
                const float c = cl[k - 1] + cr[j - k - 1];
                if (cd > c) {
                    cd = c;
                    bc = pack_cnt(k, j - k);
                }
            }
            if (!(prev <= cd)) {
                prev = cd;
                prevc = bc;
            }
            st_agent(a.cost + (j - 1) * T + node, prev);
            a.cnt[(j - 1) * T + node] = prevc;
        }
        node = a.parent[node];
    }
}

// The children (binary node, slots) of a wide interior entry (b, budget):
// collapseWide's expansion.  Returns the count; the root keeps its best split
// when its cost prefers a leaf.
template <int N>
HX_DEV int expand(const WideArgs& a, int64_t T, int b, int budget, bool is_root, int* cb, int* ck) {
    if (b >= a.n - 1) {   // the one-reference tree: the root's only child is the leaf
        cb[0] = b;
        ck[0] = 1;
        return 1;
    }
    uint16_t c0 = a.cnt[(budget - 1) * T + b];
    if (c0 == kLeafCnt && is_root) {
        const int l = a.left[b], r = a.right[b];
        float cd = kMaxFloat;
        for (int k = 1; k < N; ++k) {
            const float c = a.cost[(k - 1) * T + l] + a.cost[(N - k - 1) * T + r];
            if (cd > c) {
                cd = c;
                c0 = pack_cnt(k, N - k);
            }
        }
    }
    int m = 2;
    cb[0] = a.left[b];
    ck[0] = cnt_x(c0);
    cb[1] = a.right[b];
    ck[1] = cnt_y(c0);
    while (m < N) {
        int s = -1;
        uint16_t cs = kLeafCnt;
        for (int i = 0; i < m; ++i) {
            const uint16_t cc = a.cnt[(ck[i] - 1) * T + cb[i]];
            const int x = cnt_x(cc), y = cnt_y(cc);
            if (x > 0 && y > 0 && x + y < N) {
                s = i;
                cs = cc;
                break;
            }
        }
        if (s < 0) break;
        for (int i = m; i > s + 1; --i) {
            cb[i] = cb[i - 1];
            ck[i] = ck[i - 1];
        }
        const int sb = cb[s];
        cb[s] = a.left[sb];
        ck[s] = cnt_x(cs);
        cb[s + 1] = a.right[sb];
        ck[s + 1] = cnt_y(cs);
        ++m;
    }
    return m;
}

template <int N>
HX_DEV bool child_is_leaf(const WideArgs& a, int64_t T, int b, int k) {
    return a.cnt[(k - 1) * T + b] == kLeafCnt;
}

// Per entry of a level: (children, interior children) for the level scan.
template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_count(WideArgs a, int64_t lo, int64_t hi, int2* __restrict__ out) {
    const int64_t e = lo + blockIdx.x * (int64_t)kWideBlock + threadIdx.x;
    if (e >= hi) return;
    int2 r = make_int2(0, 0);
    if (!a.e_leaf[e]) {
        const int64_t T = 2 * a.n - 1;
        int cb[N], ck[N];
        const int m = expand<N>(a, T, a.e_b[e], a.e_budget[e], e == 0, cb, ck);
        int ni = 0;
        for (int i = 0; i < m; ++i) ni += child_is_leaf<N>(a, T, cb[i], ck[i]) ? 0 : 1;
        r = make_int2(m, ni);
    }
    out[e - lo] = r;
}

// Exclusive scan of a level's int2 counts in place, total at v[m]: per-block
// sums (kScanPer items per block), one workgroup scans the sums, then each
// block scans its items from its offset.
constexpr int kScanItems = 8;
constexpr int kScanPer = kWideBlock * kScanItems;

HX_DEV int2 add2(int2 p, int2 q) { return make_int2(p.x + q.x, p.y + q.y); }

__global__ __launch_bounds__(kWideBlock) void k_scan_reduce(const int2* __restrict__ v, int64_t m, int2* __restrict__ part) {
    __shared__ int2 wt[kWideBlock / kWave];
    const int64_t base = blockIdx.x * (int64_t)kScanPer;
    int2 s = make_int2(0, 0);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kWideBlock + threadIdx.x;
        if (i < m) s = add2(s, v[i]);
    }
    int2 total;
    block_inclusive_scan<kWideBlock / kWave>(s, add2, wt, total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_parts(int2* __restrict__ part, int64_t nb) {
    __shared__ int2 wt[1024 / kWave];
    __shared__ int2 carry;
    if (threadIdx.x == 0) carry = make_int2(0, 0);
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int2 x = i < nb ? part[i] : make_int2(0, 0);
        int2 total;
        const int2 inc = block_inclusive_scan<1024 / kWave>(x, add2, wt, total);
        const int2 c0 = carry;
        if (i < nb) part[i] = make_int2(c0.x + inc.x - x.x, c0.y + inc.y - x.y);
        __syncthreads();
        if (threadIdx.x == 0) carry = add2(c0, total);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nb] = carry;
}

__global__ __launch_bounds__(kWideBlock) void k_scan_apply(int2* __restrict__ v, int64_t m, const int2* __restrict__ part) {
    __shared__ int2 wt[kWideBlock / kWave];
    const int64_t base = blockIdx.x * (int64_t)kScanPer;
    int2 carry = part[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int64_t i = base + k * kWideBlock + threadIdx.x;
        const int2 x = i < m ? v[i] : make_int2(0, 0);
        int2 total;
        const int2 inc = block_inclusive_scan<kWideBlock / kWave>(x, add2, wt, total);
        if (i < m) v[i] = make_int2(carry.x + inc.x - x.x, carry.y + inc.y - x.y);
        carry = add2(carry, total);
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) v[m] = carry;
}

// Children of the level's interiors written as the next level, breadth-first:
// an interior child gets interior rank int_base + (interior children before it),
// a leaf child leaf rank leaf_base + (leaf children before it).
template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_scatter(WideArgs a, int64_t lo, int64_t hi, const int2* __restrict__ off,
                                                             int64_t int_base, int64_t leaf_base) {
    const int64_t e = lo + blockIdx.x * (int64_t)kWideBlock + threadIdx.x;
    if (e >= hi || a.e_leaf[e]) return;
    const int64_t T = 2 * a.n - 1;
    int cb[N], ck[N];
    const int m = expand<N>(a, T, a.e_b[e], a.e_budget[e], e == 0, cb, ck);
    const int2 o = off[e - lo];
    int64_t slot = hi + o.x;
    int ir = (int)(int_base + o.y);
    int lr = (int)(leaf_base + (o.x - o.y));
    a.e_first[e] = (int32_t)slot;
    a.e_nchild[e] = m;
    for (int i = 0; i < m; ++i, ++slot) {
        const bool leaf = child_is_leaf<N>(a, T, cb[i], ck[i]);
        a.e_b[slot] = cb[i];
        a.e_budget[slot] = (int8_t)ck[i];
        a.e_leaf[slot] = leaf ? 1 : 0;
        a.e_parent[slot] = (int32_t)e;
        a.e_rank[slot] = leaf ? lr++ : ir++;
        a.e_first[slot] = -1;
        a.e_nchild[slot] = 0;
    }
}

template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_root(WideArgs a) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.e_b[0] = root_of(a);
        a.e_budget[0] = 1;
        a.e_leaf[0] = 0;
        a.e_parent[0] = -1;
        a.e_rank[0] = 0;
        a.e_first[0] = -1;
        a.e_nchild[0] = 0;
    }
}

__global__ __launch_bounds__(kWideBlock) void k_wide_zero(int32_t* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)kWideBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWideBlock) p[i] = 0;
}

// Node records (HipBVHNodeN word map): child i's box at words 4i..4i+3
// (min x, max x, min y, max y) and 4(N+i), 4(N+i)+1 (min z, max z), its link
// at 4(N+i)+2 (~index for a leaf); a leaf's begin/end at 4N+2, 4N+3 and its box
// in every slot; child count at 8N-9, size (~size for a leaf) at 8N-5, parent
// at 8N-1; unused slots of an interior repeat child 0's link with the point
// box at child 0's min corner.
template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_emit(WideArgs a, int64_t n_entries, int64_t n_interior) {
    const int64_t e = blockIdx.x * (int64_t)kWideBlock + threadIdx.x;
    if (e >= n_entries) return;
    const bool leaf = a.e_leaf[e];
    const int64_t idx = leaf ? n_interior + a.e_rank[e] : a.e_rank[e];
    int w[8 * N];
#pragma unroll
    for (int i = 0; i < 8 * N; ++i) w[i] = 0;
    auto put_box = [&](int i, float4 mn, float4 mx) {
        w[4 * i + 0] = __float_as_int(mn.x);
        w[4 * i + 1] = __float_as_int(mx.x);
        w[4 * i + 2] = __float_as_int(mn.y);
        w[4 * i + 3] = __float_as_int(mx.y);
        w[4 * (N + i) + 0] = __float_as_int(mn.z);
        w[4 * (N + i) + 1] = __float_as_int(mx.z);
    };
    const int b = a.e_b[e];
    const bool bleaf = b >= a.n - 1;
    const int first = bleaf ? (int)(b - (a.n - 1)) : a.first[b];
    const int last = bleaf ? first : a.last[b];
    int size = last - first + 1;
    int cnt = 0;
    if (leaf) {
        const float4 mn = a.nmin[b], mx = a.nmax[b];
#pragma unroll
        for (int i = 0; i < N; ++i) put_box(i, mn, mx);
        w[4 * N + 2] = first;
        w[4 * N + 3] = last + 1;
        size = ~size;
    } else {
        const int f = a.e_first[e];
        cnt = a.e_nchild[e];
        int link0 = 0;
        float4 mn0 = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int i = 0; i < cnt; ++i) {
            const int c = f + i;
            const int cbn = a.e_b[c];
            const float4 mn = a.nmin[cbn], mx = a.nmax[cbn];
            put_box(i, mn, mx);
            const int link = a.e_leaf[c] ? ~(int)(n_interior + a.e_rank[c]) : a.e_rank[c];
            w[4 * (N + i) + 2] = link;
            if (i == 0) {
                link0 = link;
                mn0 = mn;
            }
        }
        for (int i = cnt; i < N; ++i) {
            put_box(i, mn0, mn0);
            w[4 * (N + i) + 2] = link0;
        }
    }
    w[8 * N - 9] = cnt;
    w[8 * N - 5] = size;
    w[8 * N - 1] = a.e_parent[e] < 0 ? -1 : a.e_rank[a.e_parent[e]];
    int4* o = a.nodes + (int64_t)idx * 2 * N;
#pragma unroll
    for (int q = 0; q < 2 * N; ++q) o[q] = make_int4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// SAH cost (HipBVHKernels.cu:184-238) per node in FP64, block partial sums in
// node order, then one workgroup adds the partials in order.
template <int N>
__global__ __launch_bounds__(kWideBlock) void k_wide_cost(const int4* __restrict__ nodes, int64_t n_nodes, float scene_area,
                                                          float ct, float ci, double* __restrict__ partial) {
    __shared__ double s[kWideBlock];
    const int64_t i = blockIdx.x * (int64_t)kWideBlock + threadIdx.x;
    double v = 0.0;
    if (i < n_nodes) {
        const int* w = reinterpret_cast<const int*>(nodes + i * 2 * N);
        float mn[3] = {kMaxFloat, kMaxFloat, kMaxFloat}, mx[3] = {-kMaxFloat, -kMaxFloat, -kMaxFloat};
        for (int c = 0; c < N; ++c) {
            const float b[6] = {__int_as_float(w[4 * c]), __int_as_float(w[4 * c + 2]), __int_as_float(w[4 * (N + c)]),
                                __int_as_float(w[4 * c + 1]), __int_as_float(w[4 * c + 3]),
                                __int_as_float(w[4 * (N + c) + 1])};
            for (int k = 0; k < 3; ++k) {
                mn[k] = b[k] < mn[k] ? b[k] : mn[k];
                mx[k] = b[3 + k] > mx[k] ? b[3 + k] : mx[k];
            }
        }
        const float P = box_area(make_float3(mn[0], mn[1], mn[2]), make_float3(mx[0], mx[1], mx[2])) / scene_area;
        const int sz = w[8 * N - 5];
        v = sz < 0 ? (double)(ci * P * (float)(div_ceil_i(~sz, N) * N)) : (double)(ct * P);
    }
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = kWideBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_wide_cost_sum(const double* __restrict__ partial, int64_t m, double* __restrict__ out) {
    __shared__ double s[256];
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < m; i += 256) v += partial[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kWideBlock - 1) / kWideBlock); }

template <int N>
int build_wide_n(Context* c, WideArgs a, int64_t* n_interior, int64_t* n_leaves, int64_t* n_entries) {
    const int64_t T = 2 * a.n - 1;
    {
        LaunchScope ls(c, "wide_dp");
        if (a.n > 1) {
            hipLaunchKernelGGL(k_wide_zero, dim3((unsigned)std::min<int64_t>(grid_of(a.n - 1), 4096)), dim3(kWideBlock), 0,
                               c->stream, a.counters, a.n - 1);
            HX_TRY(check_launch("wide_zero"));
        }
        hipLaunchKernelGGL(k_wide_dp<N>, dim3(grid_of(a.n)), dim3(kWideBlock), 0, c->stream, a);
        HX_TRY(check_launch("wide_dp"));
    }
    (void)T;
    LaunchScope ls(c, "wide_levels");
    hipLaunchKernelGGL(k_wide_root<N>, dim3(1), dim3(64), 0, c->stream, a);
    HX_TRY(check_launch("wide_root"));
    int64_t lo = 0, hi = 1, int_total = 1, leaf_total = 0;
    int2 tot;
    while (hi > lo) {
        const int64_t m = hi - lo;
        hipLaunchKernelGGL(k_wide_count<N>, dim3(grid_of(m)), dim3(kWideBlock), 0, c->stream, a, lo, hi, a.scan);
        HX_TRY(check_launch("wide_count"));
        const int64_t nb = (m + kScanPer - 1) / kScanPer;
        hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kWideBlock), 0, c->stream, a.scan, m, a.scan_part);
        HX_TRY(check_launch("wide_scan_reduce"));
        hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(1024), 0, c->stream, a.scan_part, nb);
        HX_TRY(check_launch("wide_scan_parts"));
        hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kWideBlock), 0, c->stream, a.scan, m, a.scan_part);
        HX_TRY(check_launch("wide_scan_apply"));
        HX_HIP(hipMemcpyAsync(&tot, a.scan + m, sizeof(int2), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
        if (tot.x == 0) break;
        if (hi + tot.x > 2 * a.n) {
            set_error("wide collapse: %lld nodes exceed the %lld entries reserved", (long long)(hi + tot.x),
                      (long long)(2 * a.n));
            return HX_ERR_HIP;
        }
        hipLaunchKernelGGL(k_wide_scatter<N>, dim3(grid_of(m)), dim3(kWideBlock), 0, c->stream, a, lo, hi, a.scan,
                           int_total, leaf_total);
        HX_TRY(check_launch("wide_scatter"));
        int_total += tot.y;
        leaf_total += tot.x - tot.y;
        lo = hi;
        hi += tot.x;
    }
    *n_interior = int_total;
    *n_leaves = leaf_total;
    *n_entries = hi;
    hipLaunchKernelGGL(k_wide_emit<N>, dim3(grid_of(hi)), dim3(kWideBlock), 0, c->stream, a, hi, int_total);
    return check_launch("wide_emit");
}

}  // namespace

int launch_wide_build(Context* c, const WideArgs& a, int64_t* n_interior, int64_t* n_leaves, int64_t* n_entries) {
    if (a.N == 4) return build_wide_n<4>(c, a, n_interior, n_leaves, n_entries);
    if (a.N == 8) return build_wide_n<8>(c, a, n_interior, n_leaves, n_entries);
    set_error("wide layout: %d children per node is not supported (4 or 8)", a.N);
    return HX_ERR_INVALID_ARG;
}

int launch_wide_emit(Context* c, const WideArgs& a, int64_t n_entries, int64_t n_interior) {
    LaunchScope ls(c, "wide_emit");
    if (a.N == 4)
        hipLaunchKernelGGL(k_wide_emit<4>, dim3(grid_of(n_entries)), dim3(kWideBlock), 0, c->stream, a, n_entries, n_interior);
    else
        hipLaunchKernelGGL(k_wide_emit<8>, dim3(grid_of(n_entries)), dim3(kWideBlock), 0, c->stream, a, n_entries, n_interior);
    return check_launch("wide_emit");
}

int launch_wide_cost(Context* c, int N, const int4* nodes, int64_t n_nodes, float scene_area, float ct, float ci,
                     double* partial, double* out) {
    const unsigned g = grid_of(n_nodes);
    if (N == 4)
        hipLaunchKernelGGL(k_wide_cost<4>, dim3(g), dim3(kWideBlock), 0, c->stream, nodes, n_nodes, scene_area, ct, ci, partial);
    else
        hipLaunchKernelGGL(k_wide_cost<8>, dim3(g), dim3(kWideBlock), 0, c->stream, nodes, n_nodes, scene_area, ct, ci, partial);
    HX_TRY(check_launch("wide_cost"));
    hipLaunchKernelGGL(k_wide_cost_sum, dim3(1), dim3(256), 0, c->stream, partial, (int64_t)g, out);
    return check_launch("wide_cost_sum");
}

int wide_cost_partials(int64_t n_nodes) { return (int)grid_of(n_nodes); }

}  // namespace hx
