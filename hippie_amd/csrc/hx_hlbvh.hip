// hx_hlbvh.hip -- HLBVH on gfx950: the Karras LBVH below clusters of
// references sharing the top sah_bits Morton bits, a binned SAH split above.
//
// Reference (src/hippie/rt/bvh/): HLBVHBuilder.cpp:61-437, HLBVHBuilderKernels
// .cu:35-704, HLBVHBin.h:14-44.
//
// A run of equal Morton prefixes in sorted order is exactly one Karras
// subtree, so clusters are found by a scan over the sorted keys instead of
// the reference's state climb, invalidation climb and compaction.  The upper
// levels are built breadth first, one launch per level and one workgroup per
// task: the task's clusters are binned on the fixed prefix grid (LDS bins,
// order-preserving integer min/max), the split is the first strict SAH
// minimum over (axis, bin) as in the reference, and the task's cluster list
// is partitioned stably, so every task owns a contiguous list in Morton order.
// Upper nodes reuse the ids of the Karras nodes that span clusters; the root
// stays 0.  Each upper node carries its depth-first leaf offset, which makes
// the final relabelling of the lower Karras subtrees a shift per cluster.
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
constexpr int kMaxBins = 32;

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

struct Sum {
    HX_DEV int operator()(int a, int b) const { return a + b; }
};

HX_DEV uint32_t ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HX_DEV float unord(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

template <class K>
HX_DEV uint32_t prefix_at(const void* keys, int64_t p, int shift) {
    return (uint32_t)(static_cast<const K*>(keys)[p] >> shift);
}

// Run starts of the Morton prefix, counted per 4096-position tile.
template <class K>
__global__ __launch_bounds__(kBlock) void k_hl_start_counts(const void* __restrict__ keys, int n, int shift,
                                                            int32_t* __restrict__ partial) {
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int cnt = 0;
    for (int k = 0; k < kItems; ++k) {
        const int64_t p = base + k * kBlock + threadIdx.x;
        if (p < n) cnt += (p == 0 || prefix_at<K>(keys, p, shift) != prefix_at<K>(keys, p - 1, shift)) ? 1 : 0;
    }
    cnt = wave_sum_i32(cnt);
    __shared__ int ws[kWavesPerBlock];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) s += ws[w];
        partial[blockIdx.x] = s;
    }
}

// Single block: exclusive scan in place, total to *total.
__global__ __launch_bounds__(kBlock) void k_hl_scan(int32_t* __restrict__ partial, int count,
                                                    int32_t* __restrict__ total) {
    __shared__ int wt[kWavesPerBlock];
    int carry = 0;
    for (int base = 0; base < count; base += kBlock) {
        const int t = base + threadIdx.x;
        const int v = t < count ? partial[t] : 0;
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(v, Sum(), wt, tot);
        if (t < count) partial[t] = carry + inc - v;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// Cluster of every position; run starts and ends per cluster.
template <class K>
__global__ __launch_bounds__(kBlock) void k_hl_clusters(HlbvhArgs a, int shift, const int32_t* __restrict__ partial) {
    __shared__ int wt[kWavesPerBlock];
    const int n = (int)a.n;
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int carry = partial[blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const int64_t p = base + k * kBlock + threadIdx.x;
        int st = 0, en = 0;
        if (p < n) {
            const uint32_t me = prefix_at<K>(a.keys, p, shift);
            st = (p == 0 || me != prefix_at<K>(a.keys, p - 1, shift)) ? 1 : 0;
            en = (p == n - 1 || me != prefix_at<K>(a.keys, p + 1, shift)) ? 1 : 0;
        }
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(st, Sum(), wt, tot);
        if (p < n) {
            const int c = carry + inc - 1;
            a.cl_of[p] = c;
            if (st) a.ca[c] = (int)p;
            if (en) a.cb[c] = (int)p;
        }
        carry += tot;
    }
}

// Cluster roots (the Karras node, or leaf, whose range is the run) and the
// per-tile counts of the upper Karras nodes (ranges spanning clusters).  A
// lower node records its cluster as split = ~cluster: the upper levels
// rewrite first/last of the upper nodes, so the final shift cannot tell the
// two kinds apart by their ranges any more (every upper node gets a split >=
// 0 from its partition).
__global__ __launch_bounds__(kBlock) void k_hl_roots(HlbvhArgs a, int32_t* __restrict__ partial) {
    const int n = (int)a.n;
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int cnt = 0;
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kBlock + threadIdx.x;
        if (i < n) {
            const int c = a.cl_of[i];
            if (a.ca[c] == i && a.cb[c] == i) a.croot[c] = n - 1 + (int)i;
        }
        if (i < n - 1) {
            const int f = a.first[i], l = a.last[i];
            const int c = a.cl_of[f];
            if (a.cl_of[l] != c) {
                ++cnt;
            } else {
                a.split[i] = ~c;
                if (a.ca[c] == f && a.cb[c] == l) a.croot[c] = (int)i;
            }
        }
    }
    cnt = wave_sum_i32(cnt);
    __shared__ int ws[kWavesPerBlock];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) s += ws[w];
        partial[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_hl_upper_ids(HlbvhArgs a, const int32_t* __restrict__ partial) {
    __shared__ int wt[kWavesPerBlock];
    const int n = (int)a.n;
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int carry = partial[blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + k * kBlock + threadIdx.x;
        const int f = i < n - 1 ? (a.cl_of[a.first[i]] != a.cl_of[a.last[i]] ? 1 : 0) : 0;
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(f, Sum(), wt, tot);
        if (f) a.upper[carry + inc - 1] = (int)i;
        carry += tot;
    }
}

// Bin coordinates (mortonCodeToCoords, HLBVHBuilderKernels.cu:10-19), box and
// size of every cluster; the member list starts in Morton order.
template <class K>
__global__ __launch_bounds__(kBlock) void k_hl_cluster_init(HlbvhArgs a, int shift, int C) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const uint32_t code = prefix_at<K>(a.keys, a.ca[c], shift);
    int co[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) co[ax] += ((code >> (3 * i + ax)) & 1) << i;
    a.coord[c] = make_int4(co[0], co[1], co[2], a.cb[c] - a.ca[c] + 1);
    a.members[0][c] = c;
}

struct Bin {
    float3 mn, mx;
    int count;
};

HX_DEV void bin_include(Bin& b, const Bin& o) {
    b.mn = make_float3(fminf(b.mn.x, o.mn.x), fminf(b.mn.y, o.mn.y), fminf(b.mn.z, o.mn.z));
    b.mx = make_float3(fmaxf(b.mx.x, o.mx.x), fmaxf(b.mx.y, o.mx.y), fmaxf(b.mx.z, o.mx.z));
    b.count += o.count;
}
// HLBVHBin::area / cost (HLBVHBin.h:34-43).
HX_DEV float bin_cost(const Bin& b) {
    const float dx = b.mx.x - b.mn.x, dy = b.mx.y - b.mn.y, dz = b.mx.z - b.mn.z;
    return (float)b.count * (2.0f * (dx * dy + dx * dz + dy * dz));
}

// One workgroup per task: bin the task's clusters, choose the split (split,
// :395-497), write the upper node's box, and count the new interior children.
__global__ __launch_bounds__(kBlock) void k_hl_decide(HlbvhArgs a, int src, const int4* __restrict__ tasks,
                                                      int4* __restrict__ decision, int32_t* __restrict__ child_counts) {
    __shared__ uint32_t s_mn[3][kMaxBins][3], s_mx[3][kMaxBins][3];
    __shared__ int s_cnt[3][kMaxBins];
    __shared__ float s_best[3];
    __shared__ int s_bin[3];
    __shared__ Bin s_l[3], s_r[3], s_all;
    const int NB = a.nbins;
    const int4 tk = tasks[blockIdx.x];  // k, s, e, dfs offset
    for (int i = threadIdx.x; i < 3 * kMaxBins; i += kBlock) {
        const int ax = i / kMaxBins, b = i % kMaxBins;
        for (int k = 0; k < 3; ++k) {  // empty bin: (MAX_FLOAT, -MAX_FLOAT), as HLBVHBuilderKernels.cu:341-342
            s_mn[ax][b][k] = ord(kMaxFloat);
            s_mx[ax][b][k] = ord(-kMaxFloat);
        }
        s_cnt[ax][b] = 0;
    }
    __syncthreads();
    const int32_t* mem = a.members[src];
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) {
        const int c = mem[i];
        const int root = a.croot[c];
        const float4 bmn = a.nmin[root], bmx = a.nmax[root];
        const int4 co = a.coord[c];
        const int cc[3] = {co.x, co.y, co.z};
        const uint32_t lo[3] = {ord(bmn.x), ord(bmn.y), ord(bmn.z)};
        const uint32_t hi[3] = {ord(bmx.x), ord(bmx.y), ord(bmx.z)};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int b = min(cc[ax], NB - 1);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicMin(&s_mn[ax][b][k], lo[k]);
                atomicMax(&s_mx[ax][b][k], hi[k]);
            }
            atomicAdd(&s_cnt[ax][b], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int ax = threadIdx.x;
        auto load = [&](int b) {
            Bin r;
            r.mn = make_float3(unord(s_mn[ax][b][0]), unord(s_mn[ax][b][1]), unord(s_mn[ax][b][2]));
            r.mx = make_float3(unord(s_mx[ax][b][0]), unord(s_mx[ax][b][1]), unord(s_mx[ax][b][2]));
            r.count = s_cnt[ax][b];
            return r;
        };
        Bin right[kMaxBins];
        right[NB - 1] = load(NB - 1);
        for (int b = NB - 2; b >= 0; --b) {
            right[b] = right[b + 1];
            bin_include(right[b], load(b));
        }
        Bin acc = load(0);
        float best = kMaxFloat;
        int bb = -1;
        Bin lb = acc, rb = acc;
        for (int b = 0; b < NB - 1; ++b) {
            if (acc.count > 0 && right[b + 1].count > 0) {
                const float cst = bin_cost(acc) + bin_cost(right[b + 1]);
                if (cst < best) {
                    best = cst;
                    bb = b;
                    lb = acc;
                    rb = right[b + 1];
                }
            }
            bin_include(acc, load(b + 1));
        }
        s_best[ax] = best;
        s_bin[ax] = bb;
        s_l[ax] = lb;
        s_r[ax] = rb;
        if (ax == 2) s_all = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float best = kMaxFloat;
        int axis = -1;
        for (int ax = 0; ax < 3; ++ax)
            if (s_best[ax] < best) {
                best = s_best[ax];
                axis = ax;
            }
        const int cnt = tk.z - tk.y;
        Bin lb, rb;
        int nl, bin;
        if (axis < 0) {  // every cluster in one bin on every axis: median
            lb = s_all;
            rb = s_all;
            nl = cnt / 2;
            bin = -1;
        } else {
            lb = s_l[axis];
            rb = s_r[axis];
            nl = lb.count;
            bin = s_bin[axis];
        }
        Box3 bx = box_empty();
        box_grow(bx, lb.mn);
        box_grow(bx, lb.mx);
        box_grow(bx, rb.mn);
        box_grow(bx, rb.mx);
        const int v = a.upper[tk.x];
        a.nmin[v] = make_float4(bx.mn.x, bx.mn.y, bx.mn.z, 0.0f);
        a.nmax[v] = make_float4(bx.mx.x, bx.mx.y, bx.mx.z, 0.0f);
        decision[blockIdx.x] = make_int4(axis, bin, nl, 0);
        child_counts[blockIdx.x] = (nl > 1 ? 1 : 0) + (cnt - nl > 1 ? 1 : 0);
    }
}

// One workgroup per task: stable partition of the task's clusters, the upper
// node's links and depth-first ranges, its children (clusters or new tasks).
__global__ __launch_bounds__(kBlock) void k_hl_partition(HlbvhArgs a, int src, const int4* __restrict__ tasks,
                                                         const int4* __restrict__ decision,
                                                         const int32_t* __restrict__ child_base, int next_k,
                                                         int4* __restrict__ next_tasks) {
    __shared__ int wt[kWavesPerBlock];
    __shared__ int s_left_size;
    const int n = (int)a.n;
    const int4 tk = tasks[blockIdx.x];
    const int4 dc = decision[blockIdx.x];
    const int cnt = tk.z - tk.y, nl = dc.z;
    const int32_t* mem = a.members[src];
    int32_t* out = a.members[src ^ 1];
    if (threadIdx.x == 0) s_left_size = 0;
    __syncthreads();
    int carry = 0, lsize = 0;
    for (int base = 0; base < cnt; base += kBlock) {
        const int j = base + threadIdx.x;
        int c = -1, f = 0;
        if (j < cnt) {
            c = mem[tk.y + j];
            if (dc.x < 0) {
                f = j < nl ? 1 : 0;
            } else {
                const int4 co = a.coord[c];
                const int cc = dc.x == 0 ? co.x : (dc.x == 1 ? co.y : co.z);
                f = min(cc, a.nbins - 1) <= dc.y ? 1 : 0;
            }
        }
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(f, Sum(), wt, tot);
        if (j < cnt) {
            const int lrank = carry + inc - f;           // lefts before j
            const int dst = f ? lrank : nl + (j - lrank);  // stable
            out[tk.y + dst] = c;
            if (f) lsize += a.coord[c].w;
        }
        carry += tot;
    }
    lsize = wave_sum_i32(lsize);
    if (lane_id() == 0) atomicAdd(&s_left_size, lsize);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int sl = s_left_size;
    const int v = a.upper[tk.x];
    const int off = tk.w;
    const int size = a.task_size[tk.x];
    a.first[v] = off;
    a.last[v] = off + size - 1;
    a.split[v] = off + sl - 1;
    if (tk.x == 0) a.parent[v] = -1;
    const int base_k = next_k + child_base[blockIdx.x];
    int made = 0;
    int child[2];
    const int rs[2] = {tk.y, tk.y + nl}, re[2] = {tk.y + nl, tk.z};
    const int offs[2] = {off, off + sl}, sizes[2] = {sl, size - sl};
    for (int side = 0; side < 2; ++side) {
        if (re[side] - rs[side] == 1) {
            const int c = out[rs[side]];
            a.cl_off[c] = offs[side];
            a.cl_par[c] = v;
            const int r = a.croot[c];
            if (r >= n - 1) {
                child[side] = n - 1 + offs[side];  // a one-leaf cluster: its leaf's final id
            } else {
                child[side] = r;
                a.parent[r] = v;
            }
        } else {
            const int k = base_k + made++;
            a.task_size[k] = sizes[side];
            a.parent[a.upper[k]] = v;
            child[side] = a.upper[k];
            next_tasks[child_base[blockIdx.x] + made - 1] = make_int4(k, rs[side], re[side], offs[side]);
        }
    }
    a.left[v] = child[0];
    a.right[v] = child[1];
}

// Lower Karras nodes and leaves: shift their ranges to their cluster's
// depth-first offset.  Pass 0 reads every leaf's parent (a lower node, or the
// upper node above a one-leaf cluster) and fixes the lower nodes; pass 1
// writes the leaves at their final ids.
__global__ __launch_bounds__(kBlock) void k_hl_lower(HlbvhArgs a, int pass) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int n = (int)a.n;
    if (pass == 0) {
        if (i < n) {
            const int c = a.cl_of[i];
            a.tmp[i] = (a.ca[c] == i && a.cb[c] == i) ? a.cl_par[c] : a.parent[n - 1 + i];
        }
        if (i < n - 1) {
            const int sp = a.split[i];
            if (sp >= 0) return;  // upper node: already placed
            const int c = ~sp;
            const int f = a.first[i], l = a.last[i];
            const int sh = a.cl_off[c] - a.ca[c];
            const int lc = a.left[i], rc = a.right[i];
            const int gamma = lc >= n - 1 ? lc - (n - 1) : lc;
            a.first[i] = f + sh;
            a.last[i] = l + sh;
            a.split[i] = gamma + sh;
            if (lc >= n - 1) a.left[i] = lc + sh;
            if (rc >= n - 1) a.right[i] = rc + sh;
        }
    } else if (i < n) {
        const int c = a.cl_of[i];
        const int rank = i + a.cl_off[c] - a.ca[c];
        const int leaf = n - 1 + rank;
        a.tri_out[rank] = a.sorted_tri[i];
        a.parent[leaf] = a.tmp[i];
        a.first[leaf] = rank;
        a.last[leaf] = rank;
        a.left[leaf] = rank;
        a.right[leaf] = rank + 1;
    }
}

template <class K>
int run_hlbvh_t(Context* c, const HlbvhArgs& a0, int* clusters_out) {
    HlbvhArgs a = a0;
    const int n = (int)a.n;
    const int shift = a.bits - a.sah_bits;
    const int ntiles = (n + kTile - 1) / kTile;
    HX_TRY(c->scan_partials.reserve(sizeof(int32_t) * (size_t)(ntiles + 64)));
    int32_t* partial = c->scan_partials.as<int32_t>();
    int32_t* total = partial + ntiles;
    int C = 0;
    {
        LaunchScope ls(c, "hlbvh_clusters");
        hipLaunchKernelGGL(k_hl_start_counts<K>, dim3(ntiles), dim3(kBlock), 0, c->stream, a.keys, n, shift, partial);
        HX_TRY(check_launch("hl_start_counts"));
        hipLaunchKernelGGL(k_hl_scan, dim3(1), dim3(kBlock), 0, c->stream, partial, ntiles, total);
        HX_TRY(check_launch("hl_scan"));
        hipLaunchKernelGGL(k_hl_clusters<K>, dim3(ntiles), dim3(kBlock), 0, c->stream, a, shift, partial);
        HX_TRY(check_launch("hl_clusters"));
        HX_HIP(hipMemcpyAsync(&C, total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
    }
    *clusters_out = C;
    if (C <= 1) return HX_OK;  // one cluster: the tree is the LBVH
    {
        LaunchScope ls(c, "hlbvh_roots");
        hipLaunchKernelGGL(k_hl_roots, dim3(ntiles), dim3(kBlock), 0, c->stream, a, partial);
        HX_TRY(check_launch("hl_roots"));
        hipLaunchKernelGGL(k_hl_scan, dim3(1), dim3(kBlock), 0, c->stream, partial, ntiles, total);
        HX_TRY(check_launch("hl_scan"));
        hipLaunchKernelGGL(k_hl_upper_ids, dim3(ntiles), dim3(kBlock), 0, c->stream, a, partial);
        HX_TRY(check_launch("hl_upper_ids"));
        hipLaunchKernelGGL(k_hl_cluster_init<K>, dim3(grid_for(C)), dim3(kBlock), 0, c->stream, a, shift, C);
        HX_TRY(check_launch("hl_cluster_init"));
    }
    // Upper levels, breadth first.  Task arrays hold at most C - 1 entries.
    int4* tasks[2] = {a.tasks[0], a.tasks[1]};
    const int4 root = make_int4(0, 0, C, 0);
    HX_HIP(hipMemcpyAsync(tasks[0], &root, sizeof(int4), hipMemcpyHostToDevice, c->stream));
    HX_HIP(hipMemcpyAsync(a.task_size, &n, sizeof(int), hipMemcpyHostToDevice, c->stream));
    int count = 1, next_k = 1, src = 0, t = 0;
    {
        LaunchScope ls(c, "hlbvh_upper");
        while (count > 0) {
            hipLaunchKernelGGL(k_hl_decide, dim3(count), dim3(kBlock), 0, c->stream, a, src, tasks[t], a.decision,
                               a.child_counts);
            HX_TRY(check_launch("hl_decide"));
            hipLaunchKernelGGL(k_hl_scan, dim3(1), dim3(kBlock), 0, c->stream, a.child_counts, count, total);
            HX_TRY(check_launch("hl_scan"));
            hipLaunchKernelGGL(k_hl_partition, dim3(count), dim3(kBlock), 0, c->stream, a, src, tasks[t], a.decision,
                               a.child_counts, next_k, tasks[t ^ 1]);
            HX_TRY(check_launch("hl_partition"));
            int made = 0;
            HX_HIP(hipMemcpyAsync(&made, total, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HX_HIP(hipStreamSynchronize(c->stream));
            next_k += made;
            if (next_k > C - 1 + 1) {
                set_error("HLBVH: %d upper nodes for %d clusters", next_k, C);
                return HX_ERR_HIP;
            }
            count = made;
            src ^= 1;
            t ^= 1;
        }
    }
    {
        LaunchScope ls(c, "hlbvh_lower");
        hipLaunchKernelGGL(k_hl_lower, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, 0);
        HX_TRY(check_launch("hl_lower0"));
        hipLaunchKernelGGL(k_hl_lower, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a, 1);
        HX_TRY(check_launch("hl_lower1"));
    }
    return HX_OK;
}

}  // namespace

int launch_hlbvh(Context* c, const HlbvhArgs& a, bool keys64, int* clusters_out) {
    return keys64 ? run_hlbvh_t<uint64_t>(c, a, clusters_out) : run_hlbvh_t<uint32_t>(c, a, clusters_out);
}

}  // namespace hx
