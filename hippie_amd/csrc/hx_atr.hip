// hx_atr.hip -- ATRBVH (agglomerative treelet restructuring) on gfx950.
//
// Reference (src/hippie/rt/bvh/): ATRBuilder.cpp:23-236 (schedule, optimise
// loop, build), ATRBuilderKernels.cu:11-577 (optimize / optimizeSmall and
// their helpers), TRBuilderUtil.h:11-289 (formTreelet and the area helpers).
//
// The reference climbs from every leaf with atomic counters, recomputes every
// node's cost on the way, and lets a 32-lane warp restructure each node whose
// subtree holds at least gamma references.  Here the climb order is replaced by
// a level order: before each iteration the internal nodes are bucketed by
// (subtree height, >= gamma references), and per height one launch recomputes
// the costs of the small nodes (a thread each) and one restructures the large
// ones (32 lanes each).  Nodes of one height are never ancestors of each other
// and each treelet lies inside its root's subtree, so a height's nodes are
// independent, and every parent is processed after its children as in the
// reference.  Heights are kept current: a restructured treelet's interior
// nodes get heights (and sizes) from the merge, and every processed node
// recomputes its own from its final children.
//
// Each treelet is handled by 32 lanes (half a wave64), so the reference's
// warp-32 reductions -- whose tie-breaking depends on the shuffle tree -- are
// reproduced exactly: findLeafToExpand over 2+i lanes and findMinimumDistance
// (per-lane strided minimum over 32 lanes, then a 32-lane tree).  The
// distance matrix lives in LDS; the treelet leaves live in the lanes'
// registers and move by cross-lane permutes.
#include <algorithm>
#include <vector>

#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kG = 32;                         // lanes per treelet
constexpr int kAtrBlock = 256;                 // 8 treelets per workgroup
constexpr int kGroups = kAtrBlock / kG;
constexpr int kMaxT = 30;
constexpr int kTrm = kMaxT * (kMaxT - 1) / 2;  // 435 pairs

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// Lower-triangular matrix without its diagonal (TRBuilderUtil.h:44-49,
// ATRBuilderKernels.cu:11-20; the row comes from a float square root, as in
// the reference).
HX_DEV int trm_index(int row, int col) { return ((row - 1) * row) / 2 + col; }
HX_DEV int trm_row(int index) { return (int)((-1.0f + sqrtf(8.0f * (float)index + 1.0f)) / 2.0f) + 1; }
HX_DEV int trm_col(int index) {
    const int y = trm_row(index);
    return index - (y * (y - 1)) / 2;
}

// TRBuilderUtil.h:100-130 (x*y + x*z + y*z, no validity test).
HX_DEV float atr_area3(float3 mn, float3 mx) {
    const float x = mx.x - mn.x, y = mx.y - mn.y, z = mx.z - mn.z;
    return 2.0f * (x * y + x * z + y * z);
}
HX_DEV float atr_union_area(float3 amn, float3 amx, float3 bmn, float3 bmx) {
    const float x = fmaxf(amx.x, bmx.x) - fminf(amn.x, bmn.x);
    const float y = fmaxf(amx.y, bmx.y) - fminf(amn.y, bmn.y);
    const float z = fmaxf(amx.z, bmx.z) - fminf(amn.z, bmn.z);
    return 2.0f * (x * y + x * z + y * z);
}

template <class T>
HX_DEV T gshfl(T v, int src) { return __shfl(v, src, kG); }
template <class T>
HX_DEV T gshfl_down(T v, int d) { return __shfl_down(v, d, kG); }
HX_DEV float3 gshfl3(float3 v, int src) { return make_float3(gshfl(v.x, src), gshfl(v.y, src), gshfl(v.z, src)); }

// Orders this wave's LDS accesses (no instruction: the wave runs in lockstep).
HX_DEV void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

// formTreelet (TRBuilderUtil.h:171-226): expand the treelet leaf of largest
// area that is an internal node, t-2 times.  Lane i < t ends with treelet leaf
// i in tl, lane i < t-1 with the treelet's internal node i in inode (lane 0:
// the root X).
HX_DEV void form_treelet(const AtrArgs& a, int X, int xl, int xr, int t, int lane, int& tl, int& inode) {
    const int n = (int)a.n;
    tl = -1;
    inode = -1;
    float tl_area = 0.0f;
    if (lane == 0) {
        tl = xl;
        tl_area = a.area[xl];
        inode = X;
    }
    if (lane == 1) {
        tl = xr;
        tl_area = a.area[xr];
    }
    for (int it = 0; it < t - 2; ++it) {
        int idx = lane;
        float val = (tl >= n - 1) ? -kMaxFloat : tl_area;
        for (int k = 2 + it; k > 1; k = (k + 1) / 2) {  // findLeafToExpand
            const int s = k / 2;
            const int oi = gshfl_down(idx, s);
            const float ov = gshfl_down(val, s);
            if (ov > val) {
                val = ov;
                idx = oi;
            }
        }
        const int e = gshfl(idx, 0);
        const int rep = gshfl(tl, e);
        const int cl = a.left[rep], cr = a.right[rep];
        const float al = a.area[cl], ar = a.area[cr];
        if (lane == it + 1) inode = rep;
        if (lane == e) {
            tl = cl;
            tl_area = al;
        }
        if (lane == it + 2) {
            tl = cr;
            tl_area = ar;
        }
    }
}

// One treelet per 32-lane group; task = index into the height bucket.
__global__ __launch_bounds__(kAtrBlock) void k_atr_level(AtrArgs a, const int32_t* __restrict__ list, int count) {
    __shared__ float s_d[kGroups][kTrm];
    const int grp = threadIdx.x / kG;
    const int lane = threadIdx.x & (kG - 1);
    const int task = blockIdx.x * kGroups + grp;
    if (task >= count) return;  // whole 32-lane groups leave together
    float* D = s_d[grp];
    const int n = (int)a.n;
    const int t = a.treelet_size;
    const int X = list[task];

    // The climb step at X (ATRBuilderKernels.cu:293-310): children are final.
    const int xl = a.left[X], xr = a.right[X];
    const float costX = a.ct * a.area[X] + a.cost[xl] + a.cost[xr];
    const int hl = xl >= n - 1 ? 0 : a.height[xl];
    const int hr = xr >= n - 1 ? 0 : a.height[xr];
    const int heightX = 1 + max(hl, hr);

    int tl, inode;
    form_treelet(a, X, xl, xr, t, lane, tl, inode);

    // Treelet leaves: node, box, cost, size, height per lane.
    int node = tl, rnode = -1, refs = 1, hgt = 0, mrefs = 0, mh = 0;
    float sah = 0.0f, mcost = 0.0f;
    float3 mn = make_float3(0.0f, 0.0f, 0.0f), mx = mn;
    if (lane < t) {
        const float4 bmn = a.nmin[tl], bmx = a.nmax[tl];
        mn = make_float3(bmn.x, bmn.y, bmn.z);
        mx = make_float3(bmx.x, bmx.y, bmx.z);
        sah = a.cost[tl];
        if (tl < n - 1) {
            refs = a.refs[tl];
            hgt = a.height[tl];
        }
    }

    // calculateDistancesMatrix (ATRBuilderKernels.cu:22-64): every pair.
    const int pairs = t * (t - 1) / 2;
    for (int base = 0; base < pairs; base += kG) {
        const int e = base + lane;
        const int row = e < pairs ? trm_row(e) : 0;
        const int col = e < pairs ? e - (row * (row - 1)) / 2 : 0;
        const float3 amn = gshfl3(mn, row), amx = gshfl3(mx, row);
        const float3 bmn = gshfl3(mn, col), bmx = gshfl3(mx, col);
        if (e < pairs) D[e] = atr_union_area(amn, amx, bmn, bmx);
    }
    wave_lds_fence();

    for (int lastRow = t - 1; lastRow > 0; --lastRow) {
        // findMinimumDistance (:81-94): strided minimum, then a 32-lane tree.
        const int msize = lastRow * (lastRow + 1) / 2;
        int mi = 0;
        float md = kMaxFloat;
        for (int j = lane; j < msize; j += kG) {
            const float d = D[j];
            if (d < md) {
                md = d;
                mi = j;
            }
        }
        for (int s = kG / 2; s >= 1; s >>= 1) {
            const int oi = gshfl_down(mi, s);
            const float od = gshfl_down(md, s);
            if (od < md) {
                md = od;
                mi = oi;
            }
        }
        const int minIndex = gshfl(mi, 0);
        const int joinCol = trm_col(minIndex), joinRow = trm_row(minIndex);

        // updateState (:96-158).
        const float ssum = gshfl(sah, joinRow) + gshfl(sah, joinCol);
        const int leftIndex = gshfl(node, joinRow), rightIndex = gshfl(node, joinCol);
        float3 amn = gshfl3(mn, joinRow), amx = gshfl3(mx, joinRow);
        const float3 bmn = gshfl3(mn, joinCol), bmx = gshfl3(mx, joinCol);
        amn = make_float3(fminf(bmn.x, amn.x), fminf(bmn.y, amn.y), fminf(bmn.z, amn.z));
        amx = make_float3(fmaxf(bmx.x, amx.x), fmaxf(bmx.y, amx.y), fmaxf(bmx.z, amx.z));
        const int refsNew = gshfl(refs, joinRow) + gshfl(refs, joinCol);
        const int hNew = 1 + max(gshfl(hgt, joinRow), gshfl(hgt, joinCol));
        const int newId = gshfl(inode, lastRow - 1);
        const float c = ssum + a.ct * atr_area3(amn, amx);
        if (lane == joinCol) {
            node = newId;
            mn = amn;
            mx = amx;
            sah = c;
            refs = refsNew;
            hgt = hNew;
        }
        const int lastIndex = gshfl(node, lastRow);
        const float sahLast = gshfl(sah, lastRow);
        const float3 lmn = gshfl3(mn, lastRow), lmx = gshfl3(mx, lastRow);
        const int refsLast = gshfl(refs, lastRow), hLast = gshfl(hgt, lastRow);
        if (lane == joinRow) {
            node = lastIndex;
            sah = sahLast;
            mn = lmn;
            mx = lmx;
            refs = refsLast;
            hgt = hLast;
        }
        if (lane == lastRow) {
            node = leftIndex;
            rnode = rightIndex;
            mn = amn;
            mx = amx;
            mcost = c;
            mrefs = refsNew;
            mh = hNew;
        }

        if (lastRow > 1) {
            // updateDistancesMatrix (:160-196): row lastRow moves to joinRow,
            // then row/column joinCol is recomputed from the merged box.
            if (lane < lastRow && lane != joinRow) {
                const float d = D[trm_index(lastRow, lane)];
                D[trm_index(max(joinRow, lane), min(joinRow, lane))] = d;
            }
            wave_lds_fence();
            const bool upd = lane < lastRow && lane != joinCol;
            const int dr = upd ? max(joinCol, lane) : lane;
            const int dc = upd ? min(joinCol, lane) : lane;
            const float3 pmn = gshfl3(mn, dr), pmx = gshfl3(mx, dr);
            const float3 qmn = gshfl3(mn, dc), qmx = gshfl3(mx, dc);
            const float d = atr_union_area(pmn, pmx, qmn, qmx);
            if (upd) D[trm_index(dr, dc)] = d;
            wave_lds_fence();
        }
    }

    // updateTreelet (:198-228): keep the new treelet if its cost is lower.
    const float newCost = gshfl(mcost, 1);
    const int vid = gshfl(inode, lane >= 1 ? lane - 1 : 0);
    if (newCost < costX) {
        if (lane >= 1 && lane < t) {
            a.left[vid] = node;
            a.right[vid] = rnode;
            a.parent[node] = vid;
            a.parent[rnode] = vid;
            a.cost[vid] = mcost;
            a.nmin[vid] = make_float4(mn.x, mn.y, mn.z, 0.0f);
            a.nmax[vid] = make_float4(mx.x, mx.y, mx.z, 0.0f);
            a.area[vid] = atr_area3(mn, mx);
            a.refs[vid] = mrefs;
            a.height[vid] = mh;
        }
        if (lane == 0) atomicAdd(a.accepted, 1);
    } else if (lane == 0) {
        a.cost[X] = costX;
        a.height[X] = heightX;
    }
}

// The cost climb step for a node that is not restructured (ATRBuilderKernels.cu
// :293-310): every iteration recomputes it, which matters for interior nodes a
// kept treelet gave a cost in merge order.
__global__ __launch_bounds__(kBlock) void k_atr_small(AtrArgs a, const int32_t* __restrict__ list, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= count) return;
    const int v = list[k];
    a.cost[v] = a.ct * a.area[v] + a.cost[a.left[v]] + a.cost[a.right[v]];
}

// ---- TR: optimal treelet restructuring (TRBuilderKernels.cu:11-476) ----
//
// The treelet has 7 leaves (TRBuilder::setTreeletSize accepts 7 only) and is
// formed as ATRBVH's.  Its 128 leaf subsets get their union areas, the leaves
// their current costs, and a DP over the subsets by size gives every subset's
// optimal cost C(S) = min(ct A(S) + min_P [C(P) + C(S\P)], ci A(S) R), R the
// root's reference count.  The partitions are enumerated as the reference
// does, so every split leaves the subset's lowest leaf on the right and ties
// go as there: subsets of 2..5 leaves scan their partitions in one lane
// (first strict minimum), the 6-leaf subsets split theirs over 4 lanes and the
// whole treelet over 32, each reduced by the reference's shuffle-down tree.
constexpr int kTrT = 7;
constexpr int kTrS = 1 << kTrT;

struct TrSchedule {
    uint8_t subset[112];  // subsets of 2..5 leaves, grouped by size
    uint8_t off[5];       // group of size 2 + i: [off[i], off[i+1])
};
constexpr int tr_popc(int v) {
    int c = 0;
    for (; v; v >>= 1) c += v & 1;
    return c;
}
constexpr TrSchedule make_tr_schedule() {
    TrSchedule r{};
    int k = 0;
    for (int size = 2; size <= kTrT - 2; ++size) {
        r.off[size - 2] = (uint8_t)k;
        for (int v = 1; v < kTrS; ++v)
            if (tr_popc(v) == size) r.subset[k++] = (uint8_t)v;
    }
    r.off[kTrT - 3] = (uint8_t)k;
    return r;
}
static_assert(make_tr_schedule().off[4] == 112, "subsets of 2..5 of 7 leaves");
__constant__ TrSchedule kTrSchedule = make_tr_schedule();

// processSchedule's per-subset step (TRBuilderKernels.cu:96-116).
HX_DEV void tr_finish(float* C, uint8_t* M, int s, float opt, int best, float ct, float ci, float R) {
    const float A = C[s];
    C[s] = fminf(ct * A + opt, ci * A * R);
    M[s] = (uint8_t)best;
}

// reduceOptimal (TRBuilderKernels.cu:11-20): shuffle down by values/2 ... 1,
// a lane keeps its own pair unless the other is strictly cheaper.
HX_DEV void tr_reduce(float& opt, int& best, int values) {
    for (int d = values >> 1; d > 0; d >>= 1) {
        const float ov = gshfl_down(opt, d);
        const int ob = gshfl_down(best, d);
        if (ov < opt) {
            opt = ov;
            best = ob;
        }
    }
}

// One treelet per 32-lane group; task = index into the height bucket.
__global__ __launch_bounds__(kAtrBlock) void k_tr_level(AtrArgs a, const int32_t* __restrict__ list, int count) {
    __shared__ float s_cost[kGroups][kTrS];
    __shared__ uint8_t s_mask[kGroups][kTrS];
    __shared__ float s_box[kGroups][kTrT][6];
    __shared__ int s_leaf[kGroups][3][kTrT];       // node, references, height
    __shared__ int s_inode[kGroups][kTrT - 1];
    __shared__ int s_rec[kGroups][6][kTrT - 1];    // node, mask, left, right, cost subset, height
    __shared__ int s_nrec[kGroups];
    const int grp = threadIdx.x / kG;
    const int lane = threadIdx.x & (kG - 1);
    const int task = blockIdx.x * kGroups + grp;
    if (task >= count) return;  // whole 32-lane groups leave together
    float* C = s_cost[grp];
    uint8_t* M = s_mask[grp];
    const int n = (int)a.n;
    const int X = list[task];
    const float ct = a.ct, ci = a.ci;

    // The climb step at X (TRBuilderKernels.cu:391-407): children are final.
    const int xl = a.left[X], xr = a.right[X];
    const float costX = ct * a.area[X] + a.cost[xl] + a.cost[xr];
    const int hl = xl >= n - 1 ? 0 : a.height[xl];
    const int hr = xr >= n - 1 ? 0 : a.height[xr];
    const int heightX = 1 + max(hl, hr);
    const float R = (float)a.refs[X];

    int tl, inode;
    form_treelet(a, X, xl, xr, kTrT, lane, tl, inode);
    float leafCost = 0.0f;
    if (lane < kTrT) {
        const float4 bmn = a.nmin[tl], bmx = a.nmax[tl];
        float* b = s_box[grp][lane];
        b[0] = bmn.x;
        b[1] = bmn.y;
        b[2] = bmn.z;
        b[3] = bmx.x;
        b[4] = bmx.y;
        b[5] = bmx.z;
        s_leaf[grp][0][lane] = tl;
        s_leaf[grp][1][lane] = tl < n - 1 ? a.refs[tl] : 1;
        s_leaf[grp][2][lane] = tl < n - 1 ? a.height[tl] : 0;
        leafCost = a.cost[tl];
    }
    if (lane < kTrT - 1) s_inode[grp][lane] = inode;
    wave_lds_fence();

    // calculateSubsetSurfaceAreas (:22-81): lane l owns subsets 4l .. 4l+3;
    // unions are exact, so their order is free.
    for (int j = 0; j < kTrS / kG; ++j) {
        const int v = (kTrS / kG) * lane + j;
        float3 mn = make_float3(kMaxFloat, kMaxFloat, kMaxFloat);
        float3 mx = make_float3(-kMaxFloat, -kMaxFloat, -kMaxFloat);
        for (int i = 0; i < kTrT; ++i) {
            if (v >> i & 1) {
                const float* b = s_box[grp][i];
                mn = make_float3(fminf(b[0], mn.x), fminf(b[1], mn.y), fminf(b[2], mn.z));
                mx = make_float3(fmaxf(b[3], mx.x), fmaxf(b[4], mx.y), fmaxf(b[5], mx.z));
            }
        }
        C[v] = atr_area3(mn, mx);
    }
    wave_lds_fence();
    if (lane < kTrT) C[1 << lane] = leafCost;  // :433-437
    wave_lds_fence();

    // processSchedule (:84-114): subsets of 2..5 leaves, one size at a time.
#pragma unroll
    for (int g = 0; g < kTrT - 3; ++g) {
        for (int k = kTrSchedule.off[g] + lane; k < kTrSchedule.off[g + 1]; k += kG) {
            const int v = kTrSchedule.subset[k];
            float opt = kMaxFloat;
            int best = 0;
            const int delta = (v - 1) & v;
            int p = (-delta) & v;
            while (p != 0) {
                const float c = C[p] + C[p ^ v];
                if (c < opt) {
                    opt = c;
                    best = p;
                }
                p = (p - delta) & v;
            }
            tr_finish(C, M, v, opt, best, ct, ci, R);
        }
        wave_lds_fence();
    }

    // processSubsets (:116-196): the 7 subsets of 6 leaves over 4 lanes each.
    {
        float opt = kMaxFloat;
        int best = 0;
        const int g = lane >> 2;
        const int v = (kTrS - 1) & ~(1 << g);
        if (lane < 4 * kTrT) {
            const int leftMask = -(1 << max(g - 1, 0));
            int base = (lane & 3) + 1;
            int p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            for (int j = lane & 3; j < (1 << (kTrT - 2)) - 1; j += 4) {
                const float c = C[p] + C[p ^ v];
                if (c < opt) {
                    opt = c;
                    best = p;
                }
                base += 4;
                p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            }
        }
        tr_reduce(opt, best, 4);
        if (lane < 4 * kTrT && (lane & 3) == 0) tr_finish(C, M, v, opt, best, ct, ci, R);
    }
    wave_lds_fence();
    {
        float opt = kMaxFloat;
        int best = 0;
        int p = (lane + 1) * 2;
        for (int j = lane; j < (1 << (kTrT - 1)) - 1; j += kG) {
            const float c = C[p] + C[p ^ (kTrS - 1)];
            if (c < opt) {
                opt = c;
                best = p;
            }
            p += 2 * kG;
        }
        tr_reduce(opt, best, kG);
        if (lane == 0) tr_finish(C, M, kTrS - 1, opt, best, ct, ci, R);
    }
    wave_lds_fence();

    // updateTreelet (:198-324): keep the optimal tree if it is cheaper.  Lane 0
    // lays it out as the reference's stack rounds do: per round, the left
    // sides of the stacked nodes take the next internal nodes in stack order,
    // then the right sides; new internal nodes are stacked in that order.
    const bool keep = C[kTrS - 1] < costX;
    int(*rec)[kTrT - 1] = s_rec[grp];  // rec[field][record]
    if (keep && lane == 0) {
        const int* leafNode = s_leaf[grp][0];
        const int* leafH = s_leaf[grp][2];
        int nrec = 1, next = 1, rb = 0;
        rec[0][0] = X;
        rec[1][0] = kTrS - 1;
        rec[4][0] = 0;
        while (rb < nrec) {
            const int re = nrec;
            for (int side = 0; side < 2; ++side) {
                for (int i = rb; i < re; ++i) {
                    const int part = M[rec[1][i]];
                    const int v = side == 0 ? part : (part ^ rec[1][i]);
                    int child;
                    // A finite kept cost gives every subset on the tree a
                    // split; the bounds only keep non-finite boxes in range.
                    if (__popc(v) > 1 && nrec < kTrT - 1) {
                        child = s_inode[grp][next++];
                        rec[0][nrec] = child;
                        rec[1][nrec] = v;
                        rec[4][nrec] = part;  // the reference's costs[partition], both sides
                        ++nrec;
                    } else {
                        child = leafNode[max(__ffs(v) - 1, 0)];
                    }
                    rec[2 + side][i] = child;
                }
            }
            rb = re;
        }
        s_nrec[grp] = nrec;
        // Heights bottom-up: every record's children come after it.
        for (int i = nrec - 1; i >= 0; --i) {
            const int part = M[rec[1][i]];
            int h[2];
            for (int side = 0; side < 2; ++side) {
                const int v = side == 0 ? part : (part ^ rec[1][i]);
                int hv = 0;
                if (__popc(v) > 1) {
                    for (int k = i + 1; k < nrec; ++k)
                        if (rec[1][k] == v) hv = rec[5][k];
                } else {
                    hv = leafH[max(__ffs(v) - 1, 0)];
                }
                h[side] = hv;
            }
            rec[5][i] = 1 + max(h[0], h[1]);
        }
    }
    wave_lds_fence();
    if (keep && lane < s_nrec[grp]) {
        const int node = rec[0][lane], v = rec[1][lane], l = rec[2][lane], r = rec[3][lane];
        a.left[node] = l;
        a.right[node] = r;
        a.parent[l] = node;
        a.parent[r] = node;
        int refs = 0;
        float3 mn = make_float3(kMaxFloat, kMaxFloat, kMaxFloat);
        float3 mx = make_float3(-kMaxFloat, -kMaxFloat, -kMaxFloat);
        for (int i = 0; i < kTrT; ++i) {
            if (v >> i & 1) {
                const float* b = s_box[grp][i];
                mn = make_float3(fminf(b[0], mn.x), fminf(b[1], mn.y), fminf(b[2], mn.z));
                mx = make_float3(fmaxf(b[3], mx.x), fmaxf(b[4], mx.y), fmaxf(b[5], mx.z));
                refs += s_leaf[grp][1][i];
            }
        }
        a.refs[node] = refs;
        a.height[node] = rec[5][lane];
        if (lane > 0) {  // the root keeps its box and its climb cost
            a.nmin[node] = make_float4(mn.x, mn.y, mn.z, 0.0f);
            a.nmax[node] = make_float4(mx.x, mx.y, mx.z, 0.0f);
            a.area[node] = atr_area3(mn, mx);
            a.cost[node] = C[rec[4][lane]];
        }
    }
    if (lane == 0) {
        a.cost[X] = costX;
        if (keep) atomicAdd(a.accepted, 1);
        else a.height[X] = heightX;
    }
}

// Bucket of an internal node: 2 * height + (references >= gamma).
constexpr int kBuckets = 2 * kAtrMaxLevels;

HX_DEV int bucket_of(const int32_t* refs, const int32_t* height, int v, int gamma) {
    return min(2 * height[v] + (refs[v] >= gamma ? 1 : 0), kBuckets);
}

// Bucket counts: per-tile LDS counts, then one global add per (tile, bucket).
__global__ __launch_bounds__(kBlock) void k_atr_count(const int32_t* __restrict__ refs,
                                                      const int32_t* __restrict__ height, int m, int gamma,
                                                      int32_t* __restrict__ hist) {
    __shared__ int s_cnt[kBuckets + 1];
    for (int i = threadIdx.x; i <= kBuckets; i += kBlock) s_cnt[i] = 0;
    __syncthreads();
    const int64_t base = blockIdx.x * (int64_t)kTile;
    for (int k = 0; k < kItems; ++k) {
        const int64_t v = base + k * kBlock + threadIdx.x;
        if (v < m) atomicAdd(&s_cnt[bucket_of(refs, height, (int)v, gamma)], 1);
    }
    __syncthreads();
    for (int i = threadIdx.x; i <= kBuckets; i += kBlock)
        if (s_cnt[i]) atomicAdd(&hist[i], s_cnt[i]);
}

// Scatter into the buckets (order inside a bucket is free: its nodes are
// independent).  cursor[h] starts at the bucket's offset.
__global__ __launch_bounds__(kBlock) void k_atr_scatter(const int32_t* __restrict__ refs,
                                                        const int32_t* __restrict__ height, int m, int gamma,
                                                        int32_t* __restrict__ cursor, int32_t* __restrict__ list) {
    __shared__ int s_cnt[kBuckets];
    __shared__ int s_base[kBuckets];
    for (int i = threadIdx.x; i < kBuckets; i += kBlock) s_cnt[i] = 0;
    __syncthreads();
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int lvl[kItems], loc[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t v = base + k * kBlock + threadIdx.x;
        lvl[k] = -1;
        if (v < m) {
            lvl[k] = bucket_of(refs, height, (int)v, gamma);
            loc[k] = atomicAdd(&s_cnt[lvl[k]], 1);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kBuckets; i += kBlock)
        if (s_cnt[i]) s_base[i] = atomicAdd(&cursor[i], s_cnt[i]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (lvl[k] >= 0) list[s_base[lvl[k]] + loc[k]] = (int)(base + k * kBlock + threadIdx.x);
}

__global__ void k_atr_zero(int32_t* __restrict__ p, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = 0;
}

// Buckets the internal nodes by (height, >= gamma references); returns the
// per-bucket counts on the host and leaves the nodes in a.list.
int bucket_by_height(Context* c, const AtrArgs& a, int gamma, std::vector<int>& counts) {
    const int m = (int)a.n - 1;
    hipLaunchKernelGGL(k_atr_zero, dim3(4), dim3(kBlock), 0, c->stream, a.hist, kBuckets + 1);
    HX_TRY(check_launch("atr_zero"));
    const unsigned tiles = grid_for(m, kTile);
    hipLaunchKernelGGL(k_atr_count, dim3(tiles), dim3(kBlock), 0, c->stream, a.refs, a.height, m, gamma, a.hist);
    HX_TRY(check_launch("atr_count"));
    counts.assign(kBuckets + 1, 0);
    HX_HIP(hipMemcpyAsync(counts.data(), a.hist, sizeof(int) * (kBuckets + 1), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    if (counts[kBuckets] != 0) {
        set_error("a subtree is deeper than %d levels", kAtrMaxLevels - 1);
        return HX_ERR_UNSUPPORTED;
    }
    std::vector<int> offs(kBuckets, 0);
    int run = 0;
    for (int b = 0; b < kBuckets; ++b) {
        offs[b] = run;
        run += counts[b];
    }
    HX_HIP(hipMemcpyAsync(a.cursor, offs.data(), sizeof(int) * kBuckets, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_atr_scatter, dim3(tiles), dim3(kBlock), 0, c->stream, a.refs, a.height, m, gamma,
                       a.cursor, a.list);
    HX_TRY(check_launch("atr_scatter"));
    // The host vector must outlive the asynchronous upload.
    HX_HIP(hipStreamSynchronize(c->stream));
    return HX_OK;
}

}  // namespace

int launch_relabel_by_height(Context* c, const AtrArgs& a, const PlocArgs& relabel) {
    // Depth-first relabelling, parents first: buckets by height, top down.
    LaunchScope ls(c, "relabel");
    std::vector<int> counts;
    HX_TRY(bucket_by_height(c, a, 0, counts));
    HX_HIP(hipMemcpyAsync(relabel.last, a.refs, sizeof(int32_t) * (size_t)(a.n - 1), hipMemcpyDeviceToDevice,
                          c->stream));
    std::vector<int> offs(kBuckets + 1, 0);
    for (int b = 0; b < kBuckets; ++b) offs[b + 1] = offs[b] + counts[b];
    // Both buckets of a height are contiguous: one step per height.
    std::vector<int2> segs;
    for (int h = kAtrMaxLevels - 1; h >= 0; --h) {
        const int cnt = counts[2 * h] + counts[2 * h + 1];
        if (cnt > 0) segs.push_back(make_int2(offs[2 * h], cnt));
    }
    return launch_tree_down_steps(c, relabel, a.list, segs);
}

namespace {

// ATRBuilder::optimize (ATRBuilder.cpp:80-177) and TRBuilder::optimize
// (TRBuilder.cpp:140-191) share the loop: gamma starts at the treelet size and
// doubles per iteration, stopping once it exceeds the reference count.
int launch_treelets(Context* c, const AtrArgs& a, const PlocArgs& relabel, int64_t* accepted_out, bool tr) {
    const int64_t n = a.n;
    hipLaunchKernelGGL(k_atr_zero, dim3(1), dim3(64), 0, c->stream, a.accepted, 1);
    HX_TRY(check_launch("atr_zero"));
    std::vector<int> counts;
    int64_t gamma = a.treelet_size;
    {
        LaunchScope ls(c, tr ? "tr_optimize" : "atr_optimize");
        for (int i = 0; i < a.iterations; ++i) {
            HX_TRY(bucket_by_height(c, a, (int)std::min<int64_t>(gamma, INT32_MAX), counts));
            int off = 0;
            for (int b = 0; b < kBuckets; ++b) {
                const int cnt = counts[b];
                if (cnt > 0 && (b & 1)) {
                    if (tr) {
                        hipLaunchKernelGGL(k_tr_level, dim3(grid_for(cnt, kGroups)), dim3(kAtrBlock), 0,
                                           c->stream, a, (const int32_t*)(a.list + off), cnt);
                        HX_TRY(check_launch("tr_level"));
                    } else {
                        hipLaunchKernelGGL(k_atr_level, dim3(grid_for(cnt, kGroups)), dim3(kAtrBlock), 0,
                                           c->stream, a, (const int32_t*)(a.list + off), cnt);
                        HX_TRY(check_launch("atr_level"));
                    }
                } else if (cnt > 0 && i > 0) {
                    // The first iteration's small-node costs are the prepared ones.
                    hipLaunchKernelGGL(k_atr_small, dim3(grid_for(cnt, kBlock)), dim3(kBlock), 0, c->stream, a,
                                       (const int32_t*)(a.list + off), cnt);
                    HX_TRY(check_launch("atr_small"));
                }
                off += cnt;
            }
            gamma *= 2;  // ATRBuilder.cpp:163-171
            if (gamma > n) break;
        }
    }
    HX_TRY(launch_relabel_by_height(c, a, relabel));
    if (accepted_out) {
        int32_t acc = 0;
        HX_HIP(hipMemcpyAsync(&acc, a.accepted, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
        *accepted_out = acc;
    }
    return HX_OK;
}

}  // namespace

int launch_atr(Context* c, const AtrArgs& a, const PlocArgs& relabel, int64_t* accepted_out) {
    return launch_treelets(c, a, relabel, accepted_out, false);
}

int launch_tr(Context* c, const AtrArgs& a, const PlocArgs& relabel, int64_t* accepted_out) {
    if (a.treelet_size != kTrT) return HX_ERR_INVALID_ARG;
    return launch_treelets(c, a, relabel, accepted_out, true);
}

}  // namespace hx
