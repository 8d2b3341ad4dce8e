// hx_presplit.hip -- budgeted triangle presplitting on gfx950.
//
// Reference (src/hippie/rt/bvh/): Presplitter.cpp:124-326 (priorities, the
// bisection for D, split tasks), PresplitterKernels.cu:11-419 (the kernels).
// In the reference this GPU variant is compiled out in favour of a CPU SBVH
// (Presplitter.cpp:11); here it is the presplitter.
//
//   priority  per triangle, (2^i (A_box - A_ideal))^(1/3) with i from the
//             highest differing bit of the 63-bit codes of its box corners;
//   D         the largest multiplier (6 bisection rounds) whose rounded budget
//             sum floor(D p) stays within beta * n;
//   split     each triangle becomes floor(D p) + 1 references by recursive
//             clipping at Morton-grid planes.
//
// Determinism: the reference appends references and tasks with atomics, so
// the reference order depends on timing.  Here a triangle's references take
// the slots [o_t, o_t + c_t] (o from a scan of c + 1), and every split hands
// its left child the first lc + 1 slots of its range, so each reference's
// slot is fixed by the split tree alone, whatever order the tasks run in.
// The float sum of the priorities (truncated to int by the reference) is an
// exact 2^-32 fixed-point sum here, and the rounded sums are 64-bit integers.
//
// Batches: with a mesh per triangle, every mesh is presplit as if alone -- its
// own box frames the codes, its own budget beta * n_m and multiplier D_m --
// and the bisection runs for all meshes at once (per-mesh sums each round).
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

inline unsigned grid_for(int64_t n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

struct Sum {
    HX_DEV int operator()(int a, int b) const { return a + b; }
};

// x^y in double precision from the textbook series (the host oracle runs the
// same operations): ln x = e ln2 + 2 atanh((m-1)/(m+1)) with m in
// [sqrt(1/2), sqrt(2)); e^y = 2^k e^r with |r| <= ln2/2 and the Taylor series
// of e^r.
constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLn2Hi = 0.693147180369123816490;
constexpr double kLn2Lo = 1.90821492927058770002e-10;

HX_DEV double series_ln(double x) {
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double t = 1.0 / 25.0;
    for (int k = 23; k >= 1; k -= 2) t = t * s2 + 1.0 / (double)k;
    return (double)e * kLn2 + 2.0 * s * t;
}

HX_DEV double series_exp(double y) {
    const double k = rint(y / kLn2);
    const double r = (y - k * kLn2Hi) - k * kLn2Lo;
    double p = 1.0;
    for (int j = 20; j >= 1; --j) p = 1.0 + p * r / (double)j;
    return ldexp(p, (int)k);
}

HX_DEV float series_pow(float x, float y) {
    if (!(x > 0.0f) || !(x <= kMaxFloat)) return x == 0.0f ? 0.0f : x;
    return (float)series_exp((double)y * series_ln((double)x));
}

HX_DEV float ax(float3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
HX_DEV void set_ax(float3& v, int a, float f) {
    if (a == 0) v.x = f;
    else if (a == 1) v.y = f;
    else v.z = f;
}
HX_DEV int ax_i(int3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// 21 bits per axis, axis 0 in bit 0 (PresplitterKernels.cu:44-55).
HX_DEV uint64_t code63(int3 q) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 21; ++i) {
        c |= (uint64_t)((q.x >> i) & 1) << (3 * i + 0);
        c |= (uint64_t)((q.y >> i) & 1) << (3 * i + 1);
        c |= (uint64_t)((q.z >> i) & 1) << (3 * i + 2);
    }
    return c;
}

HX_DEV int diff_bit(uint64_t a, uint64_t b) {
    const uint64_t d = a ^ b;
    return d ? 63 - __clzll((long long)d) : -1;
}

struct SceneFrame {
    float3 mn, dg, scale;
};

HX_DEV SceneFrame frame_of(const float* sb) {
    SceneFrame s;
    s.mn = make_float3(sb[0], sb[1], sb[2]);
    s.dg = make_float3(sb[3] - sb[0], sb[4] - sb[1], sb[5] - sb[2]);
    s.scale = make_float3(1.0f / s.dg.x, 1.0f / s.dg.y, 1.0f / s.dg.z);
    return s;
}

HX_DEV void tri_verts(const PsArgs& a, int t, float3 v[3]) {
    const int32_t* tr = a.tris + 3 * (int64_t)t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float* p = a.verts + 3 * (int64_t)tr[k];
        v[k] = make_float3(p[0], p[1], p[2]);
    }
}

HX_DEV Box3 tri_bounds(const float3 v[3]) {
    Box3 b = box_empty();
    box_grow(b, v[0]);
    box_grow(b, v[1]);
    box_grow(b, v[2]);
    return b;
}

HX_DEV int mesh_at(const PsArgs& a, int t) { return a.mesh_of ? a.mesh_of[t] : 0; }
HX_DEV const float* frame_box(const PsArgs& a, int m) { return a.mesh_of ? a.mesh_box + 6 * (int64_t)m : a.box; }

// Per-mesh sums acc[2 * m + k] accumulated by a wave over its iterations:
// while the wave's live lanes share the mesh of the running sum it grows in
// registers, and it goes out with one device-scope atomic when the mesh
// changes or the wave ends (device atomics on one word serialise, so a
// per-wave atomic over 16M references would take milliseconds).  Waves
// whose lanes span meshes add per lane.
struct MeshSum {
    unsigned long long* acc;
    int k;
    int m = -1;
    unsigned long long s = 0;
    HX_DEV void flush() {
        if (m >= 0 && s && lane_id() == 0) atomicAdd(acc + 2 * (int64_t)m + k, s);
        s = 0;
    }
    HX_DEV void add(bool live, int mesh, unsigned long long v) {
        const uint64_t act = ballot64(live);
        if (!act) return;
        const int m0 = __shfl(mesh, __builtin_ctzll(act), kWave);
        if (__all(!live || mesh == m0)) {
            if (!live) v = 0;
#pragma unroll
            for (int sh = 1; sh < kWave; sh <<= 1) {
                const uint32_t lo = __shfl_xor((uint32_t)v, sh, kWave);
                const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), sh, kWave);
                v += ((unsigned long long)hi << 32) | lo;
            }
            if (m0 != m) {
                flush();
                m = m0;
            }
            s += v;
        } else if (live && v) {
            atomicAdd(acc + 2 * (int64_t)mesh + k, v);
        }
    }
};

// Each workgroup takes kSumChunk consecutive references.
constexpr int kSumChunk = 64 * kBlock;

// computePriorities (:11-71).
HX_DEV float priority(const float3 v[3], const SceneFrame& S) {
    const Box3 box = tri_bounds(v);
    const int3 qmn = make_int3(f2i_sat(((box.mn.x - S.mn.x) * S.scale.x) * 2097151.9f),
                               f2i_sat(((box.mn.y - S.mn.y) * S.scale.y) * 2097151.9f),
                               f2i_sat(((box.mn.z - S.mn.z) * S.scale.z) * 2097151.9f));
    const int3 qmx = make_int3(f2i_sat(((box.mx.x - S.mn.x) * S.scale.x) * 2097151.9f),
                               f2i_sat(((box.mx.y - S.mn.y) * S.scale.y) * 2097151.9f),
                               f2i_sat(((box.mx.z - S.mn.z) * S.scale.z) * 2097151.9f));
    const int diff = diff_bit(code63(qmn), code63(qmx));
    const int i = diff / 3 - 21;
    const float3 d1 = make_float3(v[0].x - v[1].x, v[0].y - v[1].y, v[0].z - v[1].z);
    const float3 d2 = make_float3(v[0].x - v[2].x, v[0].y - v[2].y, v[0].z - v[2].z);
    const float cx = d1.y * d2.z - d1.z * d2.y, cy = d1.z * d2.x - d1.x * d2.z, cz = d1.x * d2.y - d1.y * d2.x;
    const float ideal = fabsf(cx) + fabsf(cy) + fabsf(cz);
    return series_pow(ldexpf(1.0f, i) * (box_area(box.mn, box.mx) - ideal), 1.0f / 3.0f);
}

// Priorities and their exact fixed-point sums per mesh (macc[2m]).
__global__ __launch_bounds__(kBlock) void k_ps_priority(PsArgs a) {
    MeshSum sum{a.macc, 0};
    const int64_t end = min((int64_t)(blockIdx.x + 1) * kSumChunk, a.n);
    for (int64_t t = blockIdx.x * (int64_t)kSumChunk + threadIdx.x; t < (int64_t)(blockIdx.x + 1) * kSumChunk;
         t += kBlock) {
        const bool live = t < end;
        unsigned long long s = 0;
        int m = 0;
        if (live) {
            m = mesh_at(a, (int)t);
            const SceneFrame S = frame_of(frame_box(a, m));
            float3 v[3];
            tri_verts(a, (int)t, v);
            const float p = priority(v, S);
            a.prio[t] = p;
            if (p >= 0.0f && p <= 1.0e9f) s = (unsigned long long)((double)p * 4294967296.0);
        }
        sum.add(live, m, s);
    }
    sum.flush();
}

// sumPrioritiesRound (:116-148): per-mesh sums of floor(D_m p) (macc[2m+1]).
__global__ __launch_bounds__(kBlock) void k_ps_round_sum(PsArgs a) {
    MeshSum sum{a.macc, 1};
    const int64_t end = min((int64_t)(blockIdx.x + 1) * kSumChunk, a.n);
    for (int64_t t = blockIdx.x * (int64_t)kSumChunk + threadIdx.x; t < (int64_t)(blockIdx.x + 1) * kSumChunk;
         t += kBlock) {
        const bool live = t < end;
        unsigned long long s = 0;
        int m = 0;
        if (live) {
            m = mesh_at(a, (int)t);
            s = (unsigned long long)(long long)f2i_sat(a.Dm[m] * a.prio[t]);
        }
        sum.add(live, m, s);
    }
    sum.flush();
}

// Per-tile sums of floor(D p) + 1.
__global__ __launch_bounds__(kBlock) void k_ps_tile_sums(PsArgs a) {
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int s = 0;
    for (int k = 0; k < kItems; ++k) {
        const int64_t t = base + k * kBlock + threadIdx.x;
        if (t < a.n) s += f2i_sat(a.Dm[mesh_at(a, (int)t)] * a.prio[t]) + 1;
    }
    s = wave_sum_i32(s);
    __shared__ int ws[kWavesPerBlock];
    if (lane_id() == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < kWavesPerBlock; ++w) tot += ws[w];
        a.partial[blockIdx.x] = tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_ps_scan(int32_t* __restrict__ partial, int count) {
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
}

HX_DEV void write_ref(const PsArgs& a, int off, int tri, const Box3& b) {
    a.ref_min[off] = make_float4(b.mn.x, b.mn.y, b.mn.z, 0.0f);
    a.ref_max[off] = make_float4(b.mx.x, b.mx.y, b.mx.z, 0.0f);
    a.ref_tri[off] = tri;
}

// Append a task to queue q (wave-aggregated slot reservation; the order of
// the queue does not matter, every task carries its output slots).
HX_DEV void push_task(const PsArgs& a, int q, bool live, int tri, int count, int off, const Box3& b) {
    const uint64_t m = ballot64(live);
    if (!m) return;
    const int leader = __builtin_ctzll(m);
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(a.qcount + q, __builtin_popcountll(m));
    base = __shfl(base, leader, kWave);
    if (live) {
        const int slot = base + count_below(m);
        if (slot >= a.queue_cap) return;  // reported by the host from the count
        a.task[q][slot] = make_int4(tri, count, off, 0);
        a.tmn[q][slot] = make_float4(b.mn.x, b.mn.y, b.mn.z, 0.0f);
        a.tmx[q][slot] = make_float4(b.mx.x, b.mx.y, b.mx.z, 0.0f);
    }
}

// initSplitTasks (:150-187): slot offsets from the tile scan; an unsplit
// triangle is written as its own reference at once, the others are queued.
__global__ __launch_bounds__(kBlock) void k_ps_init(PsArgs a) {
    __shared__ int wt[kWavesPerBlock];
    const int64_t base = blockIdx.x * (int64_t)kTile;
    int carry = a.partial[blockIdx.x];
    for (int k = 0; k < kItems; ++k) {
        const int64_t t = base + k * kBlock + threadIdx.x;
        const bool in = t < a.n;
        const int cnt = in ? f2i_sat(a.Dm[mesh_at(a, (int)t)] * a.prio[t]) : 0;
        const int w = in ? cnt + 1 : 0;
        int tot;
        const int inc = block_inclusive_scan<kWavesPerBlock>(w, Sum(), wt, tot);
        const int off = carry + inc - w;
        carry += tot;
        Box3 b = box_empty();
        if (in) {
            float3 v[3];
            tri_verts(a, (int)t, v);
            b = tri_bounds(v);
            if (cnt == 0) write_ref(a, off, (int)t, b);
        }
        push_task(a, 0, in && cnt > 0, (int)t, cnt, off, b);
    }
}

// split (:189-419) of every queued task into two children; a child with no
// budget left is written as a reference, the others go to the next queue.
__global__ __launch_bounds__(kBlock) void k_ps_split(PsArgs a, int q, int count) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const bool live = k < count;
    int tri = 0, lc = 0, rc = 0, loff = 0, roff = 0;
    Box3 lb = box_empty(), rb = box_empty();
    if (live) {
        const int4 tk = a.task[q][k];
        const float4 bmn4 = a.tmn[q][k], bmx4 = a.tmx[q][k];
        Box3 box;
        box.mn = make_float3(bmn4.x, bmn4.y, bmn4.z);
        box.mx = make_float3(bmx4.x, bmx4.y, bmx4.z);
        tri = tk.x;
        const SceneFrame S = frame_of(frame_box(a, mesh_at(a, tri)));
        const int3 qmn = make_int3(f2i_sat(((box.mn.x - S.mn.x) / S.dg.x) * 2097151.9f),
                                   f2i_sat(((box.mn.y - S.mn.y) / S.dg.y) * 2097151.9f),
                                   f2i_sat(((box.mn.z - S.mn.z) / S.dg.z) * 2097151.9f));
        const int3 qmx = make_int3(f2i_sat(((box.mx.x - S.mn.x) / S.dg.x) * 2097151.9f),
                                   f2i_sat(((box.mx.y - S.mn.y) / S.dg.y) * 2097151.9f),
                                   f2i_sat(((box.mx.z - S.mn.z) / S.dg.z) * 2097151.9f));
        const int diff = diff_bit(code63(qmn), code63(qmx));
        int axis;
        float pos;
        if (diff >= 0) {
            axis = diff % 3;
            pos = 0.0f;
            float delta = 0.5f;
            const int qa = ax_i(qmx, axis);
            for (int i = 20; i >= diff / 3; --i) {
                if ((qa >> i) & 1) pos += delta;
                delta *= 0.5f;
            }
            pos = pos * ax(S.dg, axis) + ax(S.mn, axis);
            if (ax(box.mn, axis) >= pos || ax(box.mx, axis) <= pos) pos = (ax(box.mn, axis) + ax(box.mx, axis)) * 0.5f;
        } else {
            const float dx = box.mx.x - box.mn.x, dy = box.mx.y - box.mn.y, dz = box.mx.z - box.mn.z;
            axis = (dx > dy && dx > dz) ? 0 : (dy > dz ? 1 : 2);
            pos = (ax(box.mn, axis) + ax(box.mx, axis)) * 0.5f;
        }
        float3 v[3];
        tri_verts(a, tri, v);
        // Edges (v2,v0), (v0,v1), (v1,v2): each start vertex goes to the side(s)
        // it lies on, and a crossing edge adds its plane point to both.
        int j1 = 2;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int j0 = j1;
            j1 = i;
            const float3 p0 = v[j0], p1 = v[j1];
            const float pa = ax(p0, axis), pb = ax(p1, axis);
            if (pa <= pos) box_grow(lb, p0);
            if (pa >= pos) box_grow(rb, p0);
            if ((pa < pos && pb > pos) || (pa > pos && pb < pos)) {
                const float tt = maxs(mins((pos - pa) / (pb - pa), 1.0f), 0.0f);
                const float u = 1.0f - tt;
                const float3 m = make_float3(p0.x * u + p1.x * tt, p0.y * u + p1.y * tt, p0.z * u + p1.z * tt);
                box_grow(lb, m);
                box_grow(rb, m);
            }
        }
        set_ax(lb.mx, axis, pos);
        set_ax(rb.mn, axis, pos);
        lb.mn = make_float3(maxs(lb.mn.x, box.mn.x), maxs(lb.mn.y, box.mn.y), maxs(lb.mn.z, box.mn.z));
        lb.mx = make_float3(mins(lb.mx.x, box.mx.x), mins(lb.mx.y, box.mx.y), mins(lb.mx.z, box.mx.z));
        rb.mn = make_float3(maxs(rb.mn.x, box.mn.x), maxs(rb.mn.y, box.mn.y), maxs(rb.mn.z, box.mn.z));
        rb.mx = make_float3(mins(rb.mx.x, box.mx.x), mins(rb.mx.y, box.mx.y), mins(rb.mx.z, box.mx.z));
        // AABB::longest (util/AABB.h:38) returns bool: 1 unless the box is a point.
        const float ll = fmaxf(fmaxf(lb.mx.x - lb.mn.x, lb.mx.y - lb.mn.y), lb.mx.z - lb.mn.z) != 0.0f ? 1.0f : 0.0f;
        const float rl = fmaxf(fmaxf(rb.mx.x - rb.mn.x, rb.mx.y - rb.mn.y), rb.mx.z - rb.mn.z) != 0.0f ? 1.0f : 0.0f;
        lc = f2i_sat(((float)(tk.y - 1) * ll) / (ll + rl) + 0.5f);
        rc = tk.y - 1 - lc;
        loff = tk.z;
        roff = tk.z + lc + 1;
        if (lc <= 0) write_ref(a, loff, tri, lb);
        if (rc <= 0) write_ref(a, roff, tri, rb);
    }
    push_task(a, q ^ 1, live && lc > 0, tri, lc, loff, lb);
    push_task(a, q ^ 1, live && rc > 0, tri, rc, roff, rb);
}

__global__ void k_ps_zero_u64(unsigned long long* p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void k_ps_zero_q(int32_t* qcount, int q) {
    if (threadIdx.x == 0) qcount[q] = 0;
}

__global__ void k_ps_zero_both(int32_t* qcount) {
    if (threadIdx.x < 2) qcount[threadIdx.x] = 0;
}

// Per-mesh round sums of floor(D_m p) for the multipliers `D` (M entries).
unsigned sum_blocks(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kSumChunk - 1) / kSumChunk); }

int round_sums(Context* c, const PsArgs& a, int M, const std::vector<float>& D, std::vector<int64_t>& out) {
    HX_HIP(hipMemcpyAsync(a.Dm, D.data(), sizeof(float) * (size_t)M, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_ps_zero_u64, dim3(std::min<unsigned>(grid_for(2 * (int64_t)M), 1024u)), dim3(kBlock), 0,
                       c->stream, a.macc, 2 * (int64_t)M);
    HX_TRY(check_launch("ps_zero"));
    hipLaunchKernelGGL(k_ps_round_sum, dim3(sum_blocks(a.n)), dim3(kBlock), 0, c->stream, a);
    HX_TRY(check_launch("ps_round_sum"));
    std::vector<unsigned long long> h(2 * (size_t)M);
    HX_HIP(hipMemcpyAsync(h.data(), a.macc, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    out.resize(M);
    for (int m = 0; m < M; ++m) out[m] = (int64_t)h[2 * (size_t)m + 1];
    return HX_OK;
}

__global__ __launch_bounds__(kBlock) void k_refs_to_tris(int32_t* __restrict__ ids, const int32_t* __restrict__ ref_tri,
                                                         int64_t n) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p < n) ids[p] = ref_tri[ids[p]];
}

}  // namespace

int launch_refs_to_tris(Context* c, int32_t* ids, const int32_t* ref_tri, int64_t n) {
    hipLaunchKernelGGL(k_refs_to_tris, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, ids, ref_tri, n);
    return check_launch("refs_to_tris");
}

int launch_presplit_count(Context* c, const PsArgs& a, float beta, int M, const int64_t* tri_counts,
                          int64_t* ref_counts, float* D_out) {
    LaunchScope ls(c, "presplit_count");
    hipLaunchKernelGGL(k_ps_zero_u64, dim3(std::min<unsigned>(grid_for(2 * (int64_t)M), 1024u)), dim3(kBlock), 0,
                       c->stream, a.macc, 2 * (int64_t)M);
    HX_TRY(check_launch("ps_zero"));
    hipLaunchKernelGGL(k_ps_priority, dim3(sum_blocks(a.n)), dim3(kBlock), 0, c->stream, a);
    HX_TRY(check_launch("ps_priority"));
    std::vector<unsigned long long> h(2 * (size_t)M);
    HX_HIP(hipMemcpyAsync(h.data(), a.macc, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                          c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    // computeD (Presplitter.cpp:124-155) per mesh, Smax = (int)(beta * n_m);
    // a mesh whose priority sum floors to 0 gets D = 0 (no splits).
    std::vector<int64_t> smax(M), s1(M), s;
    std::vector<float> dmin(M, 0.0f), dmax(M, 0.0f), cur(M, 0.0f);
    std::vector<char> live(M, 0);
    bool any = false;
    for (int m = 0; m < M; ++m) {
        smax[m] = f2i_host(beta * (float)tri_counts[m]);
        s1[m] = (int64_t)(h[2 * (size_t)m] >> 32);
        live[m] = s1[m] > 0;
        any = any || live[m];
        if (live[m]) dmin[m] = (float)(smax[m] / s1[m]);
    }
    if (any) {
        HX_TRY(round_sums(c, a, M, dmin, s));
        for (int m = 0; m < M; ++m)
            if (live[m]) dmax[m] = (dmin[m] * (float)smax[m]) / (float)s[m];
        for (int r = 0; r < 6; ++r) {
            for (int m = 0; m < M; ++m) cur[m] = live[m] ? 0.5f * (dmin[m] + dmax[m]) : 0.0f;
            HX_TRY(round_sums(c, a, M, cur, s));
            for (int m = 0; m < M; ++m) {
                if (!live[m]) continue;
                if (s[m] > smax[m]) dmax[m] = cur[m];
                else dmin[m] = cur[m];
            }
        }
    }
    HX_TRY(round_sums(c, a, M, dmin, s));
    for (int m = 0; m < M; ++m) {
        ref_counts[m] = tri_counts[m] + s[m];
        D_out[m] = dmin[m];
    }
    return HX_OK;
}

int launch_presplit_refs(Context* c, const PsArgs& a, int64_t n_refs) {
    LaunchScope ls(c, "presplit_split");
    const int64_t n = a.n;
    const int ntiles = (int)((n + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_ps_zero_both, dim3(1), dim3(64), 0, c->stream, a.qcount);
    HX_TRY(check_launch("ps_zero"));
    hipLaunchKernelGGL(k_ps_tile_sums, dim3(ntiles), dim3(kBlock), 0, c->stream, a);
    HX_TRY(check_launch("ps_tile_sums"));
    hipLaunchKernelGGL(k_ps_scan, dim3(1), dim3(kBlock), 0, c->stream, a.partial, ntiles);
    HX_TRY(check_launch("ps_scan"));
    hipLaunchKernelGGL(k_ps_init, dim3(ntiles), dim3(kBlock), 0, c->stream, a);
    HX_TRY(check_launch("ps_init"));
    int q = 0;
    int32_t count = 0;
    HX_HIP(hipMemcpyAsync(&count, a.qcount + q, sizeof(count), hipMemcpyDeviceToHost, c->stream));
    HX_HIP(hipStreamSynchronize(c->stream));
    // Every split consumes one unit of budget, so the rounds end; the bound
    // guards against a miscount.
    for (int64_t round = 0; count > 0; ++round) {
        if ((int64_t)count > a.queue_cap || round > n_refs) {
            set_error("presplit: split queue overflow (%d tasks, capacity %lld)", count, (long long)a.queue_cap);
            return HX_ERR_HIP;
        }
        hipLaunchKernelGGL(k_ps_zero_q, dim3(1), dim3(64), 0, c->stream, a.qcount, q ^ 1);
        HX_TRY(check_launch("ps_zero_q"));
        hipLaunchKernelGGL(k_ps_split, dim3(grid_for(count)), dim3(kBlock), 0, c->stream, a, q, (int)count);
        HX_TRY(check_launch("ps_split"));
        q ^= 1;
        HX_HIP(hipMemcpyAsync(&count, a.qcount + q, sizeof(count), hipMemcpyDeviceToHost, c->stream));
        HX_HIP(hipStreamSynchronize(c->stream));
    }
    return HX_OK;
}

}  // namespace hx
