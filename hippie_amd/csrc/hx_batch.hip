// hx_batch.hip -- the pieces a batched build adds to the single-mesh path.
//
// Many independent meshes are built as one: triangles of mesh m occupy
// positions [off[m], off[m+1]) of the input, every triangle is quantised in
// its own mesh's box, the references are sorted by (mesh, Morton) with two
// stable radix sorts (Morton digits, then mesh digits), and one Karras
// construction over the composite key gives every mesh its own subtree (the
// separator nodes above the mesh roots are never emitted).  No reference
// counterpart: the reference builds one scene per HipBVH.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Float <-> order-preserving unsigned key, so min/max are integer atomics.
HX_DEV uint32_t ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
HX_DEV float unord(uint32_t k) { return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k); }

// mesh_of[t] = m for t in [off[m], off[m+1]): one workgroup per mesh writes
// its own range (offsets start at 0 and increase strictly, so the ranges tile
// [0, n)).
__global__ __launch_bounds__(kBlock) void k_mesh_of(const int32_t* __restrict__ off, int32_t* __restrict__ mesh_of) {
    const int m = blockIdx.x;
    const int end = off[m + 1];
    for (int t = off[m] + (int)threadIdx.x; t < end; t += kBlock) mesh_of[t] = m;
}

__global__ __launch_bounds__(kBlock) void k_mesh_box_init(uint32_t* __restrict__ bits, int M) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= 6 * M) return;
    bits[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
}

// Exact per-mesh min/max of the vertices of its triangles.  A workgroup
// takes kBoxChunk consecutive triangles.  When they all belong to one mesh
// (every workgroup inside a mesh of more than kBoxChunk triangles) it reduces
// them in registers, across the wave and across its waves in LDS, and issues
// six atomics; otherwise each wave whose 64 triangles of a round share a
// mesh reduces in registers and issues six, and a wave at a mesh boundary
// one set per lane.  (One set per wave for every 64 triangles serialised on
// the few addresses of a large mesh: 9.2 ms for 128M triangles.)
constexpr int kBoxItems = 16;
constexpr int kBoxChunk = kBlock * kBoxItems;

HX_DEV void tri_bounds(const float* verts, const int32_t* tris, int64_t t, float r[6]) {
    const int32_t* tr = tris + 3 * t;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float* v = verts + 3 * (int64_t)tr[k];
        r[0] = fminf(r[0], v[0]);
        r[1] = fminf(r[1], v[1]);
        r[2] = fminf(r[2], v[2]);
        r[3] = fmaxf(r[3], v[0]);
        r[4] = fmaxf(r[4], v[1]);
        r[5] = fmaxf(r[5], v[2]);
    }
}

HX_DEV void wave_box_reduce(float r[6]) {
#pragma unroll
    for (int s = 1; s < kWave; s <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], __shfl_xor(r[k], s, kWave));
#pragma unroll
        for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], __shfl_xor(r[k], s, kWave));
    }
}

HX_DEV void box_atomics(uint32_t* bits, int m, const float r[6]) {
    for (int k = 0; k < 3; ++k) atomicMin(bits + 6 * m + k, ord(r[k]));
    for (int k = 3; k < 6; ++k) atomicMax(bits + 6 * m + k, ord(r[k]));
}

__global__ __launch_bounds__(kBlock) void k_mesh_box(const float* __restrict__ verts, const int32_t* __restrict__ tris,
                                                     int n, const int32_t* __restrict__ mesh_of,
                                                     uint32_t* __restrict__ bits) {
    __shared__ float s_box[kBlock / kWave][6];
    const int64_t start = (int64_t)blockIdx.x * kBoxChunk;
    const int64_t end = min((int64_t)n, start + kBoxChunk);
    const int m_first = mesh_of[start], m_last = mesh_of[end - 1];
    if (m_first == m_last) {
        float r[6] = {kMaxFloat, kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat, -kMaxFloat};
#pragma unroll 4
        for (int i = 0; i < kBoxItems; ++i) {
            const int64_t t = start + i * kBlock + threadIdx.x;
            if (t < end) tri_bounds(verts, tris, t, r);
        }
        wave_box_reduce(r);
        const int w = threadIdx.x / kWave;
        if (lane_id() == 0)
            for (int k = 0; k < 6; ++k) s_box[w][k] = r[k];
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int ww = 1; ww < kBlock / kWave; ++ww) {
                for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], s_box[ww][k]);
                for (int k = 3; k < 6; ++k) r[k] = fmaxf(r[k], s_box[ww][k]);
            }
            box_atomics(bits, m_first, r);
        }
        return;
    }
    for (int i = 0; i < kBoxItems; ++i) {
        const int64_t t = start + i * kBlock + threadIdx.x;
        const bool live = t < end;
        int m = -1;
        float r[6] = {kMaxFloat, kMaxFloat, kMaxFloat, -kMaxFloat, -kMaxFloat, -kMaxFloat};
        if (live) {
            m = mesh_of[t];
            tri_bounds(verts, tris, t, r);
        }
        const int m0 = __shfl(m, 0, kWave);
        const bool uniform = __all(m == m0 || !live) && m0 >= 0;
        if (uniform) {
            wave_box_reduce(r);
            if (lane_id() == 0) box_atomics(bits, m0, r);
        } else if (live) {
            box_atomics(bits, m, r);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_mesh_box_final(const uint32_t* __restrict__ bits, int M,
                                                           float* __restrict__ boxes) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < 6 * M) boxes[i] = unord(bits[i]);
}

// keys[p] = the mesh of slot vals[p]: the last m with off[m] <= vals[p], by
// binary search over the offsets (staged in LDS when they fit) instead of a
// random read of a per-slot mesh map (2.4 ms for 128M slots).
constexpr int kMeshKeyLds = 8192;

template <bool LDS>
__global__ __launch_bounds__(kBlock) void k_gather_mesh_key(const int32_t* __restrict__ vals,
                                                            const int32_t* __restrict__ off, int M, int n,
                                                            uint32_t* __restrict__ keys) {
    __shared__ int32_t s_off[LDS ? kMeshKeyLds : 1];
    const int32_t* o = off;
    if (LDS) {
        for (int i = threadIdx.x; i < M; i += kBlock) s_off[i] = off[i];
        __syncthreads();
        o = s_off;
    }
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= n) return;
    const int t = vals[p];
    int lo = 0, hi = M - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (o[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    keys[p] = (uint32_t)lo;
}

__global__ __launch_bounds__(kBlock) void k_gather_key(const int32_t* __restrict__ vals,
                                                       const uint64_t* __restrict__ src, int n,
                                                       uint64_t* __restrict__ dst) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) dst[p] = src[vals[p]];
}

__global__ __launch_bounds__(kBlock) void k_tri_local(const int32_t* __restrict__ sorted_tri,
                                                      const int32_t* __restrict__ mesh_of,
                                                      const int32_t* __restrict__ off, int n,
                                                      int32_t* __restrict__ tri_local) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < n) tri_local[p] = sorted_tri[p] - off[mesh_of[p]];
}

}  // namespace

int launch_mesh_of(Context* c, const int32_t* off, int M, int64_t n, int32_t* mesh_of) {
    LaunchScope ls(c, "batch_mesh_of");
    if (M <= 0 || n <= 0) return HX_OK;
    hipLaunchKernelGGL(k_mesh_of, dim3((unsigned)M), dim3(kBlock), 0, c->stream, off, mesh_of);
    return check_launch("mesh_of");
}

int launch_mesh_boxes(Context* c, const float* verts, const int32_t* tris, int64_t n, const int32_t* mesh_of, int M,
                      uint32_t* bits, float* boxes) {
    LaunchScope ls(c, "batch_mesh_boxes");
    hipLaunchKernelGGL(k_mesh_box_init, dim3(grid_for(6 * (int64_t)M)), dim3(kBlock), 0, c->stream, bits, M);
    HX_TRY(check_launch("mesh_box_init"));
    if (n > 0)
        hipLaunchKernelGGL(k_mesh_box, dim3((unsigned)((n + kBoxChunk - 1) / kBoxChunk)), dim3(kBlock), 0, c->stream,
                           verts, tris, (int)n, mesh_of, bits);
    HX_TRY(check_launch("mesh_box"));
    hipLaunchKernelGGL(k_mesh_box_final, dim3(grid_for(6 * (int64_t)M)), dim3(kBlock), 0, c->stream, bits, M, boxes);
    return check_launch("mesh_box_final");
}

int launch_gather_mesh_keys(Context* c, const int32_t* vals, const int32_t* off, int M, int64_t n, uint32_t* keys) {
    if (n <= 0) return HX_OK;
    if (M <= kMeshKeyLds)
        hipLaunchKernelGGL(k_gather_mesh_key<true>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, vals, off, M, (int)n,
                           keys);
    else
        hipLaunchKernelGGL(k_gather_mesh_key<false>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, vals, off, M, (int)n,
                           keys);
    return check_launch("gather_mesh_key");
}

int launch_gather_keys(Context* c, const int32_t* vals, const uint64_t* src, int64_t n, uint64_t* dst) {
    hipLaunchKernelGGL(k_gather_key, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, vals, src, (int)n, dst);
    return check_launch("gather_key");
}

int launch_tri_local(Context* c, const int32_t* sorted_tri, const int32_t* mesh_of, const int32_t* off, int64_t n,
                     int32_t* tri_local) {
    hipLaunchKernelGGL(k_tri_local, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, sorted_tri, mesh_of, off, (int)n,
                       tri_local);
    return check_launch("tri_local");
}

}  // namespace hx
