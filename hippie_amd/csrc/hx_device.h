// hx_device.h -- device-side building blocks shared by the gfx950 kernels.
//
// Everything here is written for 64-lane wavefronts: ballots are 64-bit,
// lane masks come from mbcnt, and no code assumes a 32-wide warp.
// Arithmetic that feeds bit-exact outputs (boxes, Morton codes, Woop rows,
// traversal) is written as explicit selects / fmaf so that -ffp-contract=off
// builds give the same bits as the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HX_DEV __device__ __forceinline__

namespace hx {

constexpr int kWave = 64;
constexpr float kMaxFloat = 3.402823466e+38f;  // src/hippie/Globals.h:70

// src/hippie/Globals.h:73-74: MIN/MAX are selects (not fminf/fmaxf).
HX_DEV float mins(float a, float b) { return (b < a) ? b : a; }
HX_DEV float maxs(float a, float b) { return (b > a) ? b : a; }

struct Box3 {
    float3 mn, mx;
};

HX_DEV Box3 box_empty() {
    Box3 b;
    b.mn = make_float3(kMaxFloat, kMaxFloat, kMaxFloat);
    b.mx = make_float3(-kMaxFloat, -kMaxFloat, -kMaxFloat);
    return b;
}

// util/AABB.h:26-29 grow(point)
HX_DEV void box_grow(Box3& b, float3 p) {
    b.mn.x = mins(b.mn.x, p.x); b.mn.y = mins(b.mn.y, p.y); b.mn.z = mins(b.mn.z, p.z);
    b.mx.x = maxs(b.mx.x, p.x); b.mx.y = maxs(b.mx.y, p.y); b.mx.z = maxs(b.mx.z, p.z);
}

// util/AABB.h:36,40 area() with the valid() guard.
HX_DEV float box_area(float3 mn, float3 mx) {
    if (!(mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z)) return 0.0f;
    float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Saturating float->uint, NaN -> 0 (the semantics of the hardware convert).
HX_DEV uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

HX_DEV int32_t f2i_sat(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 0x7FFFFFFF;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}

// 10-bit interleave (3 x 10 -> 30 bits).  Same function as the magic-number
// spread of rt/HipUtil.h:35-49, computed with a lookup-free bit deposit.
HX_DEV uint32_t spread10(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

HX_DEV uint32_t morton30(uint32_t x, uint32_t y, uint32_t z) {
    return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}

// 60-bit code of a normalised centroid: rt/HipUtil.h:72-82,100-106.
HX_DEV uint64_t morton60(float qx, float qy, float qz) {
    const float scale = 1048575.0f;  // (1u << 20) - 1
    uint32_t x = f2u_sat(qx * scale);
    uint32_t y = f2u_sat(qy * scale);
    uint32_t z = f2u_sat(qz * scale);
    uint64_t lo = morton30(x & 1023u, y & 1023u, z & 1023u);
    uint64_t hi = morton30(x >> 10u, y >> 10u, z >> 10u);
    return (hi << 30) | lo;
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
// ---------------------------------------------------------------------------

HX_DEV int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

HX_DEV uint64_t ballot64(bool p) { return __ballot(p); }

// Number of set bits of m strictly below this lane.
HX_DEV int count_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <class T>
HX_DEV T wave_shfl_xor(T v, int mask) {
    return __shfl_xor(v, mask, kWave);
}

HX_DEV int wave_sum_i32(int v) {
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// DPP move of every 32-bit word of v (GFX9 controls: row_shr:n = 0x110 + n,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143).  Lanes of rows outside
// ROW_MASK, or whose source lies outside the pattern, keep their own value;
// callers combine only where the source is valid.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143;
template <int CTRL, int ROW_MASK, class T>
HX_DEV T dpp_move(T v) {
    static_assert(sizeof(T) % 4 == 0, "word-sized types only");
    int w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) w[k] = __builtin_amdgcn_update_dpp(w[k], w[k], CTRL, ROW_MASK, 0xF, false);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
}

// Inclusive scan over the 64 lanes of a wave, all in DPP moves (no LDS
// traffic): Hillis-Steele within each row of 16 lanes, then the row totals
// carried by row_bcast:15 (rows 1 and 3 take lanes 15 and 47) and
// row_bcast:31 (rows 2 and 3 take lane 31).  Op must be associative.
template <class T, class Op>
HX_DEV T wave_inclusive_scan(T v, Op op) {
    const int lane = lane_id();
    const int r = lane & 15;
    T o = dpp_move<kDppRowShr1, 0xF>(v);
    if (r >= 1) v = op(o, v);
    o = dpp_move<kDppRowShr2, 0xF>(v);
    if (r >= 2) v = op(o, v);
    o = dpp_move<kDppRowShr4, 0xF>(v);
    if (r >= 4) v = op(o, v);
    o = dpp_move<kDppRowShr8, 0xF>(v);
    if (r >= 8) v = op(o, v);
    o = dpp_move<kDppRowBcast15, 0xA>(v);
    if (lane & 16) v = op(o, v);
    o = dpp_move<kDppRowBcast31, 0xC>(v);
    if (lane >= 32) v = op(o, v);
    return v;
}

// Inclusive scan over a block of NW waves; every thread of the block must
// call it (it synchronises twice).  wave_tot: NW entries of shared memory.
template <int NW, class T, class Op>
HX_DEV T block_inclusive_scan(T v, Op op, T* wave_tot, T& block_total) {
    const int w = threadIdx.x / kWave;
    v = wave_inclusive_scan(v, op);
    if (lane_id() == kWave - 1) wave_tot[w] = v;
    __syncthreads();
    T pre{};
    bool have = false;
    T tot = wave_tot[0];
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        if (i < w) {
            pre = have ? op(pre, wave_tot[i]) : wave_tot[i];
            have = true;
        }
        if (i > 0) tot = op(tot, wave_tot[i]);
    }
    __syncthreads();
    block_total = tot;
    return have ? op(pre, v) : v;
}

// Device-scope atomics with explicit ordering.
HX_DEV int atomic_add_acq_rel(int* p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
HX_DEV uint32_t atomic_load_relaxed(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
HX_DEV void atomic_store_relaxed(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace hx
