// hx_sort.hip -- stable LSD radix sort of (key, int32 value) pairs for gfx950.
//
// Replaces src/radix_sort/ (RadixSort.cpp:181-238: 5-bit digits, three
// launches and a host sync per pass).  Here: RB-bit digits (8 or 10, chosen by
// the "sort_digit_bits" tunable: 10-bit digits sort 60-bit keys in 6 passes
// instead of 8), one histogram pass for all digit positions, then one
// single-pass "onesweep" kernel per digit:
//   - tiles of 4096 keys, 4 waves of 64 lanes, 16 keys per lane in
//     wave-striped order (item i of lane l = i*64 + l within the wave slice);
//   - ranking by 64-bit ballot match: lanes with the same digit form a peer
//     mask from nbits ballots; rank = running wave count + peers below;
//   - decoupled look-back over per-tile digit counts gives each tile its
//     global digit offsets; tiles take their index from an atomic counter so a
//     tile only waits on tiles that are already running;
//   - keys and values are staged in LDS in digit order and written out in
//     contiguous runs.
// The order is stable, so with identity input values the result equals a
// stable sort by key -- the semantics the reference's callers rely on.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kMaxPasses = 8;    // 64-bit keys at 8-bit digits; 7 at 10-bit

constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1u;

inline int sort_passes(int end_bit, int rb) { return (end_bit + rb - 1) / rb; }
inline int64_t sort_tiles(int64_t n, int items) { return (n + kSortBlock * items - 1) / (kSortBlock * items); }

// `gate` (may be NULL) switches a launch off on the device: the kernels of a
// sort's fallback passes are always enqueued and return at once unless the
// local pass set *gate, so the host never waits to choose between them.
HX_DEV bool gated_off(const int32_t* gate) { return gate != nullptr && *gate == 0; }

__global__ void k_zero_u32(uint32_t* p, int64_t n, const int32_t* gate) {
    if (gated_off(gate)) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_iota(int32_t* p, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int32_t)i;
}

template <class K>
HX_DEV uint32_t digit_of(K key, int shift, uint32_t mask) {
    return (uint32_t)(key >> shift) & mask;
}

// Histograms of every digit position in one read of the keys.  One
// 1024-thread workgroup per CU, four keys in flight per thread: the LDS
// histograms are merged into the global one with device-scope atomics,
// which are expensive, so few workgroups merge.
constexpr int kHistBlock = 1024;
constexpr int kHistUnroll = 4;

template <class K, int RB>
__global__ __launch_bounds__(kHistBlock) void k_sort_hist(const K* __restrict__ keys, int64_t n, int lo_bit,
                                                          int passes, int end_bit, uint32_t* __restrict__ hist,
                                                          const int32_t* gate) {
    if (gated_off(gate)) return;
    constexpr int R = 1 << RB;
    constexpr int P = (8 * (int)sizeof(K) + RB - 1) / RB;
    __shared__ uint32_t h[P][R];
    for (int i = threadIdx.x; i < P * R; i += kHistBlock) (&h[0][0])[i] = 0u;
    __syncthreads();
    // Morton keys in input order share their high digits across a wave, and
    // 64 lanes adding to one LDS word serialise: a pass whose digit is the
    // same in every live lane adds the lane count once; otherwise every lane
    // adds its own.
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * kHistBlock;
    const int64_t n_round = (n + kWave - 1) / kWave * kWave;  // whole waves agree on i < n_round
    for (int64_t i0 = blockIdx.x * (int64_t)kHistBlock + threadIdx.x; i0 < n_round; i0 += kHistUnroll * stride) {
        K key[kHistUnroll];
#pragma unroll
        for (int u = 0; u < kHistUnroll; ++u) {
            const int64_t i = i0 + u * stride;
            key[u] = i < n ? keys[i] : (K)0;
        }
#pragma unroll
        for (int u = 0; u < kHistUnroll; ++u) {
            const int64_t i = i0 + u * stride;
            if (i >= n_round) break;
            const bool valid = i < n;
            const uint64_t act = ballot64(valid);
            if (!act) continue;
            const int src = __builtin_ctzll(act);
            for (int p = 0; p < passes; ++p) {
                const int shift = lo_bit + RB * p;
                const int nb = min(RB, end_bit - shift);
                const uint32_t d = digit_of(key[u], shift, (1u << nb) - 1u);
                const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)d, src);
                const uint64_t same = ballot64(d == v) & act;
                if (same == act) {
                    if (lane == src) atomicAdd(&h[p][v], (uint32_t)__popcll(act));
                } else if (valid) {
                    atomicAdd(&h[p][d], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * R; i += kHistBlock) {
        const uint32_t c = (&h[0][0])[i];
        if (c) atomicAdd(&hist[i], c);
    }
}

// Exclusive scan of each pass's R digit counts, in place; each thread owns
// R / 256 consecutive digits.
template <int RB>
__global__ __launch_bounds__(kSortBlock) void k_sort_scan_hist(uint32_t* __restrict__ hist, const int32_t* gate) {
    if (gated_off(gate)) return;
    constexpr int R = 1 << RB;
    constexpr int DPT = R / kSortBlock;
    __shared__ uint32_t wt[kSortWaves];
    uint32_t* h = hist + blockIdx.x * R + threadIdx.x * DPT;
    uint32_t v[DPT], sum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        v[k] = h[k];
        sum += v[k];
    }
    uint32_t tot;
    uint32_t run = block_inclusive_scan<kSortWaves>(sum, [](uint32_t a, uint32_t b) { return a + b; }, wt, tot) - sum;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        h[k] = run;
        run += v[k];
    }
}

template <class K, bool IDENT, int RB, int IT, int LB>
__global__ __launch_bounds__(kSortBlock) __attribute__((amdgpu_waves_per_eu(IT <= 8 ? 6 : IT <= 16 ? 4 : 2))) void k_onesweep(const K* __restrict__ keys_in, K* __restrict__ keys_out,
                                                         const int32_t* __restrict__ vals_in,
                                                         int32_t* __restrict__ vals_out, int64_t n, int shift,
                                                         int nbits, const uint32_t* __restrict__ digit_offsets,
                                                         uint32_t* __restrict__ status,
                                                         uint32_t* __restrict__ tile_counter,
                                                         const int32_t* gate) {
    if (gated_off(gate)) return;
    constexpr int R = 1 << RB;
    constexpr int DPT = R / kSortBlock;  // digits per thread in the per-digit phase
    constexpr int kTileKeys = kSortBlock * IT;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_hist[kSortWaves][R];
    __shared__ uint32_t s_digit_start[R];
    __shared__ uint32_t s_global_base[R];
    __shared__ uint32_t s_scan[kSortWaves];
    __shared__ union {
        K keys[kTileKeys];
        int32_t vals[kTileKeys];
    } s_stage;

    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int lane = lane_id();
    const uint32_t mask = (1u << nbits) - 1u;

    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < kSortWaves * R; i += kSortBlock) (&s_wave_hist[0][0])[i] = 0u;
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t tile_base = tile * kTileKeys;
    const int64_t wave_base = tile_base + (int64_t)w * kWave * IT;

    // Keys and values are loaded up front so that the value loads' latency
    // hides behind the ranking.
    K key[IT];
    int32_t val[IT];
    uint32_t pos[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int64_t idx = wave_base + i * kWave + lane;
        key[i] = idx < n ? keys_in[idx] : (K)0;
        val[i] = IDENT ? (int32_t)idx : (idx < n ? vals_in[idx] : 0);
    }

    // Rank within the wave, items in order, lanes in order: stable.  Phase 1
    // finds each item's peers (lanes with the same digit) by ballots and packs
    // (rank among peers, first peer lane, peer count); phase 2 lets the first
    // peer of every item add the count to the wave's digit counter -- all
    // items' LDS atomics issue back to back and execute in program order, so
    // each returns the count of the earlier items; phase 3 hands the returned
    // base to the other peers.
    uint32_t packed[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int64_t idx = wave_base + i * kWave + lane;
        const bool valid = idx < n;
        const uint32_t d = digit_of(key[i], shift, mask);
        uint64_t peers = ballot64(valid);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = ballot64(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t below = (uint32_t)count_below(peers);
        const uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : 0u;
        const uint32_t cnt = (uint32_t)__popcll(peers);
        packed[i] = valid ? (below | (leader << 8) | (cnt << 16)) : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        uint32_t old = 0u;
        if (packed[i] != 0xFFFFFFFFu && (packed[i] & 0xFFu) == 0u)
            old = atomicAdd(&s_wave_hist[w][digit_of(key[i], shift, mask)], packed[i] >> 16);
        pos[i] = old;
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const uint32_t base = __shfl(pos[i], (int)((packed[i] >> 8) & 0xFFu), kWave);
        pos[i] = base + (packed[i] & 0xFFu);
    }
    __syncthreads();

    // Per digit (thread tid owns digits tid*DPT .. tid*DPT+DPT-1): exclusive
    // offsets across waves, tile totals, tile-local starts.
    uint32_t total[DPT], tsum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        uint32_t t = 0;
#pragma unroll
        for (int ww = 0; ww < kSortWaves; ++ww) {
            const uint32_t c = s_wave_hist[ww][d];
            s_wave_hist[ww][d] = t;
            t += c;
        }
        total[k] = t;
        tsum += t;
    }
    uint32_t tile_tot;
    uint32_t run = block_inclusive_scan<kSortWaves>(tsum, [](uint32_t a, uint32_t b) { return a + b; }, s_scan,
                                                    tile_tot) - tsum;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        s_digit_start[tid * DPT + k] = run;
        run += total[k];
    }

    // Decoupled look-back for each owned digit: publish this tile's count,
    // then walk back to a tile whose inclusive prefix is known.
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        atomic_store_relaxed(status + tile * R + d, (tile == 0 ? kFlagPrefix : kFlagAggregate) | total[k]);
    }
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        const int d = tid * DPT + k;
        uint32_t excl = 0;
        if (tile != 0) {
            // Read LB predecessors per round trip, then consume them in order
            // until a prefix (done) or an unpublished tile (wait, retry).
            int64_t j = tile - 1;
            while (true) {
                uint32_t sv[LB];
#pragma unroll
                for (int q = 0; q < LB; ++q)
                    sv[q] = j - q >= 0 ? atomic_load_relaxed(status + (j - q) * R + d) : kFlagPrefix;
                bool done = false, stall = false;
#pragma unroll
                for (int q = 0; q < LB; ++q) {
                    if (done || stall) continue;
                    const uint32_t flag = sv[q] & ~kValueMask;
                    if (flag == 0u) {
                        stall = true;
                        continue;
                    }
                    excl += sv[q] & kValueMask;
                    --j;
                    done = flag == kFlagPrefix;
                }
                if (done) break;
                if (stall) __builtin_amdgcn_s_sleep(1);
            }
            atomic_store_relaxed(status + tile * R + d, kFlagPrefix | (excl + total[k]));
        }
        s_global_base[d] = digit_offsets[d] + excl;
    }
    __syncthreads();

    // Stage keys in digit order: all LDS offsets first, then the writes.
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const uint32_t dd = digit_of(key[i], shift, mask);
        pos[i] += s_digit_start[dd] + s_wave_hist[w][dd];
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int64_t idx = wave_base + i * kWave + lane;
        if (idx < n) s_stage.keys[pos[i]] = key[i];
    }
    __syncthreads();

    // Write out in contiguous digit runs: read the staged keys, then their
    // global slots, then store.
    const int tile_count = (int)min((int64_t)kTileKeys, n - tile_base);
    uint32_t dst[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int slot = i * kSortBlock + tid;
        key[i] = slot < tile_count ? s_stage.keys[slot] : (K)0;
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int slot = i * kSortBlock + tid;
        const uint32_t dd = digit_of(key[i], shift, mask);
        dst[i] = s_global_base[dd] + ((uint32_t)slot - s_digit_start[dd]);
    }
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int slot = i * kSortBlock + tid;
        if (slot < tile_count) keys_out[dst[i]] = key[i];
    }
    __syncthreads();

#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int64_t idx = wave_base + i * kWave + lane;
        if (idx < n) s_stage.vals[pos[i]] = val[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const int slot = i * kSortBlock + tid;
        if (slot < tile_count) vals_out[dst[i]] = s_stage.vals[slot];
    }
}

// The low-bit pass of a sort whose high bits [lo_bit, end_bit) are already
// sorted stably: each key is ranked inside its run of equal high bits by its
// low bits, ties by position (so the order stays stable), and written to run
// start + rank.  Key bits at and above end_bit are not part of the sort key
// and are ignored, as the digit passes ignore them.  A run longer than `limit` keys sets *overflow and nothing of
// it is written; the caller then runs the remaining digit passes instead.
// Each workgroup stages its kLocalTile keys and `limit` keys either side in
// LDS, which covers every run it ranks.
constexpr int kLocalTile = 4 * kSortBlock;
constexpr int kLocalMaxRun = 2048;

__global__ __launch_bounds__(kSortBlock) void k_sort_local(const uint64_t* __restrict__ keys_in,
                                                           const int32_t* __restrict__ vals_in,
                                                           uint64_t* __restrict__ keys_out,
                                                           int32_t* __restrict__ vals_out, int64_t n, int lo_bit,
                                                           int end_bit, int limit, int32_t* __restrict__ overflow) {
    extern __shared__ uint64_t s_keys[];  // [base - limit, base + kLocalTile + limit)
    const int64_t base = blockIdx.x * (int64_t)kLocalTile;
    const int64_t lo = base - limit;
    const int span = kLocalTile + 2 * limit;
    // The values of this thread's keys are loaded with the staging so their
    // latency is not paid after each rank.
    int32_t val[kLocalTile / kSortBlock];
#pragma unroll
    for (int u = 0; u < kLocalTile / kSortBlock; ++u) {
        const int64_t p = base + u * kSortBlock + threadIdx.x;
        val[u] = p < n ? vals_in[p] : 0;
    }
    for (int k = threadIdx.x; k < span; k += kSortBlock) {
        const int64_t q = lo + k;
        s_keys[k] = (q >= 0 && q < n) ? keys_in[q] : 0ull;
    }
    __syncthreads();
    const uint64_t low_mask = (1ull << lo_bit) - 1ull;
    const uint64_t key_mask = end_bit >= 64 ? ~0ull : (1ull << end_bit) - 1ull;
    auto high = [&](uint64_t k) { return (k & key_mask) >> lo_bit; };
#pragma unroll
    for (int u = 0; u < kLocalTile / kSortBlock; ++u) {
        const int64_t p = base + u * kSortBlock + threadIdx.x;
        if (p >= n) break;
        const int lp = (int)(p - lo);
        const uint64_t k = s_keys[lp];
        const uint64_t hi = high(k);
        // A run of more than limit keys starting at p reaches p + limit: one
        // probe finds every such run at its first key without walking it.
        if (p + limit < n && high(s_keys[lp + limit]) == hi) {
            atomicOr(overflow, 1);
            continue;
        }
        int ls = lp, le = lp + 1;
        while (lo + ls > 0 && ls > 0 && high(s_keys[ls - 1]) == hi) --ls;
        while (lo + le < n && le < span && high(s_keys[le]) == hi) ++le;
        if (le - ls > limit || (ls == 0 && lo > 0) || (le == span && lo + le < n)) {
            atomicOr(overflow, 1);  // longer than limit (or reaching past the staged keys)
            continue;
        }
        const uint64_t lk = k & low_mask;
        int r = 0;
        for (int q = ls; q < le; ++q) {
            const uint64_t kq = s_keys[q] & low_mask;
            r += (kq < lk) || (kq == lk && q < lp);
        }
        const int64_t dst = lo + ls + r;
        keys_out[dst] = k;
        vals_out[dst] = val[u];
    }
}

__global__ void k_sort_flag_clear(int32_t* flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// Gated copy of a fallback sort's output into the buffers the local pass
// writes, so that both paths leave the result in the same place.
__global__ void k_sort_copy_gated(const uint64_t* __restrict__ ks, const int32_t* __restrict__ vs,
                                  uint64_t* __restrict__ kd, int32_t* __restrict__ vd, int64_t n,
                                  const int32_t* gate) {
    if (gated_off(gate)) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

// Status words (look-back flags, digit histograms, tile counters) of an LSD
// sort of n keys over `bits` key bits.
template <int RB, int IT>
int64_t sort_status_words(int64_t n, int bits) {
    const int passes = sort_passes(bits, RB);
    return (int64_t)passes * sort_tiles(n, IT) * (1 << RB) + (int64_t)passes * (1 << RB) + kMaxPasses;
}

// LSD passes over key bits [lo_bit, end_bit) from buffers 0; *result = the
// buffer index holding the output.  With a gate every launch is enqueued and
// does its work only if *gate is set (status must then be reserved already).
template <class K, int RB, int IT, int LB>
int sort_impl(Context* c, void* keys[2], int32_t* vals[2], int64_t n, int lo_bit, int end_bit, bool vals_identity,
              int* result, const int32_t* gate = nullptr, const SortWorkspace* ws = nullptr) {
    constexpr int R = 1 << RB;
    const int passes = sort_passes(end_bit - lo_bit, RB);
    *result = 0;
    if (n <= 1 || passes == 0) {
        if (vals_identity && n > 0 && gate == nullptr) {
            hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, vals[0], n);
            HX_TRY(check_launch("iota"));
        }
        return HX_OK;
    }
    const int64_t ntiles = sort_tiles(n, IT);
    const int64_t status_words = (int64_t)passes * ntiles * R;
    const int64_t words = status_words + (int64_t)passes * R + kMaxPasses;
    uint32_t* status = nullptr;
    if (ws != nullptr) {
        if (ws->status_words < words) {
            set_error("radix sort: %lld keys need %lld status words, the workspace holds %lld", (long long)n,
                      (long long)words, (long long)ws->status_words);
            return HX_ERR_INVALID_ARG;
        }
        status = ws->status;
    } else {
        if (gate != nullptr && c->sort_status.bytes < sizeof(uint32_t) * (size_t)words) {
            set_error("radix sort: gated passes need their status words reserved in advance");
            return HX_ERR_INVALID_ARG;
        }
        HX_TRY(c->sort_status.reserve(sizeof(uint32_t) * (size_t)words));
        status = c->sort_status.as<uint32_t>();
    }
    uint32_t* hist = status + status_words;
    uint32_t* counters = hist + passes * R;
    {
        LaunchScope ls(c, "sort_hist");
        const unsigned zb = (unsigned)std::min<int64_t>((words + 255) / 256, 4096);
        hipLaunchKernelGGL(k_zero_u32, dim3(zb), dim3(256), 0, c->stream, status, words, gate);
        HX_TRY(check_launch("sort_zero"));
        const unsigned hb = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>((n + kHistBlock * kHistUnroll - 1) / (kHistBlock * kHistUnroll), c->cu_count));
        hipLaunchKernelGGL((k_sort_hist<K, RB>), dim3(hb), dim3(kHistBlock), 0, c->stream,
                           static_cast<const K*>(keys[0]), n, lo_bit, passes, end_bit, hist, gate);
        HX_TRY(check_launch("sort_hist"));
        hipLaunchKernelGGL(k_sort_scan_hist<RB>, dim3(passes), dim3(kSortBlock), 0, c->stream, hist, gate);
        HX_TRY(check_launch("sort_scan_hist"));
    }
    for (int p = 0; p < passes; ++p) {
        const int in = p & 1, out = in ^ 1;
        const int shift = lo_bit + RB * p;
        const int nb = std::min(RB, end_bit - shift);
        LaunchScope ls(c, "sort_onesweep");
        if (p == 0 && vals_identity)
            hipLaunchKernelGGL((k_onesweep<K, true, RB, IT, LB>), dim3((unsigned)ntiles), dim3(kSortBlock), 0, c->stream,
                               static_cast<const K*>(keys[in]), static_cast<K*>(keys[out]), (const int32_t*)nullptr,
                               vals[out], n, shift, nb, hist + p * R, status + (int64_t)p * ntiles * R, counters + p,
                               gate);
        else
            hipLaunchKernelGGL((k_onesweep<K, false, RB, IT, LB>), dim3((unsigned)ntiles), dim3(kSortBlock), 0, c->stream,
                               static_cast<const K*>(keys[in]), static_cast<K*>(keys[out]), vals[in], vals[out], n,
                               shift, nb, hist + p * R, status + (int64_t)p * ntiles * R, counters + p, gate);
        HX_TRY(check_launch("sort_onesweep"));
    }
    *result = passes & 1;
    return HX_OK;
}

constexpr int64_t kLocalMinKeys = 1 << 18;

// High passes, then the local pass into buffers 1.  A run longer than the
// local pass ranks sets a flag on the device; the low passes over the high
// passes' output are always enqueued behind it and run only when the flag is
// set (a stable LSD sort of any order in which equal keys keep their input
// order gives the same result, so they may start from the partly sorted
// buffers), followed by a gated copy that puts their output where the local
// pass writes.  The host never reads the flag: no synchronisation, and a
// build stays capturable in a graph.
template <int RB, int IT, int LB>
int sort_split(Context* c, void* keys[2], int32_t* vals[2], int64_t n, int lo_bit, int end_bit, bool vals_identity,
               int* result) {
    // Every buffer the gated passes touch is reserved before anything runs.
    const int64_t words = std::max(sort_status_words<RB, IT>(n, end_bit), sort_status_words<RB, IT>(n, end_bit - lo_bit));
    HX_TRY(c->sort_status.reserve(sizeof(uint32_t) * (size_t)words));
    HX_TRY(c->sort_flag.reserve(64));
    int32_t* flag = c->sort_flag.as<int32_t>();
    int hi_res = 0;
    HX_TRY((sort_impl<uint64_t, RB, IT, LB>(c, keys, vals, n, lo_bit, end_bit, vals_identity, &hi_res)));
    void* k2[2] = {keys[hi_res], keys[hi_res ^ 1]};
    int32_t* v2[2] = {vals[hi_res], vals[hi_res ^ 1]};
    {
        LaunchScope ls(c, "sort_local");
        hipLaunchKernelGGL(k_sort_flag_clear, dim3(1), dim3(64), 0, c->stream, flag);
        HX_TRY(check_launch("sort_flag_clear"));
        const int limit = std::min(c->tune.sort_local_run, kLocalMaxRun);
        hipLaunchKernelGGL(k_sort_local, dim3((unsigned)((n + kLocalTile - 1) / kLocalTile)), dim3(kSortBlock),
                           sizeof(uint64_t) * (size_t)(kLocalTile + 2 * limit), c->stream,
                           static_cast<const uint64_t*>(k2[0]), v2[0], static_cast<uint64_t*>(k2[1]), v2[1], n,
                           lo_bit, end_bit, limit, flag);
        HX_TRY(check_launch("sort_local"));
    }
    int lo_res = 0;
    HX_TRY((sort_impl<uint64_t, RB, IT, LB>(c, k2, v2, n, 0, end_bit, false, &lo_res, flag)));
    if (lo_res == 0) {
        const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_sort_copy_gated, dim3(g), dim3(256), 0, c->stream, static_cast<const uint64_t*>(k2[0]),
                           v2[0], static_cast<uint64_t*>(k2[1]), v2[1], n, (const int32_t*)flag);
        HX_TRY(check_launch("sort_copy_gated"));
    }
    *result = hi_res ^ 1;
    return HX_OK;
}

}  // namespace

int sort_pass_count(const Context* c, int end_bit, int64_t n, bool keys64) {
    const int bits = c->tune.sort_digit_bits;
    const int lo_bit = end_bit - 4 * bits;
    if (keys64 && c->tune.sort_local_run > 0 && n >= kLocalMinKeys && lo_bit >= bits) return 5;
    return sort_passes(end_bit, bits);
}

int sort_status_words_max(const Context* c, int64_t n, int end_bit) {
    const int rb = c->tune.sort_digit_bits, it = c->tune.sort_items;
    const int passes = sort_passes(end_bit, rb);
    return (int)((int64_t)passes * sort_tiles(n, it) * (1 << rb) + (int64_t)passes * (1 << rb) + kMaxPasses);
}

int launch_sort_pairs(Context* c, void* keys[2], int32_t* vals[2], int64_t n, int end_bit, bool keys64,
                      bool vals_identity, int* result_index, bool allow_split, const SortWorkspace* ws) {
    if (n >= (int64_t)1 << 30) {
        set_error("radix sort: n = %lld exceeds 2^30 - 1", (long long)n);
        return HX_ERR_INVALID_ARG;
    }
    const int bits = c->tune.sort_digit_bits, items = c->tune.sort_items, lb = c->tune.sort_lookback;
    // Long keys, many of them: four digit passes over the high bits leave runs
    // of equal high bits that are short for Morton codes (at most 19 keys for
    // the 16.7M-triangle benchmark scene at 32 high bits), which one pass
    // ranks in place; a long run falls back to the remaining passes.
    const int lo_bit = end_bit - 4 * bits;
    const bool split = allow_split && ws == nullptr && keys64 && c->tune.sort_local_run > 0 && n >= kLocalMinKeys &&
                       lo_bit >= bits;
#define HX_SORT_CASE(RB, IT, LB)                                                                                   \
    if (bits == RB && items == IT && lb == LB) {                                                                  \
        if (split) return sort_split<RB, IT, LB>(c, keys, vals, n, lo_bit, end_bit, vals_identity, result_index); \
        if (keys64)                                                                                                \
            return sort_impl<uint64_t, RB, IT, LB>(c, keys, vals, n, 0, end_bit, vals_identity, result_index,     \
                                                   nullptr, ws);                                                  \
        return sort_impl<uint32_t, RB, IT, LB>(c, keys, vals, n, 0, end_bit, vals_identity, result_index, nullptr, \
                                               ws);                                                               \
    }
    HX_SORT_CASE(8, 8, 4)
    HX_SORT_CASE(8, 8, 8)
    HX_SORT_CASE(8, 8, 16)
    HX_SORT_CASE(8, 16, 4)
    HX_SORT_CASE(8, 16, 8)
    HX_SORT_CASE(8, 16, 16)
    HX_SORT_CASE(8, 32, 4)
    HX_SORT_CASE(8, 32, 8)
    HX_SORT_CASE(8, 32, 16)
    HX_SORT_CASE(10, 8, 4)
    HX_SORT_CASE(10, 8, 8)
    HX_SORT_CASE(10, 8, 16)
    HX_SORT_CASE(10, 16, 4)
    HX_SORT_CASE(10, 16, 8)
    HX_SORT_CASE(10, 16, 16)
    HX_SORT_CASE(10, 32, 4)
    HX_SORT_CASE(10, 32, 8)
    HX_SORT_CASE(10, 32, 16)
#undef HX_SORT_CASE
    set_error("radix sort: no kernel for %d-bit digits, %d keys per thread, look-back %d", bits, items, lb);
    return HX_ERR_INVALID_ARG;
}

}  // namespace hx
