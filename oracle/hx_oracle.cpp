/*
 * hx_oracle.cpp -- CPU restatement of hippie's LBVH path (TEST INFRASTRUCTURE).
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg; the product never links it.  Built with -O3 -ffp-contract=off -fopenmp
 * (oracle/Makefile).  Reference citations are relative to src/hippie/.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   scene box, setup boxes, refit  -- exact min/max vs brute-force unions.
 *   Morton codes                   -- bit-by-bit interleave closed form over
 *                                     all 2^10 values per half, end points.
 *   sort                           -- numpy stable argsort (library routine).
 *   construct                      -- brute-force top-down split at the highest
 *                                     differing bit (middle index for ties).
 *   collapse                       -- brute-force recursive SAH decision on the
 *                                     binary tree (same formula) + validate().
 *   woopify                        -- FP64 check that the transform maps
 *                                     v0,v1,v2 to (u,v) = (1,0),(0,1),(0,0).
 *   cost                           -- FP64 closed form over the canonical tree.
 *   trace                          -- brute force over all triangles; FP64
 *                                     Moller-Trumbore for t.
 *   ray keys                       -- bit loops re-derived in numpy.
 */
#include "hx_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

/* Globals.h:70-74 -- MIN/MAX are selects, not fminf/fmaxf (signed zeros and
 * NaN follow the select). */
constexpr float kMaxFloat = 3.402823466e+38f;
inline float MINs(float a, float b) { return (b < a) ? b : a; }
inline float MAXs(float a, float b) { return (b > a) ? b : a; }

struct V3 {
    float x, y, z;
};

/* util/AABB.h:14-49 */
struct Box {
    V3 mn{kMaxFloat, kMaxFloat, kMaxFloat};
    V3 mx{-kMaxFloat, -kMaxFloat, -kMaxFloat};
    void grow(const V3& p) {
        mn.x = MINs(mn.x, p.x); mn.y = MINs(mn.y, p.y); mn.z = MINs(mn.z, p.z);
        mx.x = MAXs(mx.x, p.x); mx.y = MAXs(mx.y, p.y); mx.z = MAXs(mx.z, p.z);
    }
    bool valid() const { return mn.x <= mx.x && mn.y <= mx.y && mn.z <= mx.z; }
    float area() const {
        if (!valid()) return 0.0f;
        float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

inline V3 load3(const float* p) { return V3{p[0], p[1], p[2]}; }
inline V3 load4(const float* p) { return V3{p[0], p[1], p[2]}; }
inline void store4(float* p, const V3& v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = 0.0f; }

/* Float -> unsigned / int conversions with the saturating semantics of the
 * GPU convert instructions (NaN -> 0).  Inputs on the path are in range; the
 * helpers only make the out-of-range cases defined and equal on both sides. */
inline uint32_t f2u(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}
inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 0x7FFFFFFF;
    if (x <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)x;
}

/* rt/HipUtil.h:35-49 */
inline uint32_t morton10(uint32_t x, uint32_t y, uint32_t z) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    y = (y | (y << 16)) & 0x030000FFu;
    y = (y | (y << 8)) & 0x0300F00Fu;
    y = (y | (y << 4)) & 0x030C30C3u;
    y = (y | (y << 2)) & 0x09249249u;
    z = (z | (z << 16)) & 0x030000FFu;
    z = (z | (z << 8)) & 0x0300F00Fu;
    z = (z | (z << 4)) & 0x030C30C3u;
    z = (z | (z << 2)) & 0x09249249u;
    return x | (y << 1) | (z << 2);
}

/* rt/HipUtil.h:72-82,100-106 */
inline uint64_t morton60(const V3& c) {
    const float scale = (float)((1u << 20) - 1);
    uint32_t x = f2u(c.x * scale);
    uint32_t y = f2u(c.y * scale);
    uint32_t z = f2u(c.z * scale);
    uint64_t lo = morton10(x & 1023u, y & 1023u, z & 1023u);
    uint64_t hi = morton10(x >> 10u, y >> 10u, z >> 10u);
    return (hi << 30) | lo;
}

/* rt/bvh/HipBVHUtil.h:17-24 */
inline int delta(int64_t i, int64_t j, int64_t n, const uint64_t* k) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return __builtin_clzll(a ^ b);
    return __builtin_clzll((uint64_t)(i ^ j)) + 64;
}

inline int sgn(int v) { return (0 < v) - (v < 0); }

template <class F>
void parallel_for(int64_t n, F f) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) f(i);
}

/* util/Math.h:318-346 (det of a 3x3 built from MatBase (r,c) accessors). */
inline float det3(const float m[3][3]) {
    /* v(r,c) = m[r][c] */
    return m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1] +
           m[1][0] * m[2][1] * m[0][2] - m[1][0] * m[2][2] * m[0][1] +
           m[2][0] * m[0][1] * m[1][2] - m[2][0] * m[0][2] * m[1][1];
}

/* util/Math.h:349-369 inverse() of a 4x4 via cofactors; d sums all 16
 * cofactor products (= 4 det), result = r * rcp(d) * 4. m[r][c]. */
void inverse4(const float v[4][4], float r[4][4]) {
    float d = 0.0f;
    float si = 1.0f;
    for (int i = 0; i < 4; ++i) {
        float sj = si;
        for (int j = 0; j < 4; ++j) {
            float sub[3][3];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    sub[k][l] = v[(k < j) ? k : k + 1][(l < i) ? l : l + 1];
            float dd = det3(sub) * sj;
            r[i][j] = dd;
            d += dd * v[j][i];
            sj = -sj;
        }
        si = -si;
    }
    float rd = (d != 0.0f) ? 1.0f / d : 0.0f;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) r[a][b] = (r[a][b] * rd) * 4.0f;
}

/* rt/bvh/HipBVHKernels.cu:12-33,59-61 */
void woop_rows(const V3& v0, const V3& v1, const V3& v2, float out[12]) {
    V3 v02{v0.x - v2.x, v0.y - v2.y, v0.z - v2.z};
    V3 v12{v1.x - v2.x, v1.y - v2.y, v1.z - v2.z};
    /* util/Math.h cross() */
    V3 nrm{v02.y * v12.z - v02.z * v12.y, v02.z * v12.x - v02.x * v12.z,
           v02.x * v12.y - v02.y * v12.x};
    /* m[col][row] in the reference; here M[row][col]. */
    float M[4][4];
    M[0][0] = v02.x; M[1][0] = v02.y; M[2][0] = v02.z; M[3][0] = 0.0f;
    M[0][1] = v12.x; M[1][1] = v12.y; M[2][1] = v12.z; M[3][1] = 0.0f;
    M[0][2] = nrm.x; M[1][2] = nrm.y; M[2][2] = nrm.z; M[3][2] = 0.0f;
    M[0][3] = v2.x;  M[1][3] = v2.y;  M[2][3] = v2.z;  M[3][3] = 1.0f;
    float I[4][4];
    inverse4(M, I);
    out[0] = I[2][0]; out[1] = I[2][1]; out[2] = I[2][2]; out[3] = -I[2][3];
    out[4] = I[0][0]; out[5] = I[0][1]; out[6] = I[0][2]; out[7] = I[0][3];
    out[8] = I[1][0]; out[9] = I[1][1]; out[10] = I[1][2]; out[11] = I[1][3];
}

inline void tri_verts(const float* verts, const int32_t* tris, int64_t t, V3& a, V3& b, V3& c) {
    const int32_t* tr = tris + 3 * t;
    a = load3(verts + 3 * (int64_t)tr[0]);
    b = load3(verts + 3 * (int64_t)tr[1]);
    c = load3(verts + 3 * (int64_t)tr[2]);
}

/* Bin node accessors: rt/bvh/HipBVHNode.h:117-216. */
inline float w2f(int32_t w) { float f; std::memcpy(&f, &w, 4); return f; }
inline int32_t f2w(float f) { int32_t w; std::memcpy(&w, &f, 4); return w; }

void set_child_box(int32_t* node, int i, const V3& mn, const V3& mx) {
    node[4 * i + 0] = f2w(mn.x);
    node[4 * i + 2] = f2w(mn.y);
    node[2 * i + 8] = f2w(mn.z);
    node[4 * i + 1] = f2w(mx.x);
    node[4 * i + 3] = f2w(mx.y);
    node[2 * i + 9] = f2w(mx.z);
}

/* Trace helpers: rt/tracer/HipTracerKernels.h:124-195 -- fminf/fmaxf with
 * the semantics of the GPU min/max instructions (a NaN operand yields the
 * other operand; -0 orders below +0), then the integer min/max on float bits. */
inline float fmin_s(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return std::signbit(a) ? a : b;
    return a < b ? a : b;
}
inline float fmax_s(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    if (a == b) return std::signbit(a) ? b : a;
    return a > b ? a : b;
}
inline int32_t fb(float f) { return f2w(f); }
inline float bf(int32_t i) { return w2f(i); }
inline float i_min_max(float a, float b, float c) {  /* max(min(a,b),c) on bits */
    int32_t x = std::min(fb(a), fb(b));
    return bf(std::max(x, fb(c)));
}
inline float i_max_min(float a, float b, float c) {  /* min(max(a,b),c) on bits */
    int32_t x = std::max(fb(a), fb(b));
    return bf(std::min(x, fb(c)));
}
inline float i_max_max(float a, float b, float c) { return bf(std::max(std::max(fb(a), fb(b)), fb(c))); }
inline float i_min_min(float a, float b, float c) { return bf(std::min(std::min(fb(a), fb(b)), fb(c))); }
inline float span_begin(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    return i_max_max(fmin_s(a0, a1), fmin_s(b0, b1), i_min_max(c0, c1, d));
}
inline float span_end(float a0, float a1, float b0, float b1, float c0, float c1, float d) {
    return i_min_min(fmax_s(a0, a1), fmax_s(b0, b1), i_max_min(c0, c1, d));
}

constexpr int32_t kSentinel = 0x76543210;  /* HipTracerKernels.h:11 */
constexpr int kStackSize = 64;             /* HipTracerBinKernels.cu:5 */

struct RayState {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
    float idx, idy, idz, oodx, oody, oodz;
};

inline void setup_ray(const float* r, RayState& s) {
    s.ox = r[0]; s.oy = r[1]; s.oz = r[2]; s.tmin = r[3];
    s.dx = r[4]; s.dy = r[5]; s.dz = r[6]; s.tmax = r[7];
    const float ooeps = 8.2718061255302767e-25f; /* exp2f(-80.0f) */
    s.idx = 1.0f / (std::fabs(s.dx) > ooeps ? s.dx : std::copysign(ooeps, s.dx));
    s.idy = 1.0f / (std::fabs(s.dy) > ooeps ? s.dy : std::copysign(ooeps, s.dy));
    s.idz = 1.0f / (std::fabs(s.dz) > ooeps ? s.dz : std::copysign(ooeps, s.dz));
    s.oodx = s.ox * s.idx;
    s.oody = s.oy * s.idy;
    s.oodz = s.oz * s.idz;
}

/* Woop test of HipTracerBinKernels.cu:215-268 written with the fused
 * multiply-adds an FMA-contracting compiler forms from it. Returns true and
 * writes t,u,v on a hit within (tmin, hitT]. */
inline bool woop_test(const RayState& s, const float* w, float hitT, float& t, float& u, float& v) {
    float Oz = std::fmaf(-s.oz, w[2], std::fmaf(-s.oy, w[1], std::fmaf(-s.ox, w[0], w[3])));
    float den = std::fmaf(s.dz, w[2], std::fmaf(s.dy, w[1], s.dx * w[0]));
    float invDz = 1.0f / den;
    t = Oz * invDz;
    if (!(t > s.tmin && t <= hitT)) return false;
    float Ox = std::fmaf(s.oz, w[6], std::fmaf(s.oy, w[5], std::fmaf(s.ox, w[4], w[7])));
    float Dx = std::fmaf(s.dz, w[6], std::fmaf(s.dy, w[5], s.dx * w[4]));
    u = std::fmaf(t, Dx, Ox);
    if (!(u >= 0.0f)) return false;
    float Oy = std::fmaf(s.oz, w[10], std::fmaf(s.oy, w[9], std::fmaf(s.ox, w[8], w[11])));
    float Dy = std::fmaf(s.dz, w[10], std::fmaf(s.dy, w[9], s.dx * w[8]));
    v = std::fmaf(t, Dy, Oy);
    return v >= 0.0f && u + v <= 1.0f;
}

}  // namespace

extern "C" {

void hxo_set_threads(int threads) {
    if (threads > 0) omp_set_num_threads(threads);
}

int hxo_get_threads(void) { return omp_get_max_threads(); }

/* SceneLoader.cpp:134-141 after AssimpImporter.cpp:99-113: the box of every
 * vertex, scale = 1 / qMax(qMax(d.x, d.y), d.z) (qMax(a, b) = a < b ? b : a),
 * then v -= mn and v *= scale per component; box6 receives the vertex box. */
void hxo_normalize(float* verts, int64_t nv, float box6[6]) {
    Box b;
    for (int64_t i = 0; i < nv; ++i) b.grow(V3{verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]});
    const float dx = b.mx.x - b.mn.x, dy = b.mx.y - b.mn.y, dz = b.mx.z - b.mn.z;
    const float dxy = dx < dy ? dy : dx;
    const float scale = 1.0f / (dxy < dz ? dz : dxy);
    const float mn[3] = {b.mn.x, b.mn.y, b.mn.z};
    for (int64_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) {
            const float d = verts[3 * i + k] - mn[k];
            verts[3 * i + k] = d * scale;
        }
    box6[0] = b.mn.x; box6[1] = b.mn.y; box6[2] = b.mn.z;
    box6[3] = b.mx.x; box6[4] = b.mx.y; box6[5] = b.mx.z;
}

void hxo_scene_box(const float* verts, const int32_t* tris, int64_t n, float out[6]) {
    Box scene;
    for (int64_t t = 0; t < n; ++t) {
        V3 a, b, c;
        tri_verts(verts, tris, t, a, b, c);
        Box box;
        box.grow(a); box.grow(b); box.grow(c);
        scene.grow(box.mn);
        scene.grow(box.mx);
    }
    out[0] = scene.mn.x; out[1] = scene.mn.y; out[2] = scene.mn.z;
    out[3] = scene.mx.x; out[4] = scene.mx.y; out[5] = scene.mx.z;
}

void hxo_setup_boxes(const float* verts, const int32_t* tris, int64_t n, float* bmin, float* bmax) {
    parallel_for(n, [&](int64_t t) {
        V3 a, b, c;
        tri_verts(verts, tris, t, a, b, c);
        Box box;
        box.grow(a); box.grow(b); box.grow(c);
        store4(bmin + 4 * t, box.mn);
        store4(bmax + 4 * t, box.mx);
    });
}

void hxo_morton_codes(const float* bmin, const float* bmax, int64_t n, const float sb[6], int bits,
                      uint64_t* keys) {
    V3 smn{sb[0], sb[1], sb[2]};
    V3 diag{sb[3] - sb[0], sb[4] - sb[1], sb[5] - sb[2]};
    V3 scale{1.0f / diag.x, 1.0f / diag.y, 1.0f / diag.z};
    const int shift = 60 - bits;
    parallel_for(n, [&](int64_t r) {
        Box box;
        box.grow(load4(bmin + 4 * r));
        box.grow(load4(bmax + 4 * r));
        V3 c{(box.mn.x + box.mx.x) * 0.5f, (box.mn.y + box.mx.y) * 0.5f, (box.mn.z + box.mx.z) * 0.5f};
        V3 q{(c.x - smn.x) * scale.x, (c.y - smn.y) * scale.y, (c.z - smn.z) * scale.z};
        keys[r] = morton60(q) >> shift;
    });
}

void hxo_sort_pairs(uint64_t* keys, int32_t* vals, int64_t n, int end_bit) {
    if (n <= 1 || end_bit <= 0) return;
    std::vector<uint64_t> k2(n);
    std::vector<int32_t> v2(n);
    uint64_t* ks = keys; int32_t* vs = vals;
    uint64_t* kd = k2.data(); int32_t* vd = v2.data();
    const int T = omp_get_max_threads();
    std::vector<int64_t> hist((size_t)T * 256);
    for (int bit = 0; bit < end_bit; bit += 8) {
        const int nb = std::min(8, end_bit - bit);
        const uint64_t mask = (1ull << nb) - 1;
        std::fill(hist.begin(), hist.end(), 0);
#pragma omp parallel num_threads(T)
        {
            int tid = omp_get_thread_num();
            int64_t lo = n * tid / T, hi = n * (tid + 1) / T;
            int64_t* h = hist.data() + (size_t)tid * 256;
            for (int64_t i = lo; i < hi; ++i) h[(ks[i] >> bit) & mask]++;
#pragma omp barrier
#pragma omp single
            {
                int64_t sum = 0;
                for (int d = 0; d < 256; ++d)
                    for (int t = 0; t < T; ++t) {
                        int64_t c = hist[(size_t)t * 256 + d];
                        hist[(size_t)t * 256 + d] = sum;
                        sum += c;
                    }
            }
            for (int64_t i = lo; i < hi; ++i) {
                int64_t dst = h[(ks[i] >> bit) & mask]++;
                kd[dst] = ks[i];
                vd[dst] = vs[i];
            }
        }
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    if (ks != keys) {
        std::memcpy(keys, ks, sizeof(uint64_t) * n);
        std::memcpy(vals, vs, sizeof(int32_t) * n);
    }
}

void hxo_construct(const uint64_t* k, int64_t n, int32_t* left, int32_t* right, int32_t* parent) {
    /* setupLeaves (LBVHBuilderKernels.cu:185-187): leaf node n-1+i has
     * left = i, right = i + 1. */
    parallel_for(n, [&](int64_t i) {
        left[n - 1 + i] = (int32_t)i;
        right[n - 1 + i] = (int32_t)(i + 1);
    });
    parent[0] = -1;
    parallel_for(n - 1, [&](int64_t i) {
        const int d = sgn(delta(i, i + 1, n, k) - delta(i, i - 1, n, k));
        const int dmin = delta(i, i - d, n, k);
        int64_t lmax = 2;
        while (delta(i, i + lmax * d, n, k) > dmin) lmax <<= 1;
        int64_t l = 0;
        for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(i, i + (l + t) * d, n, k) > dmin) l += t;
        const int64_t j = i + l * d;
        const int dnode = delta(i, j, n, k);
        int64_t s = 0, kk = 2, t;
        do {
            t = (l + kk - 1) / kk;
            kk <<= 1;
            if (delta(i, i + (s + t) * d, n, k) > dnode) s += t;
        } while (t > 1);
        const int64_t gamma = i + s * d + std::min(d, 0);
        int64_t lc = gamma, rc = gamma + 1;
        if (std::min(i, j) == gamma) lc += n - 1;
        if (std::max(i, j) == gamma + 1) rc += n - 1;
        left[i] = (int32_t)lc;
        right[i] = (int32_t)rc;
        parent[lc] = (int32_t)i;
        parent[rc] = (int32_t)i;
    });
}

void hxo_refit(const int32_t* sv, const float* rmin, const float* rmax, int64_t n, const int32_t* left,
               const int32_t* right, const int32_t* parent, float* nmin, float* nmax) {
    const int64_t N = 2 * n - 1;
    /* setupLeaves: leaf box = reference box of the sorted reference. */
    parallel_for(n, [&](int64_t i) {
        int64_t r = sv[i];
        std::memcpy(nmin + 4 * (n - 1 + i), rmin + 4 * r, 16);
        std::memcpy(nmax + 4 * (n - 1 + i), rmax + 4 * r, 16);
    });
    if (n < 2) return;
    /* refit (LBVHBuilderKernels.cu:254-305): second arriver unions children. */
    std::vector<std::atomic<int>> counters(n - 1);
    for (auto& c : counters) c.store(0, std::memory_order_relaxed);
    parallel_for(n, [&](int64_t t) {
        int64_t node = parent[t + (N >> 1)];
        while (node >= 0 && counters[node].fetch_add(1, std::memory_order_acq_rel) > 0) {
            int64_t l = left[node], r = right[node];
            Box box;
            box.grow(load4(nmin + 4 * l));
            box.grow(load4(nmin + 4 * r));
            V3 mn = box.mn;
            box.grow(load4(nmax + 4 * l));
            box.grow(load4(nmax + 4 * r));
            store4(nmin + 4 * node, mn);
            store4(nmax + 4 * node, box.mx);
            std::atomic_thread_fence(std::memory_order_release);
            node = parent[node];
        }
    });
}

/* PLOC clustering: PLOCBuilder::clustering (PLOCBuilder.cpp:21-159) with
 * generateNeighbours / merge / compact (PLOCBuilderKernels.cu:116-389).
 * Input: leaf boxes of the sorted references in nmin/nmax[n-1+i] (as written
 * by hxo_refit's leaf setup).  Output: left/right/parent and boxes of the
 * internal nodes 0..n-2.  The reference numbers the nodes a round creates
 * with warp atomics; here they are numbered in cluster order. */
int hxo_ploc(int64_t n, int radius, int max_radius, int adaptive, int32_t* left, int32_t* right, int32_t* parent,
             float* nmin, float* nmax) {
    std::vector<int32_t> cl(n), next(n), nb(n);
    for (int64_t i = 0; i < n; ++i) {
        cl[i] = (int32_t)(n - 1 + i);
        left[n - 1 + i] = (int32_t)i;
        right[n - 1 + i] = (int32_t)(i + 1);
    }
    parent[n - 1] = -1;
    int64_t m = n;
    int rounds = 0;
    int r = radius;
    while (m > 1) {
        ++rounds;
        /* Adaptive radius (PLOCBuilder.cpp:57-58). */
        if (adaptive) r = std::min(r + 1, max_radius);
        /* generateNeighbours: leftmost minimum of area(union) in [i-r, i+r]. */
        parallel_for(m, [&](int64_t i) {
            const int64_t ni = cl[i];
            V3 bmn = load4(nmin + 4 * ni), bmx = load4(nmax + 4 * ni);
            int32_t best = -1;
            float bestD = kMaxFloat;
            const int64_t lo = std::max<int64_t>(0, i - r), hi = std::min<int64_t>(m, i + r + 1);
            for (int64_t j = lo; j < hi; ++j) {
                if (j == i) continue;
                const int64_t nj = cl[j];
                Box nbx;
                nbx.mn = load4(nmin + 4 * nj);
                nbx.mx = load4(nmax + 4 * nj);
                nbx.grow(bmn);
                nbx.grow(bmx);
                const float d = nbx.area();
                if (bestD > d) { bestD = d; best = (int32_t)j; }
            }
            nb[i] = best;
        });
        /* merge + compact, merges numbered in cluster order from m-2 down. */
        int64_t nodeOffset = m - 2, merges = 0, out = 0;
        for (int64_t i = 0; i < m; ++i) {
            const int64_t j = nb[i];
            const bool mutual = j >= 0 && nb[j] == i;
            if (mutual && i > j) continue;          /* absorbed by its partner */
            if (!mutual) { next[out++] = cl[i]; continue; }
            const int64_t node = nodeOffset - merges++;
            const int32_t ln = cl[i], rn = cl[j];
            V3 lmn = load4(nmin + 4 * (int64_t)ln), rmn = load4(nmin + 4 * (int64_t)rn);
            V3 lmx = load4(nmax + 4 * (int64_t)ln), rmx = load4(nmax + 4 * (int64_t)rn);
            store4(nmin + 4 * node, V3{MINs(lmn.x, rmn.x), MINs(lmn.y, rmn.y), MINs(lmn.z, rmn.z)});
            store4(nmax + 4 * node, V3{MAXs(lmx.x, rmx.x), MAXs(lmx.y, rmx.y), MAXs(lmx.z, rmx.z)});
            parent[ln] = (int32_t)node;
            parent[rn] = (int32_t)node;
            left[node] = ln;
            right[node] = rn;
            next[out++] = (int32_t)node;
        }
        std::swap(cl, next);
        if (merges == 0) return -1; /* no finite union area: the reference reads nb[-1] */
        m = out;
    }
    parent[cl[0]] = -1;
    return rounds;
}

/* ------------------------------------------------------------------------ */
/* ATRBVH (ATRBuilder.cpp:23-177, ATRBuilderKernels.cu:11-577,
 * TRBuilderUtil.h:11-289).  The reference restructures one treelet per
 * 32-lane warp; its two warp reductions (findLeafToExpand over 2+i lanes,
 * findMinimum over 32 lanes of strided minima) decide ties by the shape of
 * the shuffle tree, so both are emulated lane for lane here.  Nodes are
 * processed in post-order, which is every order the reference's climb can
 * take (a node is processed once both children are final, and its treelet
 * lies inside its own subtree). */
}  // extern "C"

namespace {

/* TRBuilderUtil.h:112-121 / 123-144: note the x*y + x*z + y*z order. */
inline float atr_area(const V3& mn, const V3& mx) {
    const float x = mx.x - mn.x, y = mx.y - mn.y, z = mx.z - mn.z;
    return 2.0f * (x * y + x * z + y * z);
}
inline float atr_union_area(const V3& amn, const V3& amx, const V3& bmn, const V3& bmx) {
    const float x = std::fmax(amx.x, bmx.x) - std::fmin(amn.x, bmn.x);
    const float y = std::fmax(amx.y, bmx.y) - std::fmin(amn.y, bmn.y);
    const float z = std::fmax(amx.z, bmx.z) - std::fmin(amn.z, bmn.z);
    return 2.0f * (x * y + x * z + y * z);
}
inline int trm_index(int row, int col) { return ((row - 1) * row) / 2 + col; }
/* ATRBuilderKernels.cu:11-20 (float sqrt, as the kernel computes it). */
inline int trm_row(int index) { return (int)((-1.0f + std::sqrt(8.0f * (float)index + 1.0f)) / 2.0f) + 1; }
inline int trm_col(int index) {
    const int y = trm_row(index);
    return index - (y * (y - 1)) / 2;
}

/* Shuffle-down tree over k lanes (TRBuilderUtil.h:157-169, ATRBuilderKernels.cu
 * :67-79): every lane reads its partner's value from before the step; a lane
 * keeps its own pair unless the partner's is strictly better. */
template <class Better>
void lane_tree(int k, int* idx, float* val, Better better) {
    int oi[64];
    float ov[64];
    while (k > 1) {
        const int s = k / 2;
        const int live = (k + 1) / 2;
        for (int i = 0; i < live; ++i) {
            oi[i] = idx[i + s];
            ov[i] = val[i + s];
        }
        for (int i = 0; i < live; ++i)
            if (better(ov[i], val[i])) {
                val[i] = ov[i];
                idx[i] = oi[i];
            }
        k = live;
    }
}

struct AtrTree {
    int64_t n;
    int32_t* left;
    int32_t* right;
    int32_t* parent;
    float* nmin;
    float* nmax;
    std::vector<float> area, cost;
    std::vector<int32_t> refs;
    bool is_leaf(int32_t v) const { return v >= n - 1; }
    V3 bmin(int32_t v) const { return load4(nmin + 4 * (int64_t)v); }
    V3 bmax(int32_t v) const { return load4(nmax + 4 * (int64_t)v); }
};

/* One treelet at root X (formTreelet, calculateDistancesMatrix, the
 * lastRow loop with findMinimumDistance / updateState / updateDistancesMatrix,
 * updateTreelet).  Returns true if the restructured treelet was kept. */
bool atr_treelet(AtrTree& T, int32_t X, int t, float ct) {
    int internal[32], leaves[32];
    float areas[32];
    internal[0] = X;
    leaves[0] = T.left[X];
    leaves[1] = T.right[X];
    areas[0] = T.area[leaves[0]];
    areas[1] = T.area[leaves[1]];
    for (int it = 0; it < t - 2; ++it) {
        int idx[64];
        float val[64];
        const int k = 2 + it;
        for (int i = 0; i < k; ++i) {
            idx[i] = i;
            val[i] = T.is_leaf(leaves[i]) ? -kMaxFloat : areas[i];
        }
        lane_tree(k, idx, val, [](float o, float m) { return o > m; });
        const int e = idx[0];
        const int32_t rep = leaves[e];
        internal[it + 1] = rep;
        leaves[e] = T.left[rep];
        leaves[it + 2] = T.right[rep];
        areas[e] = T.area[leaves[e]];
        areas[it + 2] = T.area[leaves[it + 2]];
    }
    int node[32], rnode[32];
    float sah[32], mcost[32];
    V3 mn[32], mx[32];
    for (int i = 0; i < t; ++i) {
        node[i] = leaves[i];
        sah[i] = T.cost[leaves[i]];
        mn[i] = T.bmin(leaves[i]);
        mx[i] = T.bmax(leaves[i]);
    }
    float D[32 * 31 / 2];
    for (int a = 1; a < t; ++a)
        for (int b = 0; b < a; ++b) D[trm_index(a, b)] = atr_union_area(mn[a], mx[a], mn[b], mx[b]);
    for (int lastRow = t - 1; lastRow > 0; --lastRow) {
        /* findMinimumDistance: 32 lanes, strided, strict '<'. */
        int idx[64];
        float val[64];
        const int msize = lastRow * (lastRow + 1) / 2;
        for (int l = 0; l < 32; ++l) {
            idx[l] = 0;
            val[l] = kMaxFloat;
            for (int j = l; j < msize; j += 32)
                if (D[j] < val[l]) {
                    val[l] = D[j];
                    idx[l] = j;
                }
        }
        lane_tree(32, idx, val, [](float o, float m) { return o < m; });
        const int minIndex = idx[0];
        const int joinCol = trm_col(minIndex), joinRow = trm_row(minIndex);
        /* updateState (ATRBuilderKernels.cu:96-158). */
        const float s = sah[joinRow] + sah[joinCol];
        const int leftIndex = node[joinRow], rightIndex = node[joinCol];
        V3 amn = mn[joinRow], amx = mx[joinRow];
        const V3 bmn = mn[joinCol], bmx = mx[joinCol];
        amn = V3{std::fmin(bmn.x, amn.x), std::fmin(bmn.y, amn.y), std::fmin(bmn.z, amn.z)};
        amx = V3{std::fmax(bmx.x, amx.x), std::fmax(bmx.y, amx.y), std::fmax(bmx.z, amx.z)};
        const float c = s + ct * atr_area(amn, amx);
        node[joinCol] = internal[lastRow - 1];
        mn[joinCol] = amn;
        mx[joinCol] = amx;
        mcost[lastRow] = c;
        sah[joinCol] = c;
        const int lastIndex = node[lastRow];
        const float sahLast = sah[lastRow];
        const V3 lmn = mn[lastRow], lmx = mx[lastRow];
        node[joinRow] = lastIndex;
        sah[joinRow] = sahLast;
        mn[joinRow] = lmn;
        mx[joinRow] = lmx;
        node[lastRow] = leftIndex;
        rnode[lastRow] = rightIndex;
        mn[lastRow] = amn;
        mx[lastRow] = amx;
        if (lastRow > 1) {
            /* updateDistancesMatrix (:160-196). */
            for (int l = 0; l < lastRow; ++l)
                if (l != joinRow)
                    D[trm_index(std::max(joinRow, l), std::min(joinRow, l))] = D[trm_index(lastRow, l)];
            for (int l = 0; l < lastRow; ++l)
                if (l != joinCol) {
                    const int r = std::max(joinCol, l), cc = std::min(joinCol, l);
                    D[trm_index(r, cc)] = atr_union_area(mn[r], mx[r], mn[cc], mx[cc]);
                }
        }
    }
    /* updateTreelet (:198-228). */
    if (!(mcost[1] < T.cost[X])) return false;
    for (int l = 1; l < t; ++l) {
        const int32_t v = internal[l - 1];
        T.left[v] = node[l];
        T.right[v] = rnode[l];
        T.parent[node[l]] = v;
        T.parent[rnode[l]] = v;
        T.cost[v] = mcost[l];
        store4(T.nmin + 4 * (int64_t)v, mn[l]);
        store4(T.nmax + 4 * (int64_t)v, mx[l]);
        T.area[v] = atr_area(mn[l], mx[l]);
    }
    return true;
}

/* TR (TRBuilder.cpp:140-347, TRBuilderKernels.cu:11-476): the treelet of 7
 * leaves formed as ATRBVH's, then the optimal binary tree over it by a DP over
 * its 128 leaf subsets: C(S) = min(ct A(S) + min_P [C(P) + C(S\P)],
 * ci A(S) R) with R the treelet root's reference count (the reference's leaf
 * option counts every reference of the treelet).  Subsets of 2..5 leaves take
 * the first strict minimum in processSchedule's partition order; the 6-leaf
 * subsets take it over 4 lanes and the whole treelet over 32 lanes, with the
 * reference's shuffle-down trees deciding ties.  The kept structure is laid
 * out as updateTreelet does (internal nodes handed out left sides first in
 * stack order, then right sides).  Every enumerated partition leaves out the
 * subset's lowest leaf, so that leaf always goes right.  The costs written
 * matter: later treelets of the same pass that reach into this one read them.
 * A new right child's cost is the reference's costs[partition] -- the left
 * side's -- and the root keeps its climb cost. */
constexpr int kTrT = 7;

float tr_area(const V3& mn, const V3& mx) { return atr_area(mn, mx); }

bool tr_treelet(AtrTree& T, int32_t X, float ct, float ci) {
    const int t = kTrT;
    int internal[kTrT], leaves[kTrT];
    float areas[kTrT];
    internal[0] = X;
    leaves[0] = T.left[X];
    leaves[1] = T.right[X];
    areas[0] = T.area[leaves[0]];
    areas[1] = T.area[leaves[1]];
    for (int it = 0; it < t - 2; ++it) {
        int idx[64];
        float val[64];
        const int k = 2 + it;
        for (int i = 0; i < k; ++i) {
            idx[i] = i;
            val[i] = T.is_leaf(leaves[i]) ? -kMaxFloat : areas[i];
        }
        lane_tree(k, idx, val, [](float o, float m) { return o > m; });
        const int e = idx[0];
        const int32_t rep = leaves[e];
        internal[it + 1] = rep;
        leaves[e] = T.left[rep];
        leaves[it + 2] = T.right[rep];
        areas[e] = T.area[leaves[e]];
        areas[it + 2] = T.area[leaves[it + 2]];
    }
    /* calculateSubsetSurfaceAreas (:22-81). */
    const int S = 1 << t;
    std::vector<V3> smn(S), smx(S);
    std::vector<float> cost(S);
    std::vector<int> mask(S, 0);
    for (int s = 0; s < S; ++s) {
        V3 mn{kMaxFloat, kMaxFloat, kMaxFloat}, mx{-kMaxFloat, -kMaxFloat, -kMaxFloat};
        for (int i = 0; i < t; ++i)
            if (s & (1 << i)) {
                const V3 a = T.bmin(leaves[i]), b = T.bmax(leaves[i]);
                mn = V3{std::fmin(a.x, mn.x), std::fmin(a.y, mn.y), std::fmin(a.z, mn.z)};
                mx = V3{std::fmax(b.x, mx.x), std::fmax(b.y, mx.y), std::fmax(b.z, mx.z)};
            }
        smn[s] = mn;
        smx[s] = mx;
        cost[s] = tr_area(mn, mx);
    }
    for (int i = 0; i < t; ++i) cost[1 << i] = T.cost[leaves[i]];
    const float R = (float)T.refs[X];
    auto finish = [&](int s, float opt, int part) {
        cost[s] = std::fmin(ct * cost[s] + opt, ci * cost[s] * R);
        mask[s] = part;
    };
    /* processSchedule (:84-114): subsets of 2..t-2 leaves, smaller first. */
    for (int size = 2; size <= t - 2; ++size)
        for (int s = 1; s < S; ++s) {
            if (__builtin_popcount(s) != size) continue;
            float opt = kMaxFloat;
            int best = 0;
            const int delta = (s - 1) & s;
            int p = (-delta) & s;
            while (p != 0) {
                const float c = cost[p] + cost[p ^ s];
                if (c < opt) {
                    opt = c;
                    best = p;
                }
                p = (p - delta) & s;
            }
            finish(s, opt, best);
        }
    /* processSubsets (:116-196): the t subsets of t-1 leaves over 4 lanes each. */
    for (int g = 0; g < t; ++g) {
        const int s = (S - 1) & ~(1 << g);
        int idx[64];
        float val[64];
        for (int q = 0; q < 4; ++q) {
            const int lane = 4 * g + q;
            const int leftMask = -(1 << std::max(lane / 4 - 1, 0));
            int base = (lane & 3) + 1;
            float opt = kMaxFloat;
            int best = 0;
            const int parts = (1 << (t - 2)) - 1;
            int p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            for (int j = (lane & 3); j < parts; j += 4) {
                const float c = cost[p] + cost[p ^ s];
                if (c < opt) {
                    opt = c;
                    best = p;
                }
                base += 4;
                p = (((base & leftMask) << 1) | (base & ~leftMask)) << 1;
            }
            idx[q] = best;
            val[q] = opt;
        }
        lane_tree(4, idx, val, [](float o, float m) { return o < m; });
        finish(s, val[0], idx[0]);
    }
    {
        const int s = S - 1;
        int idx[64];
        float val[64];
        for (int lane = 0; lane < 32; ++lane) {
            float opt = kMaxFloat;
            int best = 0;
            int p = (lane + 1) * 2;
            for (int j = lane; j < (1 << (t - 1)) - 1; j += 32) {
                const float c = cost[p] + cost[p ^ s];
                if (c < opt) {
                    opt = c;
                    best = p;
                }
                p += 64;
            }
            idx[lane] = best;
            val[lane] = opt;
        }
        lane_tree(32, idx, val, [](float o, float m) { return o < m; });
        finish(s, val[0], idx[0]);
    }
    /* updateTreelet (:198-324). */
    if (!(cost[S - 1] < T.cost[X])) return false;
    int st_node[kTrT], st_mask[kTrT], nst = 1, next = 1;
    st_node[0] = X;
    st_mask[0] = S - 1;
    while (nst > 0) {
        int nn[kTrT], nm[kTrT], k = 0;
        auto place = [&](int root, int part, bool left) {
            int32_t child;
            if (__builtin_popcount(part) > 1) {
                child = internal[next++];
                store4(T.nmin + 4 * (int64_t)child, smn[part]);
                store4(T.nmax + 4 * (int64_t)child, smx[part]);
                T.area[child] = tr_area(smn[part], smx[part]);
                nn[k] = child;
                nm[k] = part;
                ++k;
            } else {
                child = leaves[__builtin_ctz(part)];
            }
            if (left) T.left[root] = child;
            else T.right[root] = child;
            T.parent[child] = root;
            return child;
        };
        std::vector<std::pair<int32_t, int>> lefts, rights;
        for (int i = 0; i < nst; ++i) {
            const int part = mask[st_mask[i]];
            const int32_t c = place(st_node[i], part, true);
            if (__builtin_popcount(part) > 1) T.cost[c] = cost[part];
        }
        for (int i = 0; i < nst; ++i) {
            const int part = mask[st_mask[i]];
            const int comp = part ^ st_mask[i];
            const int32_t c = place(st_node[i], comp, false);
            if (__builtin_popcount(comp) > 1) T.cost[c] = cost[part];   /* the reference's costs[partition] */
        }
        for (int i = 0; i < k; ++i) {
            st_node[i] = nn[i];
            st_mask[i] = nm[i];
        }
        nst = k;
    }
    return true;
}

}  // namespace

extern "C" {

/* TRBuilder::optimize (TRBuilder.cpp:140-191): gamma = 7, doubling; post-order
 * climb per iteration as the ATRBVH restatement's. */
int64_t hxo_tr(int64_t n, int iterations, float ct, float ci, int32_t* left, int32_t* right, int32_t* parent,
               float* nmin, float* nmax) {
    if (n < 2) return 0;
    const int64_t N = 2 * n - 1;
    AtrTree T{n, left, right, parent, nmin, nmax, std::vector<float>(N), std::vector<float>(N),
              std::vector<int32_t>(N)};
    for (int64_t v = 0; v < N; ++v) T.area[v] = atr_area(T.bmin((int32_t)v), T.bmax((int32_t)v));
    int64_t accepted = 0;
    int64_t gamma = kTrT;
    std::vector<int32_t> stack;
    for (int i = 0; i < iterations; ++i) {
        for (int64_t p = 0; p < n; ++p) {
            T.refs[n - 1 + p] = 1;
            T.cost[n - 1 + p] = ci * T.area[n - 1 + p];
        }
        stack.assign(1, 0);
        std::vector<uint8_t> seen(n - 1, 0);
        while (!stack.empty()) {
            const int32_t v = stack.back();
            if (!seen[v]) {
                seen[v] = 1;
                if (!T.is_leaf(T.right[v])) stack.push_back(T.right[v]);
                if (!T.is_leaf(T.left[v])) stack.push_back(T.left[v]);
                continue;
            }
            stack.pop_back();
            const int32_t l = T.left[v], r = T.right[v];
            T.refs[v] = T.refs[l] + T.refs[r];
            T.cost[v] = ct * T.area[v] + T.cost[l] + T.cost[r];
            if (T.refs[v] >= gamma && tr_treelet(T, v, ct, ci)) ++accepted;
        }
        gamma *= 2;
        if (gamma > n) break;
    }
    return accepted;
}

int64_t hxo_atr(int64_t n, int treelet_size, int iterations, float ct, float ci, int32_t* left,
                           int32_t* right, int32_t* parent, float* nmin, float* nmax) {
    if (n < 2 || treelet_size < 4 || treelet_size > 30) return 0;
    const int64_t N = 2 * n - 1;
    AtrTree T{n, left, right, parent, nmin, nmax, std::vector<float>(N), std::vector<float>(N),
              std::vector<int32_t>(N)};
    /* computeSurfaceAreas (:560-577), once per build. */
    for (int64_t v = 0; v < N; ++v) T.area[v] = atr_area(T.bmin((int32_t)v), T.bmax((int32_t)v));
    int64_t accepted = 0;
    int64_t gamma = treelet_size;
    std::vector<int32_t> stack;
    for (int i = 0; i < iterations; ++i) {
        for (int64_t p = 0; p < n; ++p) {
            T.refs[n - 1 + p] = 1;
            T.cost[n - 1 + p] = ci * T.area[n - 1 + p];
        }
        /* Post-order over the current tree (explicit stack: (node, visited)). */
        stack.assign(1, 0);
        std::vector<uint8_t> seen(n - 1, 0);
        while (!stack.empty()) {
            const int32_t v = stack.back();
            if (!seen[v]) {
                seen[v] = 1;
                if (!T.is_leaf(T.right[v])) stack.push_back(T.right[v]);
                if (!T.is_leaf(T.left[v])) stack.push_back(T.left[v]);
                continue;
            }
            stack.pop_back();
            const int32_t l = T.left[v], r = T.right[v];
            T.refs[v] = T.refs[l] + T.refs[r];
            T.cost[v] = ct * T.area[v] + T.cost[l] + T.cost[r];
            if (T.refs[v] >= gamma && atr_treelet(T, v, treelet_size, ct)) ++accepted;
        }
        gamma *= 2;
        if (gamma > n) break;
    }
    return accepted;
}

int64_t hxo_collapse(int64_t n, const int32_t* sv, const int32_t* left, const int32_t* right,
                     const int32_t* parent, const float* nmin, const float* nmax, int adaptive,
                     int max_leaf_size, float ct, float ci, int32_t* nodes, int32_t* tri_out,
                     int64_t* n_interior, int64_t* n_leaves) {
    if (n == 1) {
        /* A single reference: the reference indexes termCounters[-1]
         * (BVHCollapserKernels.cu:29-32).  Emit an interior root whose two
         * children are the one leaf. */
        V3 mn = load4(nmin), mx = load4(nmax);
        int32_t* root = nodes;
        int32_t* leaf = nodes + 16;
        std::memset(nodes, 0, 32 * sizeof(int32_t));
        set_child_box(root, 0, mn, mx);
        set_child_box(root, 1, mn, mx);
        root[12] = ~1; root[13] = ~1; root[14] = 1; root[15] = -1;
        set_child_box(leaf, 0, mn, mx);
        set_child_box(leaf, 1, mn, mx);
        leaf[12] = 0; leaf[13] = 1; leaf[14] = ~1; leaf[15] = 0;
        tri_out[0] = sv[0];
        *n_interior = 1;
        *n_leaves = 1;
        return 2;
    }
    const int64_t N = 2 * n - 1;
    std::vector<int32_t> sizes(N), states(N), leafIdx(n), offsets(N, 0), refOffsets(N, 0);
    std::vector<float> costs(N);
    std::vector<std::atomic<int>> counters(n - 1);
    auto clear = [&]() { for (auto& c : counters) c.store(0, std::memory_order_relaxed); };

    /* computeSizes (BVHCollapserKernels.cu:11-56). */
    clear();
    parallel_for(n, [&](int64_t t) {
        int64_t leafNode = t + (N >> 1);
        sizes[leafNode] = 1;
        std::atomic_thread_fence(std::memory_order_release);
        int64_t node = parent[leafNode];
        while (counters[node].fetch_add(1, std::memory_order_acq_rel) > 0) {
            sizes[node] = sizes[left[node]] + sizes[right[node]];
            std::atomic_thread_fence(std::memory_order_release);
            if (node == 0) break;
            node = parent[node];
        }
    });

    /* computeNodeStatesAdaptive (:58-136) / computeNodeStates (:138-184). */
    clear();
    parallel_for(n, [&](int64_t t) {
        int64_t node = t + n - 1;
        if (adaptive) {
            Box box;
            box.grow(load4(nmin + 4 * node));
            box.grow(load4(nmax + 4 * node));
            costs[node] = ci * box.area();
        }
        states[node] = 0;
        std::atomic_thread_fence(std::memory_order_release);
        node = parent[node];
        while (counters[node].fetch_add(1, std::memory_order_acq_rel) > 0) {
            if (adaptive) {
                Box box;
                box.grow(load4(nmin + 4 * node));
                box.grow(load4(nmax + 4 * node));
                float area = box.area();
                float cost = ct * area + costs[left[node]] + costs[right[node]];
                float costAsLeaf = ci * area * (float)sizes[node];
                if (costAsLeaf < cost) { costs[node] = costAsLeaf; states[node] = 0; }
                else { costs[node] = cost; states[node] = 1; }
            } else {
                states[node] = sizes[node] <= max_leaf_size ? 0 : 1;
            }
            std::atomic_thread_fence(std::memory_order_release);
            if (node == 0) break;
            node = parent[node];
        }
    });
    /* The root is always interior: computeLeafIndices (:205) never stops at the
     * root, and compact (:463-464) places a state-0 root at an unwritten slot. */
    states[0] = 1;

    /* computeLeafIndices (:186-217). */
    parallel_for(n, [&](int64_t t) {
        int64_t node = t + n - 1;
        int64_t li = node;
        int64_t p = parent[li];
        int st = states[p];
        while (p > 0) {
            if (st == 0) li = p;
            node = p;
            p = parent[node];
            st = states[p];
        }
        leafIdx[t] = (int32_t)li;
    });

    /* invalidateCollapsedNodes (:219-266). */
    clear();
    parallel_for(n, [&](int64_t t) {
        int64_t li = leafIdx[t];
        int64_t node = t + n - 1;
        if (node == li) return;
        states[node] = -1;
        node = parent[node];
        while (counters[node].fetch_add(1, std::memory_order_acq_rel) > 0) {
            if (node == li) break;
            states[node] = -1;
            if (node == 0) break;
            node = parent[node];
        }
    });

    /* computeNodeOffsets (:268-355, host loop BVHCollapser.cpp:211-231): BFS by
     * levels; within a level, tasks in order, left child before right.  The
     * reference orders warps by atomic arrival; here the order is sequential. */
    std::vector<int32_t> taskNodes(n);
    int64_t interiorCount = 1, leafCount = 0;
    taskNodes[0] = 0;
    offsets[0] = 0; /* the reference never writes the root slot */
    int64_t taskOffset = 0, numTasks = 1;
    while (numTasks > 0) {
        for (int64_t t = taskOffset; t < taskOffset + numTasks; ++t) {
            int64_t node = taskNodes[t];
            int32_t c[2] = {left[node], right[node]};
            for (int k = 0; k < 2; ++k)
                if (states[c[k]] > 0) {
                    offsets[c[k]] = (int32_t)interiorCount;
                    taskNodes[interiorCount++] = c[k];
                }
            for (int k = 0; k < 2; ++k)
                if (states[c[k]] == 0) offsets[c[k]] = (int32_t)(leafCount++);
        }
        taskOffset += numTasks;
        numTasks = interiorCount - taskOffset;
    }

    /* computeReferenceOffsets (:357-404): leaf triangle ranges in node order. */
    {
        int64_t sum = 0;
        for (int64_t node = 0; node < N; ++node)
            if (states[node] == 0) { refOffsets[node] = (int32_t)sum; sum += sizes[node]; }
    }

    /* compactBin (:435-526). */
    const int64_t nInt = interiorCount;
    parallel_for(N, [&](int64_t node) {
        int st = states[node];
        if (st == -1) return;
        int64_t newIdx = st > 0 ? offsets[node] : offsets[node] + nInt;
        int32_t par = parent[node];
        int32_t sz = sizes[node];
        int32_t l = left[node], r = right[node];
        if (par >= 0) par = offsets[par];
        V3 lmn, lmx, rmn, rmx;
        int32_t li, ri;
        if (st > 0) {
            int sl = states[l], sr = states[r];
            lmn = load4(nmin + 4 * (int64_t)l); lmx = load4(nmax + 4 * (int64_t)l);
            rmn = load4(nmin + 4 * (int64_t)r); rmx = load4(nmax + 4 * (int64_t)r);
            li = sl > 0 ? offsets[l] : ~(int32_t)(offsets[l] + nInt);
            ri = sr > 0 ? offsets[r] : ~(int32_t)(offsets[r] + nInt);
        } else {
            lmn = rmn = load4(nmin + 4 * node);
            lmx = rmx = load4(nmax + 4 * node);
            li = refOffsets[node];
            ri = li + sz;
            sz = ~sz;
        }
        int32_t* out = nodes + 16 * newIdx;
        set_child_box(out, 0, lmn, lmx);
        set_child_box(out, 1, rmn, rmx);
        out[12] = li; out[13] = ri; out[14] = sz; out[15] = par;
    });

    /* reorderTriangleIndices (:406-433): within a leaf, sorted order. */
    for (int64_t t = 0; t < n; ++t) {
        int64_t refIdx = left[t + n - 1];
        int64_t off = refOffsets[leafIdx[t]]++;
        tri_out[off] = sv[refIdx];
    }
    *n_interior = nInt;
    *n_leaves = leafCount;
    return nInt + leafCount;
}

void hxo_woopify(const float* verts, const int32_t* tris, const int32_t* tri_idx, int64_t n, float* woop) {
    parallel_for(n, [&](int64_t r) {
        V3 a, b, c;
        tri_verts(verts, tris, tri_idx[r], a, b, c);
        woop_rows(a, b, c, woop + 12 * r);
    });
}

/* HipBVH::refit (HipBVH.cpp:36-93) with refitLeaves / refitInteriors
 * (HipBVHKernels.cu:240-326).  The reference launches refitLeaves with 4 of
 * its 6 parameters (HipBVH.cpp:43-48); this follows the kernels' intent.
 * Interiors are processed children-first (post-order), the order a counter
 * climb also guarantees; the child box is the child's getBoundingBox(), the
 * grow-union of its two child boxes. */
void hxo_refit_final(int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                     const float* verts, const int32_t* tris) {
    const int64_t nn = n_interior + n_leaves;
    auto node_box = [&](int64_t i) {
        const int32_t* nd = nodes + 16 * i;
        Box b;
        for (int c = 0; c < 2; ++c) {
            Box cb;
            cb.mn = V3{w2f(nd[4 * c + 0]), w2f(nd[4 * c + 2]), w2f(nd[2 * c + 8])};
            cb.mx = V3{w2f(nd[4 * c + 1]), w2f(nd[4 * c + 3]), w2f(nd[2 * c + 9])};
            b.grow(cb.mn);
            b.grow(cb.mx);
        }
        return b;
    };
    parallel_for(n_leaves, [&](int64_t k) {
        int32_t* nd = nodes + 16 * (n_interior + k);
        Box b;
        for (int32_t r = nd[12]; r < nd[13]; ++r) {
            V3 a, bb, c;
            tri_verts(verts, tris, tri_idx[r], a, bb, c);
            b.grow(a); b.grow(bb); b.grow(c);
        }
        set_child_box(nd, 0, b.mn, b.mx);
        set_child_box(nd, 1, b.mn, b.mx);
    });
    std::vector<int64_t> order;
    std::vector<int64_t> stack{0};
    while (!stack.empty()) {
        int64_t i = stack.back();
        stack.pop_back();
        order.push_back(i);
        const int32_t* nd = nodes + 16 * i;
        for (int c = 0; c < 2; ++c)
            if (nd[12 + c] >= 0) stack.push_back(nd[12 + c]);
    }
    for (auto it = order.rbegin(); it != order.rend(); ++it) {
        int32_t* nd = nodes + 16 * (*it);
        for (int c = 0; c < 2; ++c) {
            int64_t child = nd[12 + c] < 0 ? ~(int64_t)nd[12 + c] : nd[12 + c];
            Box cb = node_box(child);
            set_child_box(nd, c, cb.mn, cb.mx);
        }
    }
    (void)nn;
}

double hxo_cost(const int32_t* nodes, int64_t n_nodes, float scene_area, float ct, float ci) {
    double sum = 0.0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        const int32_t* nd = nodes + 16 * i;
        Box box;
        for (int c = 0; c < 2; ++c) {
            V3 mn{w2f(nd[4 * c + 0]), w2f(nd[4 * c + 2]), w2f(nd[2 * c + 8])};
            V3 mx{w2f(nd[4 * c + 1]), w2f(nd[4 * c + 3]), w2f(nd[2 * c + 9])};
            box.grow(mn);
            box.grow(mx);
        }
        float P = box.area() / scene_area;
        float term;
        if (nd[14] < 0) term = ci * P * (float)(~nd[14]);
        else term = ct * P;
        sum += (double)term;
    }
    return sum;
}

int64_t hxo_trace(const int32_t* nodes, const float* woop, const int32_t* tri_idx, const float* rays,
                  int64_t n_rays, int any_hit, int32_t* results, int32_t* stats) {
    int64_t overflows = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflows)
    for (int64_t ri = 0; ri < n_rays; ++ri) {
        RayState s;
        setup_ray(rays + 8 * ri, s);
        int32_t stack[kStackSize];
        stack[0] = kSentinel;
        int sp = 0;
        bool overflow = false;
        int32_t leafAddr = 0, nodeAddr = 0, hitIndex = -1, hitId = -1;
        float hitT = s.tmax, hitU = 0.0f, hitV = 0.0f;
        int32_t nNodes = 0, nTris = 0;
        while (nodeAddr != kSentinel) {
            while ((uint32_t)nodeAddr < (uint32_t)kSentinel) {
                const int32_t* nd = nodes + 16 * (int64_t)nodeAddr;
                const float* f = (const float*)nd;
                nNodes++;
                float c0lox = std::fmaf(f[0], s.idx, -s.oodx);
                float c0hix = std::fmaf(f[1], s.idx, -s.oodx);
                float c0loy = std::fmaf(f[2], s.idy, -s.oody);
                float c0hiy = std::fmaf(f[3], s.idy, -s.oody);
                float c0loz = std::fmaf(f[8], s.idz, -s.oodz);
                float c0hiz = std::fmaf(f[9], s.idz, -s.oodz);
                float c1loz = std::fmaf(f[10], s.idz, -s.oodz);
                float c1hiz = std::fmaf(f[11], s.idz, -s.oodz);
                float c0min = span_begin(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, s.tmin);
                float c0max = span_end(c0lox, c0hix, c0loy, c0hiy, c0loz, c0hiz, hitT);
                float c1lox = std::fmaf(f[4], s.idx, -s.oodx);
                float c1hix = std::fmaf(f[5], s.idx, -s.oodx);
                float c1loy = std::fmaf(f[6], s.idy, -s.oody);
                float c1hiy = std::fmaf(f[7], s.idy, -s.oody);
                float c1min = span_begin(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, s.tmin);
                float c1max = span_end(c1lox, c1hix, c1loy, c1hiy, c1loz, c1hiz, hitT);
                bool swp = c1min < c0min;
                bool t0 = c0max >= c0min;
                bool t1 = c1max >= c1min;
                int32_t cx = nd[12], cy = nd[13];
                if (!t0 && !t1) {
                    nodeAddr = stack[sp--];
                } else {
                    nodeAddr = t0 ? cx : cy;
                    if (t0 && t1) {
                        if (swp) std::swap(nodeAddr, cy);
                        if (sp + 1 >= kStackSize) { overflow = true; nodeAddr = kSentinel; leafAddr = 0; break; }
                        stack[++sp] = cy;
                    }
                }
                if (nodeAddr < 0 && leafAddr >= 0) {
                    leafAddr = nodeAddr;
                    nodeAddr = stack[sp--];
                }
                if (!(leafAddr >= 0)) break; /* one lane: __any() over itself */
            }
            while (leafAddr < 0) {
                const int32_t* leaf = nodes + 16 * (int64_t)(~leafAddr);
                for (int32_t tri = leaf[12]; tri < leaf[13]; ++tri) {
                    nTris++;
                    float t, u, v;
                    if (woop_test(s, woop + 12 * (int64_t)tri, hitT, t, u, v)) {
                        int32_t id = tri_idx[tri];
                        if (t < hitT || id < hitId) {
                            hitT = t; hitIndex = tri; hitId = id; hitU = u; hitV = v;
                            if (any_hit) { nodeAddr = kSentinel; break; }
                        }
                    }
                }
                leafAddr = nodeAddr;
                if (nodeAddr < 0) nodeAddr = stack[sp--];
            }
        }
        int32_t* res = results + 4 * ri;
        if (overflow) overflows++;
        if (hitIndex == -1) {
            res[0] = -1;
            res[1] = f2w(hitT);
        } else {
            res[0] = hitId; res[1] = f2w(hitT); res[2] = f2w(hitU); res[3] = f2w(hitV);
        }
        if (stats) { stats[2 * ri] = nNodes; stats[2 * ri + 1] = nTris; }
    }
    return overflows;
}

void hxo_trace_brute(const float* verts, const int32_t* tris, int64_t n, const float* rays, int64_t n_rays,
                     int32_t* results) {
    std::vector<float> w(12 * n);
    std::vector<int32_t> ident(n);
    for (int64_t i = 0; i < n; ++i) ident[i] = (int32_t)i;
    hxo_woopify(verts, tris, ident.data(), n, w.data());
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ri = 0; ri < n_rays; ++ri) {
        RayState s;
        setup_ray(rays + 8 * ri, s);
        float hitT = s.tmax, hitU = 0.0f, hitV = 0.0f;
        int32_t hitId = -1;
        for (int64_t t = 0; t < n; ++t) {
            float tt, u, v;
            if (woop_test(s, w.data() + 12 * t, hitT, tt, u, v) && (tt < hitT || t < hitId)) {
                hitT = tt; hitId = (int32_t)t; hitU = u; hitV = v;
            }
        }
        int32_t* res = results + 4 * ri;
        res[0] = hitId; res[1] = f2w(hitT);
        if (hitId >= 0) { res[2] = f2w(hitU); res[3] = f2w(hitV); }
    }
}

int64_t hxo_bin_canonical(const int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                          uint32_t* out, int64_t capacity) {
    std::vector<uint32_t> buf;
    buf.reserve((size_t)(n_interior * 13 + n_leaves * 8));
    std::vector<int32_t> stack;
    stack.push_back(0);
    std::vector<int32_t> ids;
    while (!stack.empty()) {
        int32_t idx = stack.back();
        stack.pop_back();
        const int32_t* nd = nodes + 16 * (int64_t)idx;
        if (nd[14] < 0) {
            buf.push_back(1u);
            for (int w : {0, 1, 2, 3, 8, 9}) buf.push_back((uint32_t)nd[w]);
            int32_t b = nd[12], e = nd[13];
            buf.push_back((uint32_t)(e - b));
            ids.assign(tri_idx + b, tri_idx + e);
            std::sort(ids.begin(), ids.end());
            for (int32_t id : ids) buf.push_back((uint32_t)id);
        } else {
            buf.push_back(0u);
            for (int w = 0; w < 12; ++w) buf.push_back((uint32_t)nd[w]);
            int32_t c0 = nd[12] < 0 ? ~nd[12] : nd[12];
            int32_t c1 = nd[13] < 0 ? ~nd[13] : nd[13];
            stack.push_back(c1);
            stack.push_back(c0);
        }
    }
    if (out) {
        int64_t m = std::min<int64_t>(capacity, (int64_t)buf.size());
        std::memcpy(out, buf.data(), sizeof(uint32_t) * m);
    }
    return (int64_t)buf.size();
}

int hxo_validate_bin_refs(const int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                          int64_t n_refs, int64_t n_tris, const float* verts, const int32_t* tris) {
    const bool split = n_tris != n_refs;
    const int64_t nn = n_interior + n_leaves;
    if (n_interior < 1) return 1;
    std::vector<int> seen(nn, 0), refSeen(n_refs, 0);
    for (int64_t i = 0; i < n_interior; ++i) {
        const int32_t* nd = nodes + 16 * i;
        if (nd[14] < 0) return 2;                       /* interior first */
        if (i == 0 && nd[15] != -1) return 3;           /* root parent */
        int64_t sum = 0;
        for (int c = 0; c < 2; ++c) {
            int32_t ci = nd[12 + c];
            int64_t child = ci < 0 ? ~(int64_t)ci : ci;
            if (child <= 0 || child >= nn) return 4;
            if ((ci < 0) != (child >= n_interior)) return 5;  /* leaf flag */
            const int32_t* ch = nodes + 16 * child;
            if (ch[15] != (int32_t)i) return 6;        /* parent pointer */
            bool dup = (c == 1 && nd[12] == nd[13] && n_refs == 1);
            if (!dup) { if (seen[child]) return 7; seen[child] = 1; }
            int32_t sz = ch[14] < 0 ? ~ch[14] : ch[14];
            if (!dup) sum += sz;
            /* child box stored in the parent equals the child's own box */
            for (int w = 0; w < 6; ++w) {
                static const int wi[2][6] = {{0, 1, 2, 3, 8, 9}, {4, 5, 6, 7, 10, 11}};
                float pv = w2f(nd[wi[c][w]]);
                const int32_t* cb = ch;
                float lo = w2f(cb[wi[0][w]]), lo2 = w2f(cb[wi[1][w]]);
                float v = (w % 2 == 0) ? std::min(lo, lo2) : std::max(lo, lo2);
                if (!(pv == v)) return 8;
            }
        }
        if (n_refs > 1 && sum != nd[14]) return 9;      /* sizes */
    }
    for (int64_t i = 1; i < nn; ++i)
        if (!seen[i]) return 10;                        /* every node referenced once */
    for (int64_t i = n_interior; i < nn; ++i) {
        const int32_t* nd = nodes + 16 * i;
        if (nd[14] >= 0) return 11;
        int32_t b = nd[12], e = nd[13];
        if (b < 0 || e > n_refs || e - b != ~nd[14]) return 12;
        V3 mn{w2f(nd[0]), w2f(nd[2]), w2f(nd[8])}, mx{w2f(nd[1]), w2f(nd[3]), w2f(nd[9])};
        for (int32_t r = b; r < e; ++r) {
            if (refSeen[r]) return 13;
            refSeen[r] = 1;
            V3 a, bb, c;
            tri_verts(verts, tris, tri_idx[r], a, bb, c);
            if (split) {
                /* a presplit reference is clipped: its leaf box meets the triangle's box */
                Box tb;
                tb.grow(a);
                tb.grow(bb);
                tb.grow(c);
                if (tb.mx.x < mn.x || tb.mx.y < mn.y || tb.mx.z < mn.z || tb.mn.x > mx.x || tb.mn.y > mx.y ||
                    tb.mn.z > mx.z)
                    return 14;
                continue;
            }
            for (const V3* p : {&a, &bb, &c})
                if (p->x < mn.x || p->y < mn.y || p->z < mn.z || p->x > mx.x || p->y > mx.y || p->z > mx.z)
                    return 14;                          /* leaf box contains its triangles */
        }
    }
    for (int64_t r = 0; r < n_refs; ++r)
        if (!refSeen[r]) return 15;
    /* Every triangle is referenced (HipBVH.cpp:290-303); without presplitting
     * the references are a permutation of the triangle ids. */
    std::vector<int> triSeen(n_tris, 0);
    for (int64_t r = 0; r < n_refs; ++r) {
        int32_t id = tri_idx[r];
        if (id < 0 || id >= n_tris || (!split && triSeen[id])) return 16;
        triSeen[id] = 1;
    }
    for (int64_t t = 0; t < n_tris; ++t)
        if (!triSeen[t]) return 17;
    return 0;
}

int hxo_validate_bin(const int32_t* nodes, int64_t n_interior, int64_t n_leaves, const int32_t* tri_idx,
                     int64_t n_refs, const float* verts, const int32_t* tris) {
    return hxo_validate_bin_refs(nodes, n_interior, n_leaves, tri_idx, n_refs, n_refs, verts, tris);
}

/* ------------------------------------------------------------------------ */
/* GPU presplitter (Presplitter.cpp:124-326, PresplitterKernels.cu:11-419).  */
/* ------------------------------------------------------------------------ */
}  // extern "C"

namespace {

/* x^y for the priority, in double precision from the textbook series (no
 * library call, so that the device and the host run the same operations):
 * ln x = e ln2 + 2 atanh((m-1)/(m+1)) with m in [sqrt(1/2), sqrt(2)); e^y =
 * 2^k e^r with |r| <= ln2/2 and the Taylor series of e^r. */
constexpr double kPsLn2 = 0.69314718055994530942;
constexpr double kPsLn2Hi = 0.693147180369123816490;  /* 21 low bits zero */
constexpr double kPsLn2Lo = 1.90821492927058770002e-10;

double ps_ln(double x) {
    int e = 0;
    double m = std::frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double t = 1.0 / 25.0;
    for (int k = 23; k >= 1; k -= 2) t = t * s2 + 1.0 / (double)k;
    return (double)e * kPsLn2 + 2.0 * s * t;
}

double ps_exp(double y) {
    const double k = std::nearbyint(y / kPsLn2);
    const double r = (y - k * kPsLn2Hi) - k * kPsLn2Lo;
    double p = 1.0;
    for (int j = 20; j >= 1; --j) p = 1.0 + p * r / (double)j;
    return std::ldexp(p, (int)k);
}

float ps_pow(float x, float y) {
    if (!(x > 0.0f) || !(x <= kMaxFloat)) return x == 0.0f ? 0.0f : x;
    return (float)ps_exp((double)y * ps_ln((double)x));
}

float axis_of(const V3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
void set_axis(V3& v, int a, float f) {
    if (a == 0) v.x = f;
    else if (a == 1) v.y = f;
    else v.z = f;
}

/* 21-bit per axis interleave, axis 0 in bit 0 (PresplitterKernels.cu:44-55). */
uint64_t ps_morton63(const int32_t q[3]) {
    uint64_t c = 0;
    for (int i = 0; i < 21; ++i)
        for (int a = 0; a < 3; ++a) c |= (uint64_t)((q[a] >> i) & 1) << (3 * i + a);
    return c;
}

int ps_diff_bit(uint64_t a, uint64_t b) {
    const uint64_t d = a ^ b;
    return d ? 63 - __builtin_clzll(d) : -1;
}

struct PsScene {
    V3 mn, dg, scale;
};

PsScene ps_scene(const float* sb) {
    PsScene s;
    s.mn = V3{sb[0], sb[1], sb[2]};
    s.dg = V3{sb[3] - sb[0], sb[4] - sb[1], sb[5] - sb[2]};
    s.scale = V3{1.0f / s.dg.x, 1.0f / s.dg.y, 1.0f / s.dg.z};
    return s;
}

/* computePriorities (:11-71): (2^i (A_box - A_ideal))^(1/3), i from the
 * highest differing bit of the 63-bit codes of the box corners. */
float ps_priority(const V3& v0, const V3& v1, const V3& v2, const PsScene& S) {
    Box box;
    box.grow(v0);
    box.grow(v1);
    box.grow(v2);
    int32_t qmn[3], qmx[3];
    for (int a = 0; a < 3; ++a) {
        qmn[a] = f2i(((axis_of(box.mn, a) - axis_of(S.mn, a)) * axis_of(S.scale, a)) * 2097151.9f);
        qmx[a] = f2i(((axis_of(box.mx, a) - axis_of(S.mn, a)) * axis_of(S.scale, a)) * 2097151.9f);
    }
    const int diff = ps_diff_bit(ps_morton63(qmn), ps_morton63(qmx));
    const int i = diff / 3 - 21;
    const V3 d1{v0.x - v1.x, v0.y - v1.y, v0.z - v1.z};
    const V3 d2{v0.x - v2.x, v0.y - v2.y, v0.z - v2.z};
    const V3 c{d1.y * d2.z - d1.z * d2.y, d1.z * d2.x - d1.x * d2.z, d1.x * d2.y - d1.y * d2.x};
    const float ideal = std::fabs(c.x) + std::fabs(c.y) + std::fabs(c.z);
    return ps_pow(std::ldexp(1.0f, i) * (box.area() - ideal), 1.0f / 3.0f);
}

struct PsTask {
    int32_t tri, count;
    int64_t off;
    Box box;
};

/* split (:189-419) of one task into two; the budget count-1 is shared by the
 * children's longest() values, which the reference's AABB::longest returns as
 * bool (util/AABB.h:38), i.e. 1 unless the child is a point. */
void ps_split(const PsTask& t, const V3 v[3], const PsScene& S, PsTask& L, PsTask& R) {
    const Box& box = t.box;
    int32_t qmn[3], qmx[3];
    for (int a = 0; a < 3; ++a) {
        qmn[a] = f2i(((axis_of(box.mn, a) - axis_of(S.mn, a)) / axis_of(S.dg, a)) * 2097151.9f);
        qmx[a] = f2i(((axis_of(box.mx, a) - axis_of(S.mn, a)) / axis_of(S.dg, a)) * 2097151.9f);
    }
    const int diff = ps_diff_bit(ps_morton63(qmn), ps_morton63(qmx));
    int axis;
    float pos;
    if (diff >= 0) {
        axis = diff % 3;
        pos = 0.0f;
        float delta = 0.5f;
        for (int i = 20; i >= diff / 3; --i) {
            if ((qmx[axis] >> i) & 1) pos += delta;
            delta *= 0.5f;
        }
        pos = pos * axis_of(S.dg, axis) + axis_of(S.mn, axis);
        if (axis_of(box.mn, axis) >= pos || axis_of(box.mx, axis) <= pos)
            pos = (axis_of(box.mn, axis) + axis_of(box.mx, axis)) * 0.5f;
    } else {
        const V3 d{box.mx.x - box.mn.x, box.mx.y - box.mn.y, box.mx.z - box.mn.z};
        axis = (d.x > d.y && d.x > d.z) ? 0 : (d.y > d.z ? 1 : 2);
        pos = (axis_of(box.mn, axis) + axis_of(box.mx, axis)) * 0.5f;
    }
    Box lb, rb;
    const V3* p1 = &v[2];
    for (int i = 0; i < 3; ++i) {
        const V3* p0 = p1;
        p1 = &v[i];
        const float a = axis_of(*p0, axis), b = axis_of(*p1, axis);
        if (a <= pos) lb.grow(*p0);
        if (a >= pos) rb.grow(*p0);
        if ((a < pos && b > pos) || (a > pos && b < pos)) {
            const float tt = MAXs(MINs((pos - a) / (b - a), 1.0f), 0.0f);
            const float u = 1.0f - tt;
            const V3 q{p0->x * u + p1->x * tt, p0->y * u + p1->y * tt, p0->z * u + p1->z * tt};
            lb.grow(q);
            rb.grow(q);
        }
    }
    set_axis(lb.mx, axis, pos);
    set_axis(rb.mn, axis, pos);
    for (Box* x : {&lb, &rb}) {
        x->mn = V3{MAXs(x->mn.x, box.mn.x), MAXs(x->mn.y, box.mn.y), MAXs(x->mn.z, box.mn.z)};
        x->mx = V3{MINs(x->mx.x, box.mx.x), MINs(x->mx.y, box.mx.y), MINs(x->mx.z, box.mx.z)};
    }
    auto longest = [](const Box& x) {
        const float m = std::fmax(std::fmax(x.mx.x - x.mn.x, x.mx.y - x.mn.y), x.mx.z - x.mn.z);
        return m != 0.0f ? 1.0f : 0.0f;
    };
    const float ll = longest(lb), rl = longest(rb);
    const int32_t lc = f2i(((float)(t.count - 1) * ll) / (ll + rl) + 0.5f);
    L = PsTask{t.tri, lc, t.off, lb};
    R = PsTask{t.tri, t.count - 1 - lc, t.off + lc + 1, rb};
}

/* Exact fixed-point sum (2^-32 units, truncated per term) of the priorities,
 * floored: the reference's float atomic sum truncated to int. */
int64_t ps_sum_floor(const std::vector<float>& p) {
    uint64_t s = 0;
    for (float x : p)
        if (x >= 0.0f && x <= 1.0e9f) s += (uint64_t)((double)x * 4294967296.0);
    return (int64_t)(s >> 32);
}

int64_t ps_sum_round(const std::vector<float>& p, float D) {
    int64_t s = 0;
    for (float x : p) s += f2i(D * x);
    return s;
}

}  // namespace

extern "C" {

float hxo_ps_pow(float x, float y) { return ps_pow(x, y); }

/* One split of a task (triangle v[9], box mn.xyz mx.xyz, budget count) in the
 * scene box sb; writes the children's boxes and budgets. */
void hxo_ps_split(const float* v, const float* box, int32_t count, const float* sb, float* lbox, float* rbox,
                  int32_t* lcount, int32_t* rcount) {
    const PsScene S = ps_scene(sb);
    const V3 tv[3] = {load3(v), load3(v + 3), load3(v + 6)};
    PsTask t{0, count, 0, Box{}};
    t.box.mn = load3(box);
    t.box.mx = load3(box + 3);
    PsTask l, r;
    ps_split(t, tv, S, l, r);
    const float lb[6] = {l.box.mn.x, l.box.mn.y, l.box.mn.z, l.box.mx.x, l.box.mx.y, l.box.mx.z};
    const float rb[6] = {r.box.mn.x, r.box.mn.y, r.box.mn.z, r.box.mx.x, r.box.mx.y, r.box.mx.z};
    std::memcpy(lbox, lb, sizeof(lb));
    std::memcpy(rbox, rb, sizeof(rb));
    *lcount = l.count;
    *rcount = r.count;
}

int64_t hxo_presplit(const float* verts, const int32_t* tris, int64_t n, const float* scene_box, float beta,
                     float* ref_min, float* ref_max, int32_t* ref_tri, int64_t cap, float* D_out,
                     float* priorities) {
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n, sb);
    const PsScene S = ps_scene(sb);
    std::vector<float> prio(n);
    auto tri_v = [&](int64_t t, V3 v[3]) {
        for (int k = 0; k < 3; ++k) v[k] = load3(verts + 3 * (int64_t)tris[3 * t + k]);
    };
    parallel_for(n, [&](int64_t t) {
        V3 v[3];
        tri_v(t, v);
        prio[t] = ps_priority(v[0], v[1], v[2], S);
    });
    if (priorities) std::copy(prio.begin(), prio.end(), priorities);
    /* computeD (Presplitter.cpp:124-155), with Smax = (int)(beta * n). */
    const int64_t smax = f2i(beta * (float)n);
    const int64_t s1 = ps_sum_floor(prio);
    float D = 0.0f;
    if (s1 > 0) {
        float dmin = (float)(smax / s1);
        int64_t s = ps_sum_round(prio, dmin);
        float dmax = (dmin * (float)smax) / (float)s;
        for (int r = 0; r < 6; ++r) {
            const float davg = 0.5f * (dmin + dmax);
            s = ps_sum_round(prio, davg);
            if (s > smax) dmax = davg;
            else dmin = davg;
        }
        D = dmin;
    }
    if (D_out) *D_out = D;
    std::vector<int64_t> off(n);
    int64_t total = 0;
    for (int64_t t = 0; t < n; ++t) {
        off[t] = total;
        total += (int64_t)f2i(D * prio[t]) + 1;
    }
    if (!ref_tri || total > cap) return total;
    /* Each triangle's references are the leaves of its split tree in
     * depth-first order, at the triangle's offset. */
    parallel_for(n, [&](int64_t t) {
        V3 v[3];
        tri_v(t, v);
        Box b;
        b.grow(v[0]);
        b.grow(v[1]);
        b.grow(v[2]);
        std::vector<PsTask> stack{PsTask{(int32_t)t, f2i(D * prio[t]), off[t], b}};
        while (!stack.empty()) {
            const PsTask k = stack.back();
            stack.pop_back();
            if (k.count <= 0) {
                store4(ref_min + 4 * k.off, k.box.mn);
                store4(ref_max + 4 * k.off, k.box.mx);
                ref_tri[k.off] = k.tri;
                continue;
            }
            PsTask l, r;
            ps_split(k, v, S, l, r);
            stack.push_back(r);
            stack.push_back(l);
        }
    });
    return total;
}

}  // extern "C"

namespace {

/* setupReferences (LBVHBuilder.cpp:26-74): the triangles' boxes, or the
 * presplitter's references (beta >= 0). */
struct RefSet {
    int64_t n = 0;
    bool split = false;
    std::vector<float> mn, mx;
    std::vector<int32_t> tri;
};

RefSet setup_refs(const float* verts, const int32_t* tris, int64_t n_tri, const float sb[6], float beta) {
    RefSet r;
    if (!(beta >= 0.0f)) {
        r.n = n_tri;
        r.mn.resize(4 * n_tri);
        r.mx.resize(4 * n_tri);
        hxo_setup_boxes(verts, tris, n_tri, r.mn.data(), r.mx.data());
        return r;
    }
    r.split = true;
    r.n = hxo_presplit(verts, tris, n_tri, sb, beta, nullptr, nullptr, nullptr, 0, nullptr, nullptr);
    r.mn.resize(4 * r.n);
    r.mx.resize(4 * r.n);
    r.tri.resize(r.n);
    hxo_presplit(verts, tris, n_tri, sb, beta, r.mn.data(), r.mx.data(), r.tri.data(), r.n, nullptr, nullptr);
    return r;
}

/* Leaf slots hold reference ids until the end; the output holds triangles. */
void refs_to_tris(const RefSet& r, int32_t* tri_out) {
    if (!r.split) return;
    for (int64_t i = 0; i < r.n; ++i) tri_out[i] = r.tri[tri_out[i]];
}

}  // namespace

extern "C" {

int64_t hxo_build_lbvh(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                       int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes, int32_t* tri_out,
                       float* woop, int64_t* n_interior, int64_t* n_leaves, double sec[8], float presplit_beta) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    auto lap = [&](int i) {
        auto t1 = clk::now();
        sec[i] = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
    };
    for (int i = 0; i < 8; ++i) sec[i] = 0.0;
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    lap(0);
    RefSet R = setup_refs(verts, tris, n_tri, sb, presplit_beta);
    const int64_t n = R.n;
    std::vector<float>& rmin = R.mn;
    std::vector<float>& rmax = R.mx;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(rmin.data(), rmax.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    lap(1);
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    lap(2);
    const int64_t N = 2 * n - 1;
    std::vector<int32_t> left(N), right(N), parent(N);
    hxo_construct(keys.data(), n, left.data(), right.data(), parent.data());
    lap(3);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    hxo_refit(vals.data(), rmin.data(), rmax.data(), n, left.data(), right.data(), parent.data(), nmin.data(),
              nmax.data());
    lap(4);
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    lap(5);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    lap(6);
    return nn;
}

/* PLOCBuilder::build (PLOCBuilder.cpp:161-208): LBVH front half, clustering,
 * collapse, woopify.  Returns the node count; sec[0..6] as hxo_build_lbvh. */
int64_t hxo_build_ploc(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                       int radius, int max_radius, int ploc_adaptive, int adaptive, int max_leaf_size, float ct,
                       float ci, int32_t* nodes, int32_t* tri_out, float* woop, int64_t* n_interior,
                       int64_t* n_leaves, double sec[8], float presplit_beta) {
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    auto lap = [&](int i) {
        auto t1 = clk::now();
        sec[i] = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
    };
    for (int i = 0; i < 8; ++i) sec[i] = 0.0;
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    lap(0);
    RefSet R = setup_refs(verts, tris, n_tri, sb, presplit_beta);
    const int64_t n = R.n;
    std::vector<float>& rmin = R.mn;
    std::vector<float>& rmax = R.mx;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(rmin.data(), rmax.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    lap(1);
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    lap(2);
    const int64_t N = 2 * n - 1;
    std::vector<int32_t> left(N), right(N), parent(N);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    parallel_for(n, [&](int64_t i) {  /* setupLeaves */
        std::memcpy(nmin.data() + 4 * (n - 1 + i), rmin.data() + 4 * (int64_t)vals[i], 16);
        std::memcpy(nmax.data() + 4 * (n - 1 + i), rmax.data() + 4 * (int64_t)vals[i], 16);
    });
    if (hxo_ploc(n, radius, max_radius, ploc_adaptive, left.data(), right.data(), parent.data(), nmin.data(),
                 nmax.data()) < 0)
        return -1;
    lap(3);
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    lap(5);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    lap(6);
    return nn;
}

/* HLBVHBuilder::build (HLBVHBuilder.cpp:61-437, HLBVHBuilderKernels.cu:35-704):
 * the LBVH below clusters of references that share the top sah_bits Morton
 * bits (a run of equal prefixes is exactly one Karras subtree), and a binned
 * SAH split over the clusters above them, in breadth-first task order.  Bins
 * are the fixed grid cells of the Morton prefix (HLBVHBin.h:14-44).  Where the
 * reference numbers clusters and median-split members with atomics, the order
 * here is the clusters' Morton order.  Upper nodes reuse the ids of the Karras
 * nodes that span clusters (sorted), so the root stays 0.  Returns the number
 * of clusters; with a single cluster the tree is the LBVH. */
static int64_t hlbvh_upper(int64_t n, int bits, int sah_bits, const uint64_t* keys, int32_t* left, int32_t* right,
                           int32_t* parent, float* nmin, float* nmax) {
    const int64_t N = 2 * n - 1;
    /* Karras ranges. */
    std::vector<int32_t> first(N), last(N);
    {
        std::vector<int32_t> order;
        order.reserve(N);
        std::vector<int32_t> st(1, 0);
        while (!st.empty()) {
            const int32_t v = st.back();
            st.pop_back();
            order.push_back(v);
            if (v < n - 1) {
                st.push_back(left[v]);
                st.push_back(right[v]);
            }
        }
        for (int64_t k = (int64_t)order.size() - 1; k >= 0; --k) {
            const int32_t v = order[k];
            if (v >= n - 1) {
                first[v] = last[v] = (int32_t)(v - (n - 1));
            } else {
                first[v] = std::min(first[left[v]], first[right[v]]);
                last[v] = std::max(last[left[v]], last[right[v]]);
            }
        }
    }
    const int shift = bits - sah_bits;
    auto prefix = [&](int64_t p) { return (uint32_t)(keys[p] >> shift); };
    std::vector<int32_t> cl_of(n);
    std::vector<int32_t> ca, cb;
    for (int64_t p = 0; p < n; ++p) {
        if (p == 0 || prefix(p) != prefix(p - 1)) {
            ca.push_back((int32_t)p);
            cb.push_back((int32_t)p);
        }
        cb.back() = (int32_t)p;
        cl_of[p] = (int32_t)ca.size() - 1;
    }
    const int64_t C = (int64_t)ca.size();
    if (C == 1) return 1;
    /* Cluster roots: the Karras node whose range is the run. */
    std::vector<int32_t> croot(C, -1);
    for (int64_t c = 0; c < C; ++c)
        if (ca[c] == cb[c]) croot[c] = (int32_t)(n - 1 + ca[c]);
    for (int64_t v = 0; v < n - 1; ++v) {
        const int32_t c = cl_of[first[v]];
        if (first[v] == ca[c] && last[v] == cb[c]) croot[c] = (int32_t)v;
    }
    /* Upper ids: Karras internal nodes spanning clusters, ascending. */
    std::vector<int32_t> upper;
    for (int64_t v = 0; v < n - 1; ++v)
        if (cl_of[first[v]] != cl_of[last[v]]) upper.push_back((int32_t)v);
    /* Cluster bin coordinates (mortonCodeToCoords, :10-19). */
    const int nbins = 1 << ((sah_bits + 2) / 3);
    std::vector<int32_t> coord(3 * C);
    std::vector<V3> cmn(C), cmx(C);
    for (int64_t c = 0; c < C; ++c) {
        const uint32_t code = prefix(ca[c]);
        int co[3] = {0, 0, 0};
        for (int i = 0; i < 10; ++i)
            for (int a = 0; a < 3; ++a) co[a] += ((code >> (3 * i + a)) & 1) << i;
        for (int a = 0; a < 3; ++a) coord[3 * c + a] = co[a];
        cmn[c] = load4(nmin + 4 * (int64_t)croot[c]);
        cmx[c] = load4(nmax + 4 * (int64_t)croot[c]);
    }
    struct Bin {
        V3 mn{kMaxFloat, kMaxFloat, kMaxFloat}, mx{-kMaxFloat, -kMaxFloat, -kMaxFloat};
        int count = 0;
        void include(const Bin& o) {
            mn = V3{std::min(mn.x, o.mn.x), std::min(mn.y, o.mn.y), std::min(mn.z, o.mn.z)};
            mx = V3{std::max(mx.x, o.mx.x), std::max(mx.y, o.mx.y), std::max(mx.z, o.mx.z)};
            count += o.count;
        }
        float area() const {
            const float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
            return 2.0f * (dx * dy + dx * dz + dy * dz);
        }
        float cost() const { return (float)count * area(); }
    };
    std::vector<int32_t> members(C);
    for (int64_t c = 0; c < C; ++c) members[c] = (int32_t)c;
    struct Task {
        int32_t k, s, e;
    };
    std::vector<Task> level(1, Task{0, 0, (int32_t)C}), next;
    int32_t next_k = 1;
    std::vector<int32_t> tmp;
    parent[0] = -1;
    while (!level.empty()) {
        next.clear();
        for (const Task& tk : level) {
            std::vector<Bin> bins(3 * nbins);
            for (int32_t i = tk.s; i < tk.e; ++i) {
                const int32_t c = members[i];
                for (int a = 0; a < 3; ++a) {
                    Bin& b = bins[a * nbins + coord[3 * c + a]];
                    b.mn = V3{std::min(b.mn.x, cmn[c].x), std::min(b.mn.y, cmn[c].y), std::min(b.mn.z, cmn[c].z)};
                    b.mx = V3{std::max(b.mx.x, cmx[c].x), std::max(b.mx.y, cmx[c].y), std::max(b.mx.z, cmx[c].z)};
                    b.count += 1;
                }
            }
            /* split (:395-497): first strict minimum over axes, then bins. */
            float best = kMaxFloat;
            int best_axis = -1, best_bin = -1;
            Bin lbest, rbest, all;
            for (int a = 0; a < 3; ++a) {
                std::vector<Bin> right_acc(nbins);
                right_acc[nbins - 1] = bins[a * nbins + nbins - 1];
                for (int b = nbins - 2; b >= 0; --b) {
                    right_acc[b] = right_acc[b + 1];
                    right_acc[b].include(bins[a * nbins + b]);
                }
                Bin acc = bins[a * nbins];
                for (int b = 0; b < nbins - 1; ++b) {
                    if (acc.count > 0 && right_acc[b + 1].count > 0) {
                        const float cst = acc.cost() + right_acc[b + 1].cost();
                        if (cst < best) {
                            best = cst;
                            best_axis = a;
                            best_bin = b;
                            lbest = acc;
                            rbest = right_acc[b + 1];
                        }
                    }
                    acc.include(bins[a * nbins + b + 1]);
                }
                all = acc;
            }
            const int32_t cnt = tk.e - tk.s;
            int32_t nl;
            tmp.assign(members.begin() + tk.s, members.begin() + tk.e);
            if (best == kMaxFloat) {
                /* every cluster in one bin on every axis: median by order */
                nl = cnt / 2;
                lbest = all;
                rbest = all;
            } else {
                int32_t w = tk.s;
                for (int32_t c : tmp)
                    if (coord[3 * c + best_axis] <= best_bin) members[w++] = c;
                nl = w - tk.s;
                for (int32_t c : tmp)
                    if (coord[3 * c + best_axis] > best_bin) members[w++] = c;
            }
            const int32_t v = upper[tk.k];
            Box bx;
            bx.grow(lbest.mn);
            bx.grow(lbest.mx);
            bx.grow(rbest.mn);
            bx.grow(rbest.mx);
            store4(nmin + 4 * (int64_t)v, bx.mn);
            store4(nmax + 4 * (int64_t)v, bx.mx);
            const int32_t ranges[2][2] = {{tk.s, tk.s + nl}, {tk.s + nl, tk.e}};
            int32_t child[2];
            for (int side = 0; side < 2; ++side) {
                const int32_t s = ranges[side][0], e = ranges[side][1];
                if (e - s == 1) {
                    child[side] = croot[members[s]];
                } else {
                    const int32_t k = next_k++;
                    child[side] = upper[k];
                    next.push_back(Task{k, s, e});
                }
                parent[child[side]] = v;
            }
            left[v] = child[0];
            right[v] = child[1];
        }
        std::swap(level, next);
    }
    return C;
}

int64_t hxo_hlbvh(int64_t n, int bits, int sah_bits, const uint64_t* sorted_keys, int32_t* left, int32_t* right,
                  int32_t* parent, float* nmin, float* nmax) {
    if (n < 2) return 1;
    return hlbvh_upper(n, bits, sah_bits, sorted_keys, left, right, parent, nmin, nmax);
}

int64_t hxo_build_hlbvh(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                        int sah_bits, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                        int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves, int64_t* clusters, float presplit_beta) {
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    RefSet R = setup_refs(verts, tris, n_tri, sb, presplit_beta);
    const int64_t n = R.n;
    std::vector<float>& rmin = R.mn;
    std::vector<float>& rmax = R.mx;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(rmin.data(), rmax.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    const int64_t N = std::max<int64_t>(2 * n - 1, 1);
    std::vector<int32_t> left(N), right(N), parent(N);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    if (n > 1) hxo_construct(keys.data(), n, left.data(), right.data(), parent.data());
    hxo_refit(vals.data(), rmin.data(), rmax.data(), n, left.data(), right.data(), parent.data(), nmin.data(),
              nmax.data());
    const int64_t C = n > 1 ? hlbvh_upper(n, bits, sah_bits, keys.data(), left.data(), right.data(), parent.data(),
                                          nmin.data(), nmax.data())
                            : 1;
    if (clusters) *clusters = C;
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    return nn;
}


/* ------------------------------------------------------------------------ */
/* Insertion (InsertionBuilder.cpp:22-274, InsertionBuilderKernels.cu:11-515):
 * parallel reinsertion.  Each step, every node i > 0 with i = mod*k + step
 * finds its best new sibling by the branch-and-bound search of findBestNode;
 * a node keeps its move only if it holds the lock (max of area-reduction bits
 * and index) on every node the move touches; kept moves are disjoint and are
 * applied; the tree is refit and its SAH measured.  The loop ends when the
 * cost stops changing (|dC| <= 0.01) at mod 1, and the best tree is kept.
 * The cost sum is FP64 in node order (the reference adds floats atomically). */
}  // extern "C"

namespace {

/* Termination guard: the reference loops until convergence. */
constexpr int kInsMaxSteps = 1000;

struct InsTree {
    int64_t n, N;
    std::vector<int32_t> left, right, parent;
    std::vector<float> nmin, nmax;  /* float4 per node */
    Box box(int32_t v) const {
        Box b;
        b.mn = load4(nmin.data() + 4 * (int64_t)v);
        b.mx = load4(nmax.data() + 4 * (int64_t)v);
        return b;
    }
};

inline Box box_union(const Box& a, const Box& b) {
    Box r;
    r.grow(a.mn);
    r.grow(a.mx);
    r.grow(b.mn);
    r.grow(b.mx);
    return r;
}

/* findBestNode (:11-131) for one node. */
void ins_find_best(const InsTree& T, int32_t in, int32_t& best_out, float& best_red) {
    const int32_t p = T.parent[in];
    const Box inBox = T.box(in);
    const float pArea = T.box(p).area();
    const float bound = pArea - inBox.area();
    Box highest;
    const int32_t sib = T.left[p] == in ? T.right[p] : T.left[p];
    int32_t out = sib, highestIdx = p;
    bool down = true;
    int32_t best = -1;
    float bestRed = 0.0f, red = 0.0f;
    while (true) {
        Box outBox = T.box(out);
        Box merged = box_union(inBox, outBox);
        if (down) {
            if (out != sib) {
                const float direct = pArea - merged.area();
                if (bestRed < direct + red) {
                    bestRed = direct + red;
                    best = out;
                }
            }
            const float cur = outBox.area() - merged.area();
            red += cur;
            if (out >= T.n - 1 || red + bound <= bestRed) down = false;
            else out = T.left[out];
        } else {
            int32_t op = T.parent[out];
            const float cur = outBox.area() - merged.area();
            red -= cur;
            if (op == highestIdx) {
                highest.grow(outBox.mn);
                highest.grow(outBox.mx);
                out = op;
                if (out != p && out != 0) {
                    merged = box_union(inBox, highest);
                    const float direct = pArea - merged.area();
                    if (bestRed < direct + red) {
                        bestRed = direct + red;
                        best = out;
                    }
                    outBox = T.box(out);
                    const float cur2 = outBox.area() - highest.area();
                    red += cur2;
                }
                op = T.parent[out];
                if (op < 0) break;
                highestIdx = op;
                down = true;
                out = T.left[highestIdx] == out ? T.right[highestIdx] : T.left[highestIdx];
            } else {
                if (T.left[op] == out) {
                    down = true;
                    out = T.right[op];
                } else {
                    out = op;
                }
            }
        }
    }
    best_out = best;
    best_red = bestRed;
}

/* The nodes a move of `in` to `out` touches (lockNodes, :133-185). */
int ins_lock_set(const InsTree& T, int32_t in, int32_t out, int32_t set[8]) {
    int k = 0;
    const int32_t p = T.parent[in];
    const int32_t sib = T.left[p] == in ? T.right[p] : T.left[p];
    set[k++] = in;
    set[k++] = sib;
    set[k++] = p;
    if (p > 0) {
        set[k++] = T.parent[p];
    } else if (sib < T.n - 1) {
        set[k++] = T.left[sib];
        set[k++] = T.right[sib];
    }
    set[k++] = out;
    if (out > 0) {
        set[k++] = T.parent[out];
    } else {
        set[k++] = T.left[out];
        set[k++] = T.right[out];
    }
    return k;
}

/* reinsert (:237-300). */
void ins_reinsert(InsTree& T, int32_t in, int32_t out) {
    int32_t p = T.parent[in];
    const int32_t sib = T.left[p] == in ? T.right[p] : T.left[p];
    if (p != 0) {
        const int32_t pp = T.parent[p];
        if (T.left[pp] == p) T.left[pp] = sib;
        else T.right[pp] = sib;
        T.parent[sib] = pp;
    } else {
        const int32_t sl = T.left[sib], sr = T.right[sib];
        T.left[0] = sl;
        T.right[0] = sr;
        T.parent[sl] = 0;
        T.parent[sr] = 0;
        p = sib;
    }
    if (out != 0) {
        const int32_t op = T.parent[out];
        if (T.left[op] == out) T.left[op] = p;
        else T.right[op] = p;
        T.parent[p] = op;
        T.left[p] = out;
        T.right[p] = in;
        T.parent[out] = p;
        T.parent[in] = p;
    } else {
        const int32_t ol = T.left[0], orr = T.right[0];
        T.left[p] = ol;
        T.right[p] = orr;
        T.parent[ol] = p;
        T.parent[orr] = p;
        T.parent[p] = 0;
        T.parent[in] = 0;
        T.left[0] = in;
        T.right[0] = p;
    }
}

void ins_refit(InsTree& T) {
    std::vector<int32_t> order;
    order.reserve(T.N);
    std::vector<int32_t> st(1, 0);
    while (!st.empty()) {
        const int32_t v = st.back();
        st.pop_back();
        order.push_back(v);
        if (v < T.n - 1) {
            st.push_back(T.left[v]);
            st.push_back(T.right[v]);
        }
    }
    for (int64_t k = (int64_t)order.size() - 1; k >= 0; --k) {
        const int32_t v = order[k];
        if (v >= T.n - 1) continue;
        const Box l = T.box(T.left[v]), r = T.box(T.right[v]);
        Box b;
        b.grow(l.mn);
        b.grow(r.mn);
        b.grow(l.mx);
        b.grow(r.mx);
        store4(T.nmin.data() + 4 * (int64_t)v, b.mn);
        store4(T.nmax.data() + 4 * (int64_t)v, b.mx);
    }
}

/* computeCost (:457-515) as an exact sum: each node's term c * (area / scene
 * area) is truncated to a multiple of 2^-32 and the multiples are added as
 * integers (modulo 2^64), so the total does not depend on the order of the
 * additions (the reference adds float block sums atomically).  A term that is
 * not a number in [0, 1e9] counts as 0. */
uint64_t ins_fixed(float t) {
    if (!(t >= 0.0f) || !(t <= 1.0e9f)) return 0;
    return (uint64_t)((double)t * 4294967296.0);
}

double ins_cost(const InsTree& T, float sceneArea, float ct, float ci) {
    uint64_t s = 0;
    for (int64_t v = 0; v < T.N; ++v) {
        const float P = T.box((int32_t)v).area() / sceneArea;
        s += ins_fixed((v >= T.n - 1 ? ci : ct) * P);
    }
    return (double)s * (1.0 / 4294967296.0);
}

}  // namespace

extern "C" {

void hxo_ins_find(int64_t n, const int32_t* left, const int32_t* right, const int32_t* parent, const float* nmin,
                  const float* nmax, int32_t* out, float* red) {
    const int64_t N = 2 * n - 1;
    InsTree T{n, N, std::vector<int32_t>(left, left + N), std::vector<int32_t>(right, right + N),
              std::vector<int32_t>(parent, parent + N), std::vector<float>(nmin, nmin + 4 * N),
              std::vector<float>(nmax, nmax + 4 * N)};
    out[0] = -1;
    red[0] = 0.0f;
    for (int64_t i = 1; i < N; ++i) ins_find_best(T, (int32_t)i, out[i], red[i]);
}

int64_t hxo_insertion(int64_t n, int mod, float ct, float ci, int32_t* left, int32_t* right, int32_t* parent,
                      float* nmin, float* nmax, int64_t* steps_out) {
    if (n < 2) {
        if (steps_out) *steps_out = 0;
        return 0;
    }
    const int64_t N = 2 * n - 1;
    InsTree T{n, N, std::vector<int32_t>(left, left + N), std::vector<int32_t>(right, right + N),
              std::vector<int32_t>(parent, parent + N), std::vector<float>(nmin, nmin + 4 * N),
              std::vector<float>(nmax, nmax + 4 * N)};
    const float sceneArea = T.box(0).area();
    std::vector<int32_t> bl = T.left, br = T.right, bp = T.parent;
    int steps = 0, modCur = std::max(mod, 1);
    double cost = 3.402823466e+38, bestCost = 3.402823466e+38, prevCost;
    int64_t total = 0;
    std::vector<int32_t> outv(N, -1);
    std::vector<float> redv(N, 0.0f);
    std::vector<uint64_t> locks(N);
    while (steps < kInsMaxSteps) {
        prevCost = cost;
        const int rem = steps % modCur;
        const int64_t cnt = (N - rem + modCur - 1) / modCur;
        parallel_for(cnt, [&](int64_t k) {
            const int64_t i = rem + k * modCur;
            outv[i] = -1;
            if (i > 0) ins_find_best(T, (int32_t)i, outv[i], redv[i]);
        });
        std::fill(locks.begin(), locks.end(), 0);
        int32_t set[8];
        for (int64_t i = rem; i < N; i += modCur) {
            if (i == 0 || outv[i] < 0) continue;
            uint32_t rb;
            std::memcpy(&rb, &redv[i], 4);
            const uint64_t key = ((uint64_t)rb << 32) | (uint64_t)i;
            const int k = ins_lock_set(T, (int32_t)i, outv[i], set);
            for (int j = 0; j < k; ++j) locks[set[j]] = std::max(locks[set[j]], key);
        }
        std::vector<int32_t> kept;
        for (int64_t i = rem; i < N; i += modCur) {
            if (i == 0 || outv[i] < 0) continue;
            uint32_t rb;
            std::memcpy(&rb, &redv[i], 4);
            const uint64_t key = ((uint64_t)rb << 32) | (uint64_t)i;
            const int k = ins_lock_set(T, (int32_t)i, outv[i], set);
            bool ok = true;
            for (int j = 0; j < k; ++j) ok = ok && locks[set[j]] == key;
            if (ok) kept.push_back((int32_t)i);
        }
        for (int32_t i : kept) ins_reinsert(T, i, outv[i]);
        total += (int64_t)kept.size();
        ins_refit(T);
        cost = ins_cost(T, sceneArea, ct, ci);
        if (bestCost > cost) {
            bestCost = cost;
            bl = T.left;
            br = T.right;
            bp = T.parent;
        }
        ++steps;
        const double BETA = 0.01;
        if ((std::fabs(prevCost - cost) <= BETA || kept.empty()) && modCur == 1) break;
        if (std::fabs(prevCost - cost) <= BETA) modCur = std::max(1, modCur - 1);
    }
    T.left = bl;
    T.right = br;
    T.parent = bp;
    ins_refit(T);
    std::copy(T.left.begin(), T.left.end(), left);
    std::copy(T.right.begin(), T.right.end(), right);
    std::copy(T.parent.begin(), T.parent.end(), parent);
    std::copy(T.nmin.begin(), T.nmin.end(), nmin);
    std::copy(T.nmax.begin(), T.nmax.end(), nmax);
    if (steps_out) *steps_out = steps;
    return total;
}

int64_t hxo_build_insertion(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                            int mod, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                            int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves,
                            int64_t* steps_out, float presplit_beta) {
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    RefSet R = setup_refs(verts, tris, n_tri, sb, presplit_beta);
    const int64_t n = R.n;
    std::vector<float>& rmin = R.mn;
    std::vector<float>& rmax = R.mx;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(rmin.data(), rmax.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    const int64_t N = std::max<int64_t>(2 * n - 1, 1);
    std::vector<int32_t> left(N), right(N), parent(N);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    if (n > 1) hxo_construct(keys.data(), n, left.data(), right.data(), parent.data());
    hxo_refit(vals.data(), rmin.data(), rmax.data(), n, left.data(), right.data(), parent.data(), nmin.data(),
              nmax.data());
    hxo_insertion(n, mod, ct, ci, left.data(), right.data(), parent.data(), nmin.data(), nmax.data(), steps_out);
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    return nn;
}

/* ATRBuilder::build (ATRBuilder.cpp:179-236): LBVH front half and hierarchy,
 * refit, ATR optimisation, collapse, woopify. */
int64_t hxo_build_atr(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                      int treelet_size, int iterations, int adaptive, int max_leaf_size, float ct, float ci,
                      int32_t* nodes, int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves,
                      int64_t* accepted, float presplit_beta) {
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    RefSet R = setup_refs(verts, tris, n_tri, sb, presplit_beta);
    const int64_t n = R.n;
    std::vector<float>& rmin = R.mn;
    std::vector<float>& rmax = R.mx;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(rmin.data(), rmax.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    const int64_t N = std::max<int64_t>(2 * n - 1, 1);
    std::vector<int32_t> left(N), right(N), parent(N);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    if (n > 1) hxo_construct(keys.data(), n, left.data(), right.data(), parent.data());
    hxo_refit(vals.data(), rmin.data(), rmax.data(), n, left.data(), right.data(), parent.data(), nmin.data(),
              nmax.data());
    const int64_t acc = hxo_atr(n, treelet_size, iterations, ct, ci, left.data(), right.data(), parent.data(),
                                nmin.data(), nmax.data());
    if (accepted) *accepted = acc;
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    return nn;
}

/* TRBuilder::build (TRBuilder.cpp:193-245): LBVH front half and hierarchy,
 * refit, TR optimisation, collapse, woopify. */
int64_t hxo_build_tr(const float* verts, const int32_t* tris, int64_t n_tri, const float* scene_box, int bits,
                     int iterations, int adaptive, int max_leaf_size, float ct, float ci, int32_t* nodes,
                     int32_t* tri_out, float* woop, int64_t* n_interior, int64_t* n_leaves, int64_t* accepted) {
    float sb[6];
    if (scene_box) std::memcpy(sb, scene_box, sizeof(sb));
    else hxo_scene_box(verts, tris, n_tri, sb);
    RefSet R = setup_refs(verts, tris, n_tri, sb, -1.0f);
    const int64_t n = R.n;
    std::vector<uint64_t> keys(n);
    std::vector<int32_t> vals(n);
    hxo_morton_codes(R.mn.data(), R.mx.data(), n, sb, bits, keys.data());
    for (int64_t i = 0; i < n; ++i) vals[i] = (int32_t)i;
    hxo_sort_pairs(keys.data(), vals.data(), n, bits);
    const int64_t N = std::max<int64_t>(2 * n - 1, 1);
    std::vector<int32_t> left(N), right(N), parent(N);
    std::vector<float> nmin(4 * N), nmax(4 * N);
    if (n > 1) hxo_construct(keys.data(), n, left.data(), right.data(), parent.data());
    hxo_refit(vals.data(), R.mn.data(), R.mx.data(), n, left.data(), right.data(), parent.data(), nmin.data(),
              nmax.data());
    const int64_t acc = hxo_tr(n, iterations, ct, ci, left.data(), right.data(), parent.data(), nmin.data(),
                               nmax.data());
    if (accepted) *accepted = acc;
    int64_t nn = hxo_collapse(n, vals.data(), left.data(), right.data(), parent.data(), nmin.data(), nmax.data(),
                              adaptive, max_leaf_size, ct, ci, nodes, tri_out, n_interior, n_leaves);
    refs_to_tris(R, tri_out);
    hxo_woopify(verts, tris, tri_out, n, woop);
    return nn;
}

/* rt/ray/RayBufferKernels.cu:556-614 */
void hxo_ray_keys_aila64(const float* rays, int64_t n_rays, int bits, uint64_t* keys) {
    parallel_for(n_rays, [&](int64_t i) {
        const float* r = rays + 8 * i;
        /* ray box [0,1]^3: (o - mn) / (mx - mn) */
        V3 a{(r[0] - 0.0f) / (1.0f - 0.0f), (r[1] - 0.0f) / (1.0f - 0.0f), (r[2] - 0.0f) / (1.0f - 0.0f)};
        float ls = 0.0f;
        ls += r[4] * r[4];
        ls += r[5] * r[5];
        ls += r[6] * r[6];
        float len = std::sqrt(ls);
        float rl = (len != 0.0f) ? 1.0f / len : 0.0f;
        V3 b{(r[4] * rl + 1.0f) * 0.5f, (r[5] * rl + 1.0f) * 0.5f, (r[6] * rl + 1.0f) * 0.5f};
        int32_t iax = f2i(a.x * 8191.0f), iay = f2i(a.y * 8191.0f), iaz = f2i(a.z * 8191.0f);
        int32_t ibx = f2i(b.x * 8191.0f), iby = f2i(b.y * 8191.0f), ibz = f2i(b.z * 8191.0f);
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
        for (int k = 9; k >= 1; --k) {
            code |= (uint64_t)((iax >> k) & 1) << (6 * k + 0);
            code |= (uint64_t)((iay >> k) & 1) << (6 * k - 1);
            code |= (uint64_t)((iaz >> k) & 1) << (6 * k - 2);
        }
        for (int k = 12; k >= 4; --k) {
            code |= (uint64_t)((ibx >> k) & 1) << (6 * k - 21);
            code |= (uint64_t)((iby >> k) & 1) << (6 * k - 22);
            code |= (uint64_t)((ibz >> k) & 1) << (6 * k - 23);
        }
        keys[i] = code >> (64 - bits);
    });
}

/* rt/ray/RayBufferKernels.cu:12-56 (intersect) and :492-554 (TwoPoint64). */
void hxo_ray_keys_twopoint64(const float* rays, int64_t n_rays, int bits, float ray_length, uint64_t* keys) {
    auto fmin_n = [](float a, float b) { return (a != a) ? b : (b != b) ? a : ((b < a) ? b : a); };
    auto fmax_n = [](float a, float b) { return (a != a) ? b : (b != b) ? a : ((b > a) ? b : a); };
    parallel_for(n_rays, [&](int64_t i) {
        const float* r = rays + 8 * i;
        float o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
        float tmn[3], tmx[3];
        for (int k = 0; k < 3; ++k) {
            if (d[k] >= 0.0f) { tmn[k] = (0.0f - o[k]) / d[k]; tmx[k] = (1.0f - o[k]) / d[k]; }
            else { tmn[k] = (1.0f - o[k]) / d[k]; tmx[k] = (0.0f - o[k]) / d[k]; }
        }
        float t0 = fmax_n(tmn[2], fmax_n(tmn[0], tmn[1]));
        float t1 = fmin_n(tmx[2], fmin_n(tmx[0], tmx[1]));
        float boxT = (t0 > t1) ? -1.0f : (1.0f - 1.0e-3f) * (t0 >= 0.0f ? t0 : t1);
        float len = ray_length;
        len = fmin_n(len, fmax_n(r[7], 0.0f));
        len = fmin_n(len, boxT);
        float a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = o[k];
            b[k] = o[k] + len * d[k];
            a[k] = (a[k] - 0.0f) * (1.0f / (1.0f - 0.0f));
            b[k] = (b[k] - 0.0f) * (1.0f / (1.0f - 0.0f));
        }
        int32_t ia[3], ib[3];
        for (int k = 0; k < 3; ++k) { ia[k] = f2i(a[k] * 8191.0f); ib[k] = f2i(b[k] * 8191.0f); }
        uint64_t code = 0;
        for (int k = 12; k >= 2; --k) {
            code |= (uint64_t)((ia[0] >> k) & 1) << (6 * k - 9);
            code |= (uint64_t)((ia[1] >> k) & 1) << (6 * k - 10);
            code |= (uint64_t)((ia[2] >> k) & 1) << (6 * k - 11);
        }
        for (int k = 12; k >= 3; --k) {
            code |= (uint64_t)((ib[0] >> k) & 1) << (6 * k - 12);
            code |= (uint64_t)((ib[1] >> k) & 1) << (6 * k - 13);
            code |= (uint64_t)((ib[2] >> k) & 1) << (6 * k - 14);
        }
        code |= (uint64_t)(ib[0] & 1);
        keys[i] = code >> (64 - bits);
    });
}

}  // extern "C"
