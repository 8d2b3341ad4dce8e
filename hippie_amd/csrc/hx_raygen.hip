// hx_raygen.hip -- the steps either side of a trace on gfx950: ray generation
// (primary, shadow, ambient occlusion, path), occlusion reconstruction into
// pixels, and keyframe interpolation of a dynamic scene's vertices before a
// refit.
//
// Reference (src/hippie/): rt/ray/RayGenKernels.cu:17-535 (generators and
// their samplers), rt/ray/RayGen.cpp:35-200 (launch parameters),
// rt/renderer/RendererKernels.cu:302-372 (reconstructShadow/AO),
// rt/scene/InterpolatorKernels.cu:10-46, rt/HipUtil.h:190-247 (smooth normal,
// Jenkins mix, LCG, TEA).  Vector arithmetic follows util/Math.h's operator
// order (sums start from 0, normalize = rcp(length) * v) with
// -ffp-contract=off, so everything but the trigonometric and power functions
// rounds exactly as the reference's code.
//
// One thread per input item, 256-thread workgroups.  Path rays are compacted
// in input order (the reference appends them with a warp atomic, so its order
// changes from run to run): a count pass, a scan of the per-workgroup counts
// and a scatter pass that recomputes each path's sample.
#include "hx_device.h"
#include "hx_internal.h"

namespace hx {

namespace {

constexpr int kGenBlock = 256;
constexpr float kPi = 3.14159274101257324f;   // float(M_PI)
constexpr float kMaxFloat = 3.402823466e+38f;

struct V3 {
    float x, y, z;
};
HX_DEV V3 v3(float x, float y, float z) { return V3{x, y, z}; }
HX_DEV V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
HX_DEV V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
HX_DEV V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
HX_DEV V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
HX_DEV V3 operator*(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
HX_DEV V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }
HX_DEV float dot3(V3 a, V3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
HX_DEV float length3(V3 a) { return sqrtf(dot3(a, a)); }
HX_DEV V3 normalize3(V3 a) {
    const float l = length3(a);
    const float r = l != 0.0f ? 1.0f / l : 0.0f;
    return r * a;
}
HX_DEV V3 cross3(V3 u, V3 v) { return v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x); }
HX_DEV V3 ld3(const float* p, int64_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
HX_DEV V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

// --- samplers (RayGenKernels.cu:17-58, HipUtil.h:211-247) ---------------------

HX_DEV float halton(int i, int base) {
    float r = 0.0f;
    float f = 1.0f / (float)base;
    int j = i;
    while (j > 0) {
        r += f * (float)(j % base);
        j = (int)floorf((float)j / (float)base);
        f /= (float)base;
    }
    return r;
}

HX_DEV float hammersley(int i, int num) { return ((float)i + 0.5f) / (float)num; }

HX_DEV void sobol2d(int i, float& x, float& y) {
    uint32_t r1 = 0, r2 = 0;
    for (uint32_t v1 = 1u << 31, v2 = 3u << 30; i; i >>= 1) {
        if (i & 1) {
            r1 ^= v1;
            r2 ^= v2 << 1;
        }
        v1 |= v1 >> 1;
        v2 ^= v2 >> 1;
    }
    x = (float)r1 * (1.0f / 4294967296.0f);
    y = (float)r2 * (1.0f / 4294967296.0f);
}

HX_DEV void jenkins_mix(uint32_t& a, uint32_t& b, uint32_t& c) {
    a -= b; a -= c; a ^= (c >> 13);
    b -= c; b -= a; b ^= (a << 8);
    c -= a; c -= b; c ^= (b >> 13);
    a -= b; a -= c; a ^= (c >> 12);
    b -= c; b -= a; b ^= (a << 16);
    c -= a; c -= b; c ^= (b >> 5);
    a -= b; a -= c; a ^= (c >> 3);
    b -= c; b -= a; b ^= (a << 10);
    c -= a; c -= b; c ^= (b >> 15);
}

HX_DEV float randf(uint32_t& seed) {
    seed = 1103515245u * seed + 12345u;
    return (float)(seed & 0x00FFFFFFu) / (float)0x01000000;
}

HX_DEV uint32_t tea16(uint32_t v0, uint32_t v1) {
    uint32_t s0 = 0;
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

// (hashA, hashB, hashC) of a task: seed + task index, mixed twice.
HX_DEV void task_hash(uint32_t seed, int64_t task, uint32_t& a, uint32_t& b, uint32_t& c) {
    a = seed + (uint32_t)task;
    b = 0x9e3779b9u;
    c = 0x9e3779b9u;
    jenkins_mix(a, b, c);
    jenkins_mix(a, b, c);
}

HX_DEV void write_ray(float4* rays, int64_t i, V3 o, float tmin, V3 d, float tmax) {
    rays[2 * i] = make_float4(o.x, o.y, o.z, tmin);
    rays[2 * i + 1] = make_float4(d.x, d.y, d.z, tmax);
}

// --- generators -----------------------------------------------------------------

__global__ __launch_bounds__(kGenBlock) void k_init_seeds(int64_t n, uint32_t frame, uint32_t* __restrict__ seeds) {
    const int64_t i = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (i < n) seeds[i] = tea16((uint32_t)i, frame);
}

struct PrimaryArgs {
    float m[16];   // screen -> world, row-major: m[4 * i + j] = M(i, j)
    float ox, oy, oz, max_dist;
    int sample_index, width, height;
    const int32_t* index_to_pixel;   // NULL: identity
    float4* rays;
};

__global__ __launch_bounds__(kGenBlock) void k_primary_rays(PrimaryArgs a) {
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (task >= (int64_t)a.width * a.height) return;
    const int pixel = a.index_to_pixel ? a.index_to_pixel[task] : (int)task;
    const int sx = pixel % a.width, sy = pixel / a.width;
    const float hx_ = halton(a.sample_index, 2), hy_ = halton(a.sample_index, 3);
    const float x = (float)sx + hx_;
    const float y = (float)sy + hy_;
    const float p[4] = {2.0f * x / (float)a.width - 1.0f, 2.0f * y / (float)a.height - 1.0f, 0.0f, 1.0f};
    float w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float rr = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) rr += a.m[4 * i + k] * p[k];
        w[i] = rr;
    }
    const V3 world = v3(w[0] / w[3], w[1] / w[3], w[2] / w[3]);
    const V3 o = v3(a.ox, a.oy, a.oz);
    write_ray(a.rays, task, o, 0.0f, normalize3(world - o), a.max_dist);
}

struct SecondaryArgs {
    int64_t batch_begin, batch_size;
    int n_samples;
    const uint32_t* seeds;
    const float4* in_rays;
    const int4* in_results;
    float4* out_rays;
    // shadow
    float light_radius, lx, ly, lz;
    // ambient occlusion
    float max_dist;
    const int32_t* tris;
    const float* verts;
};

__global__ __launch_bounds__(kGenBlock) void k_shadow_rays(SecondaryArgs a) {
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (task >= a.batch_size) return;
    uint32_t ha, hb, hc;
    task_hash(a.seeds[task], task, ha, hb, hc);
    const float4 ro = a.in_rays[2 * (a.batch_begin + task)], rd = a.in_rays[2 * (a.batch_begin + task) + 1];
    const int4 res = a.in_results[a.batch_begin + task];
    const V3 offset = v3((float)ha * 2.3283064365386963e-10f, (float)hb * 2.3283064365386963e-10f,
                         (float)hc * 2.3283064365386963e-10f);
    const int tri = res.x;
    V3 origin = xyz(ro) + xyz(rd) * fmaxf(__int_as_float(res.y) - 1.0e-5f, 0.0f);
    if (tri == -1) origin = v3(0.0f, 0.0f, 0.0f);
    const V3 light = v3(a.lx, a.ly, a.lz);
    int64_t out = (int64_t)a.n_samples * task;
    for (int i = 0; i < a.n_samples; ++i, ++out) {
        float sx, sy;
        sobol2d(i, sx, sy);
        V3 pos = v3(sx, sy, hammersley(i, a.n_samples)) + offset;
        if (pos.x >= 1.0f) pos.x -= 1.0f;
        if (pos.y >= 1.0f) pos.y -= 1.0f;
        if (pos.z >= 1.0f) pos.z -= 1.0f;
        pos = 2.0f * pos - v3(1.0f, 1.0f, 1.0f);
        const V3 target = light + a.light_radius * pos;
        const V3 dir = target - origin;
        write_ray(a.out_rays, out, origin, 0.0f, tri == -1 ? v3(1.0f, 0.0f, 0.0f) : normalize3(dir),
                  tri == -1 ? -1.0f : length3(dir));
    }
}

__global__ __launch_bounds__(kGenBlock) void k_ao_rays(SecondaryArgs a) {
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (task >= a.batch_size) return;
    const float4 ro = a.in_rays[2 * (a.batch_begin + task)], rd = a.in_rays[2 * (a.batch_begin + task) + 1];
    const int4 res = a.in_results[a.batch_begin + task];
    const V3 dir_in = xyz(rd);
    const V3 origin = xyz(ro) + dir_in * fmaxf(__int_as_float(res.y) - 1.0e-4f, 0.0f);
    V3 normal = v3(1.0f, 0.0f, 0.0f);
    if (res.x != -1) {
        const int32_t* t = a.tris + 3 * (int64_t)res.x;
        const V3 v0 = ld3(a.verts, t[0]), v1 = ld3(a.verts, t[1]), v2 = ld3(a.verts, t[2]);
        normal = normalize3(cross3(v1 - v0, v2 - v0));
    }
    if (dot3(normal, dir_in) > 0.0f) normal = neg(normal);
    const V3 na = v3(fabsf(normal.x), fabsf(normal.y), fabsf(normal.z));
    const float nm = fmaxf(fmaxf(na.x, na.y), na.z);
    V3 perp = v3(normal.y, -normal.x, 0.0f);
    if (nm == na.z) perp = v3(0.0f, normal.z, -normal.y);
    else if (nm == na.x) perp = v3(-normal.z, 0.0f, normal.x);
    perp = normalize3(perp);
    const V3 biperp = cross3(normal, perp);
    uint32_t ha, hb, hc;
    task_hash(a.seeds[task], task, ha, hb, hc);
    const float angle = 2.0f * kPi * (float)hc * 2.3283064365386963e-10f;
    const float ca = cosf(angle), sa = sinf(angle);
    const V3 t0 = perp * ca + biperp * sa;
    const V3 t1 = perp * -sa + biperp * ca;
    int64_t out = (int64_t)a.n_samples * task;
    for (int i = 0; i < a.n_samples; ++i, ++out) {
        float x = 0.0f, xadd = 1.0f;
        for (uint32_t hc2 = (uint32_t)i + 1; hc2 != 0; hc2 >>= 1) {
            xadd *= 0.5f;
            if (hc2 & 1) x += xadd;
        }
        float y = 0.0f, yadd = 1.0f;
        for (int hc3 = i + 1; hc3 != 0; hc3 /= 3) {
            yadd *= 1.0f / 3.0f;
            y += (float)(hc3 % 3) * yadd;
        }
        const float ang = 2.0f * kPi * y;
        const float r = sqrtf(x);
        x = r * cosf(ang);
        y = r * sinf(ang);
        const float z = sqrtf(1.0f - x * x - y * y);
        write_ray(a.out_rays, out, origin, 0.0f, normalize3(x * t0 + y * t1 + z * normal),
                  res.x == -1 ? -1.0f : a.max_dist);
    }
}

// --- path rays (RayGenKernels.cu:354-519) -------------------------------------------

struct Material {
    int tex_index;
    float shininess;
    float dr, dg, db;   // diffuse
    float sr, sg, sb;   // specular
};

struct PathArgs {
    int64_t n;
    int russian_roulette;
    uint32_t* seeds;
    const int32_t* in_pixel;
    const int32_t* mat_indices;
    const int32_t* tris;
    const float* normals;
    const Material* materials;
    const float4* in_rays;
    const int4* in_results;
    float4* out_rays;
    int32_t* out_pixel;
    float* decreases;         // float3 per pixel
    int32_t* block_counts;    // surviving paths per workgroup, then their exclusive prefix
    int32_t* n_out;
};

HX_DEV V3 arbitrary_normal(V3 n) {
    const float d2 = n.x * n.x + n.y * n.y;
    if (d2 > 0.0001f) {
        const float inv = 1.0f / sqrtf(d2);
        return v3(n.y * inv, -n.x * inv, 0.0f);
    }
    const float inv = 1.0f / sqrtf(n.z * n.z + n.x * n.x);
    return v3(-n.z * inv, 0.0f, n.x * inv);
}

HX_DEV void right_handed_base(V3& u, V3& v, V3 w) {
    v = arbitrary_normal(w);
    u = cross3(v, w);
}

// The sampled continuation of one path, or false when it terminates; the
// seed advances as the reference's does (three draws per live path).
HX_DEV bool path_sample(const PathArgs& a, int64_t task, uint32_t& seed, V3& o_out, V3& d_out, V3& weight) {
    const float4 ro = a.in_rays[2 * task], rd = a.in_rays[2 * task + 1];
    const int4 res = a.in_results[task];
    if (res.x < 0) return false;
    const int32_t* t = a.tris + 3 * (int64_t)res.x;
    const float u = __int_as_float(res.z), v = __int_as_float(res.w);
    const V3 n0 = ld3(a.normals, t[0]), n1 = ld3(a.normals, t[1]), n2 = ld3(a.normals, t[2]);
    V3 normal = normalize3(u * n0 + v * n1 + (1.0f - u - v) * n2);
    const V3 din = xyz(rd);
    float dp = dot3(normal, din);
    if (dp > 0.0f) {
        normal = neg(normal);
        dp = -dp;
    }
    const Material mat = a.materials[a.mat_indices[res.x]];
    V3 kd = v3(mat.dr, mat.dg, mat.db);   // textures are not part of this library
    V3 ks = v3(mat.sr, mat.sg, mat.sb);
    float pd = fmaxf(fmaxf(kd.x, kd.y), kd.z);
    float ps = fmaxf(fmaxf(ks.x, ks.y), ks.z);
    const float h = mat.shininess;
    if (pd + ps > 1.0f) {
        const float s = 1.0f / (pd + ps);
        kd = kd * s;
        ks = ks * s;
        pd *= s;
        ps *= s;
    }
    const float r0 = randf(seed), r1 = randf(seed), r2 = randf(seed);
    if (!a.russian_roulette) {
        const float s = 1.0f / (pd + ps);
        pd *= s;
        ps *= s;
    }
    if (!(r2 <= pd + ps)) return false;
    const V3 refl = normalize3(din - (2.0f * dp) * normal);
    const V3 hitp = xyz(ro) + din * __int_as_float(res.y);
    const float eps = 1.0e-4f;
    float pdf = 0.0f, cos_theta = 0.0f, cos_alpha = 0.0f;
    V3 dir = v3(0.0f, 0.0f, 0.0f);
    if (r2 >= pd) {
        V3 bu, bv;
        right_handed_base(bu, bv, refl);
        const float sin_t = sqrtf(1.0f - powf(r1, 2.0f / (float)(h + 1.0f)));
        const float cos_t = powf(r1, 1.0f / (float)(h + 1.0f));
        const float phi = 2.0f * kPi * r0;
        pdf = (h + 1.0f) / (2.0f * kPi) * powf(cos_t, h);
        const V3 hem = v3(cosf(phi) * sin_t, sinf(phi) * sin_t, cos_t);
        dir = normalize3(hem.x * bu + hem.y * bv + hem.z * refl);
        cos_theta = dot3(normal, dir);
        cos_alpha = dot3(refl, dir);
        const V3 brdf = ((h + 2.0f) / (2.0f * kPi)) * ks * powf(cos_alpha, h);
        weight = v3(brdf.x * cos_theta / pdf / (pd + ps), brdf.y * cos_theta / pdf / (pd + ps),
                    brdf.z * cos_theta / pdf / (pd + ps));
    }
    if (r2 < pd || (r2 >= pd && (cos_theta < 0.0f || pdf < 0.01f || cos_alpha < 0.0f))) {
        const float theta = 2.0f * kPi * r1;
        const float radius = sqrtf(r0);
        const float x = radius * sinf(theta);
        const float z = radius * cosf(theta);
        const float y = sqrtf(1.0f - x * x - z * z);
        V3 bu, bv;
        right_handed_base(bu, bv, normal);
        dir = x * bu + z * bv + y * normal;
        cos_theta = dot3(normal, dir);
        pdf = cos_theta / kPi;
        const V3 brdf = v3(kd.x / kPi, kd.y / kPi, kd.z / kPi);
        weight = v3(brdf.x * cos_theta / pdf / (pd + ps), brdf.y * cos_theta / pdf / (pd + ps),
                    brdf.z * cos_theta / pdf / (pd + ps));
    }
    o_out = hitp + dir * eps;
    d_out = dir;
    return true;
}

__global__ __launch_bounds__(kGenBlock) void k_path_count(PathArgs a) {
    __shared__ int wave_n[kGenBlock / kWave];
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    bool live = false;
    if (task < a.n) {
        uint32_t seed = a.seeds[task];
        V3 o, d, w;
        live = path_sample(a, task, seed, o, d, w);
    }
    const int cnt = __popcll(ballot64(live));
    if (lane_id() == 0) wave_n[threadIdx.x / kWave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int k = 0; k < kGenBlock / kWave; ++k) s += wave_n[k];
        a.block_counts[blockIdx.x] = s;
    }
}

// Exclusive prefix of the per-workgroup counts in place (one workgroup), and
// the total.
__global__ __launch_bounds__(1024) void k_path_scan(int32_t* __restrict__ counts, int64_t nb, int32_t* __restrict__ n_out) {
    __shared__ int wt[1024 / kWave];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < nb; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const int v = i < nb ? counts[i] : 0;
        int total;
        const int inc = block_inclusive_scan<1024 / kWave>(v, [](int x, int y) { return x + y; }, wt, total);
        const int c0 = carry;
        if (i < nb) counts[i] = c0 + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) carry = c0 + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *n_out = carry;
}

__global__ __launch_bounds__(kGenBlock) void k_path_scatter(PathArgs a) {
    __shared__ int wave_n[kGenBlock / kWave];
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    bool live = false;
    uint32_t seed = 0;
    V3 o = v3(0.0f, 0.0f, 0.0f), d = o, w = o;
    int4 res = make_int4(-1, 0, 0, 0);
    if (task < a.n) {
        seed = a.seeds[task];
        res = a.in_results[task];
        live = path_sample(a, task, seed, o, d, w);
        if (res.x >= 0) a.seeds[task] = seed;
    }
    const uint64_t m = ballot64(live);
    const int wave = threadIdx.x / kWave;
    if (lane_id() == 0) wave_n[wave] = __popcll(m);
    __syncthreads();
    int off = a.block_counts[blockIdx.x];
    for (int k = 0; k < wave; ++k) off += wave_n[k];
    if (live) {
        const int64_t slot = off + count_below(m);
        const int pixel = a.in_pixel[task];
        write_ray(a.out_rays, slot, o, 0.0f, d, kMaxFloat);
        a.out_pixel[slot] = pixel;
        float* dec = a.decreases + 3 * (int64_t)pixel;
        dec[0] *= w.x;
        dec[1] *= w.y;
        dec[2] *= w.z;
    }
}

// --- reconstruction and interpolation ------------------------------------------------

struct OcclusionArgs {
    int64_t batch_begin, batch_size;
    int n_samples, replace;
    const int4* results;
    const int32_t* index_to_pixel;
    const int32_t* index_to_slot;   // NULL: identity
    const float4* in_pixels;
    float4* out_pixels;
};

// reconstructAO / reconstructShadow (RendererKernels.cu:302-372): the pixel's
// colour scaled by the fraction of its samples that miss.
__global__ __launch_bounds__(kGenBlock) void k_reconstruct_occlusion(OcclusionArgs a) {
    const int64_t task = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (task >= a.batch_size) return;
    const int pixel = a.index_to_pixel[a.batch_begin + task];
    const float4 c = a.in_pixels[pixel];
    float f = 0.0f;
    for (int i = 0; i < a.n_samples; ++i) {
        const int64_t k = task * a.n_samples + i;
        const int64_t slot = a.index_to_slot ? a.index_to_slot[k] : k;
        if (a.results[slot].x == -1) f += 1.0f;
    }
    f /= (float)a.n_samples;
    const float4 v = make_float4(c.x * f, c.y * f, c.z * f, 1.0f);
    if (a.replace) {
        a.out_pixels[pixel] = v;
    } else {
        float4 o = a.out_pixels[pixel];
        a.out_pixels[pixel] = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
    }
}

// interpolateVertices / interpolateNormals (InterpolatorKernels.cu:10-46):
// out[n_static + i] = mix(a[i], b[i], t) = a (1 - t) + b t, normalised for normals.
__global__ __launch_bounds__(kGenBlock) void k_interpolate(int64_t n_static, int64_t n, float t, const float* __restrict__ a,
                                                           const float* __restrict__ b, float* __restrict__ out,
                                                           int normalize) {
    const int64_t i = blockIdx.x * (int64_t)kGenBlock + threadIdx.x;
    if (i >= n) return;
    const float s = 1.0f - t;
    V3 r = ld3(a, i) * s + ld3(b, i) * t;
    if (normalize) r = normalize3(r);
    float* o = out + 3 * (n_static + i);
    o[0] = r.x;
    o[1] = r.y;
    o[2] = r.z;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kGenBlock - 1) / kGenBlock); }

}  // namespace

int launch_init_seeds(Context* c, int64_t n, uint32_t frame, uint32_t* seeds) {
    if (n <= 0) return HX_OK;
    hipLaunchKernelGGL(k_init_seeds, dim3(grid_of(n)), dim3(kGenBlock), 0, c->stream, n, frame, seeds);
    return check_launch("init_seeds");
}

int launch_primary_rays(Context* c, int sample_index, const float origin[3], const float m[16], int width, int height,
                        float max_dist, const int32_t* index_to_pixel, float* rays) {
    PrimaryArgs a{};
    for (int k = 0; k < 16; ++k) a.m[k] = m[k];
    a.ox = origin[0];
    a.oy = origin[1];
    a.oz = origin[2];
    a.max_dist = max_dist;
    a.sample_index = sample_index;
    a.width = width;
    a.height = height;
    a.index_to_pixel = index_to_pixel;
    a.rays = reinterpret_cast<float4*>(rays);
    LaunchScope ls(c, "primary_rays");
    hipLaunchKernelGGL(k_primary_rays, dim3(grid_of((int64_t)width * height)), dim3(kGenBlock), 0, c->stream, a);
    return check_launch("primary_rays");
}

int launch_secondary_rays(Context* c, bool ao, int64_t begin, int64_t size, int n_samples, const uint32_t* seeds,
                          float light_radius, const float light[3], float max_dist, const float* in_rays,
                          const int32_t* in_results, const int32_t* tris, const float* verts, float* out_rays) {
    if (size <= 0) return HX_OK;
    SecondaryArgs a{};
    a.batch_begin = begin;
    a.batch_size = size;
    a.n_samples = n_samples;
    a.seeds = seeds;
    a.in_rays = reinterpret_cast<const float4*>(in_rays);
    a.in_results = reinterpret_cast<const int4*>(in_results);
    a.out_rays = reinterpret_cast<float4*>(out_rays);
    a.light_radius = light_radius;
    if (light) {
        a.lx = light[0];
        a.ly = light[1];
        a.lz = light[2];
    }
    a.max_dist = max_dist;
    a.tris = tris;
    a.verts = verts;
    LaunchScope ls(c, ao ? "ao_rays" : "shadow_rays");
    if (ao)
        hipLaunchKernelGGL(k_ao_rays, dim3(grid_of(size)), dim3(kGenBlock), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_shadow_rays, dim3(grid_of(size)), dim3(kGenBlock), 0, c->stream, a);
    return check_launch(ao ? "ao_rays" : "shadow_rays");
}

int launch_path_rays(Context* c, const PathRayArgs& p) {
    const int64_t nb = (p.n + kGenBlock - 1) / kGenBlock;
    HX_TRY(c->scan_partials.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(nb, 1)));
    PathArgs a{};
    a.n = p.n;
    a.russian_roulette = p.russian_roulette;
    a.seeds = p.seeds;
    a.in_pixel = p.in_pixel;
    a.mat_indices = p.mat_indices;
    a.tris = p.tris;
    a.normals = p.normals;
    a.materials = static_cast<const Material*>(p.materials);
    a.in_rays = reinterpret_cast<const float4*>(p.in_rays);
    a.in_results = reinterpret_cast<const int4*>(p.in_results);
    a.out_rays = reinterpret_cast<float4*>(p.out_rays);
    a.out_pixel = p.out_pixel;
    a.decreases = p.decreases;
    a.block_counts = c->scan_partials.as<int32_t>();
    a.n_out = p.n_out;
    LaunchScope ls(c, "path_rays");
    if (nb > 0) {
        hipLaunchKernelGGL(k_path_count, dim3((unsigned)nb), dim3(kGenBlock), 0, c->stream, a);
        HX_TRY(check_launch("path_count"));
    }
    hipLaunchKernelGGL(k_path_scan, dim3(1), dim3(1024), 0, c->stream, a.block_counts, nb, a.n_out);
    HX_TRY(check_launch("path_scan"));
    if (nb > 0) {
        hipLaunchKernelGGL(k_path_scatter, dim3((unsigned)nb), dim3(kGenBlock), 0, c->stream, a);
        HX_TRY(check_launch("path_scatter"));
    }
    return HX_OK;
}

int launch_reconstruct_occlusion(Context* c, int64_t begin, int64_t size, int n_samples, bool replace,
                                 const int32_t* results, const int32_t* index_to_pixel, const int32_t* index_to_slot,
                                 const float* in_pixels, float* out_pixels) {
    if (size <= 0) return HX_OK;
    OcclusionArgs a{};
    a.batch_begin = begin;
    a.batch_size = size;
    a.n_samples = n_samples;
    a.replace = replace ? 1 : 0;
    a.results = reinterpret_cast<const int4*>(results);
    a.index_to_pixel = index_to_pixel;
    a.index_to_slot = index_to_slot;
    a.in_pixels = reinterpret_cast<const float4*>(in_pixels);
    a.out_pixels = reinterpret_cast<float4*>(out_pixels);
    LaunchScope ls(c, "reconstruct_occlusion");
    hipLaunchKernelGGL(k_reconstruct_occlusion, dim3(grid_of(size)), dim3(kGenBlock), 0, c->stream, a);
    return check_launch("reconstruct_occlusion");
}

int launch_interpolate(Context* c, int64_t n_static, int64_t n, float t, const float* a, const float* b, float* out,
                       bool normalize) {
    if (n <= 0) return HX_OK;
    LaunchScope ls(c, "interpolate");
    hipLaunchKernelGGL(k_interpolate, dim3(grid_of(n)), dim3(kGenBlock), 0, c->stream, n_static, n, t, a, b, out,
                       normalize ? 1 : 0);
    return check_launch("interpolate");
}

}  // namespace hx
