/*
 * hx_oracle_raygen.cpp -- CPU restatement of hippie's ray generation, occlusion
 * reconstruction and vertex interpolation (TEST INFRASTRUCTURE).
 *
 * Loaded only by tests/ (part of liboracle.so); the product never links it.
 * Built with -ffp-contract=off.  Reference citations are relative to
 * src/hippie/.  Vector arithmetic follows util/Math.h: sums start from 0
 * (dot, lenSqr, matrix rows), normalize(v) = rcp(length(v)) * v, mix(a, b, t)
 * = a (1 - t) + b t.  The trigonometric and power functions are the host
 * libm's cosf/sinf/powf, so AO and path directions agree with a GPU only to
 * within a few float ulps; everything else is bit-exact.
 *
 * How each function is pinned (tests/test_oracle_raygen.py):
 *   tea16         -- TEA decryption (the published inverse round) undoes it.
 *   jenkins mix   -- its step-by-step inverse recovers the input.
 *   halton        -- exact radical inverse (fractions) within float rounding.
 *   sobol2d       -- the first 2^m points form a (0,m,2)-net in base 2.
 *   primary rays  -- float64 pinhole geometry of the camera (direction through
 *                    each pixel corner) within 1e-5.
 *   shadow rays   -- end on the light's square, lengths, dead rays.
 *   AO rays       -- unit length, in the hemisphere of the flipped normal,
 *                    Halton points warped by the closed form, dead rays.
 *   path rays     -- survivors by the closed-form probability, unit
 *                    directions in the normal's hemisphere, weights of the
 *                    cosine lobe kd / (pd + ps) per channel.
 *   occlusion     -- miss fraction counted directly in numpy.
 *   interpolation -- numpy float32 with the same operation order.
 */
#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

constexpr float kPi = 3.14159274101257324f;   // float(M_PI), Globals.h:65
constexpr float kMaxFloat = 3.402823466e+38f;  // Globals.h:70

struct V3 {
    float x, y, z;
};
inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 scale(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline V3 scale_r(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 mulv(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float dot(V3 a, V3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
inline float len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalize(V3 a) {
    const float l = len(a);
    return scale(l ? 1.0f / l : 0.0f, a);
}
inline V3 cross(V3 u, V3 v) { return v3(u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x); }
inline V3 ld(const float* p, int64_t i) { return v3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }

/* RayGenKernels.cu:21-31 */
float halton(int i, int base) {
    float r = 0.0f;
    float f = 1.0f / (float)base;
    int j = i;
    while (j > 0) {
        r += f * (float)(j % base);
        j = (int)std::floor((float)j / (float)base);
        f /= (float)base;
    }
    return r;
}

/* RayGenKernels.cu:17-19 */
float hammersley(int i, int num) { return ((float)i + 0.5f) / (float)num; }

/* RayGenKernels.cu:41-59 */
void sobol2d(int i, float& x, float& y) {
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

/* rt/HipUtil.h:211-221 */
void jenkins(uint32_t& a, uint32_t& b, uint32_t& c) {
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

/* rt/HipUtil.h:223-233 */
float randf(uint32_t& seed) {
    seed = 1103515245u * seed + 12345u;
    return (float)(seed & 0x00FFFFFFu) / (float)0x01000000;
}

/* rt/HipUtil.h:235-247: 16 TEA rounds; the reference keeps v0. */
void tea16_pair(uint32_t& v0, uint32_t& v1) {
    uint32_t s0 = 0;
    for (int n = 0; n < 16; ++n) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
}
uint32_t tea16(uint32_t v0, uint32_t v1) {
    tea16_pair(v0, v1);
    return v0;
}

void put_ray(float* rays, int64_t i, V3 o, float tmin, V3 d, float tmax) {
    float* r = rays + 8 * i;
    r[0] = o.x; r[1] = o.y; r[2] = o.z; r[3] = tmin;
    r[4] = d.x; r[5] = d.y; r[6] = d.z; r[7] = tmax;
}

float as_float(int32_t i) {
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

/* RayGenKernels.cu:62-104 */
V3 arbitrary_normal(V3 n) {
    const float d2 = n.x * n.x + n.y * n.y;
    if (d2 > 0.0001f) {
        const float inv = 1.0f / std::sqrt(d2);
        return v3(n.y * inv, -n.x * inv, 0.0f);
    }
    const float inv = 1.0f / std::sqrt(n.z * n.z + n.x * n.x);
    return v3(-n.z * inv, 0.0f, n.x * inv);
}

struct Material {
    int tex_index;
    float shininess;
    float diffuse[3];
    float specular[3];
};

}  // namespace

extern "C" {

/* out: (v0, v1) after the rounds per input pair. */
void hxo_tea16(const uint32_t* v0, const uint32_t* v1, int64_t n, uint32_t* out) {
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a = v0[i], b = v1[i];
        tea16_pair(a, b);
        out[2 * i] = a;
        out[2 * i + 1] = b;
    }
}

void hxo_jenkins_mix(uint32_t* abc, int64_t n) {
    for (int64_t i = 0; i < n; ++i) jenkins(abc[3 * i], abc[3 * i + 1], abc[3 * i + 2]);
}

void hxo_halton(const int32_t* idx, int64_t n, int32_t base, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = halton(idx[i], base);
}

void hxo_sobol2d(const int32_t* idx, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) sobol2d(idx[i], out[2 * i], out[2 * i + 1]);
}

/* RayGenKernels.cu:520-535 (initSeeds) */
void hxo_init_seeds(int64_t n, int32_t frame, uint32_t* seeds) {
    for (int64_t i = 0; i < n; ++i) seeds[i] = tea16((uint32_t)i, (uint32_t)frame);
}

/* RayGenKernels.cu:106-153 (generatePrimaryRays); m row-major, index_to_pixel NULL = identity */
void hxo_primary_rays(int32_t sample_index, const float* origin, const float* m, int32_t w, int32_t h,
                      float max_dist, const int32_t* index_to_pixel, float* rays) {
    const int64_t n = (int64_t)w * h;
#pragma omp parallel for schedule(static)
    for (int64_t task = 0; task < n; ++task) {
        const int pixel = index_to_pixel ? index_to_pixel[task] : (int)task;
        const int sx = pixel % w, sy = pixel / w;
        const float x = (float)sx + halton(sample_index, 2);
        const float y = (float)sy + halton(sample_index, 3);
        const float p[4] = {2.0f * x / (float)w - 1.0f, 2.0f * y / (float)h - 1.0f, 0.0f, 1.0f};
        float wv[4];
        for (int i = 0; i < 4; ++i) {
            float rr = 0.0f;
            for (int k = 0; k < 4; ++k) rr += m[4 * i + k] * p[k];
            wv[i] = rr;
        }
        const V3 world = v3(wv[0] / wv[3], wv[1] / wv[3], wv[2] / wv[3]);
        const V3 o = v3(origin[0], origin[1], origin[2]);
        put_ray(rays, task, o, 0.0f, normalize(sub(world, o)), max_dist);
    }
}

/* RayGenKernels.cu:155-229 (generateShadowRays) */
void hxo_shadow_rays(int64_t begin, int64_t size, int32_t ns, const uint32_t* seeds, float light_radius,
                     const float* light, const float* in_rays, const int32_t* in_results, float* out_rays) {
#pragma omp parallel for schedule(static)
    for (int64_t task = 0; task < size; ++task) {
        uint32_t a = seeds[task] + (uint32_t)task, b = 0x9e3779b9u, c = 0x9e3779b9u;
        jenkins(a, b, c);
        jenkins(a, b, c);
        const float* r = in_rays + 8 * (begin + task);
        const int32_t* res = in_results + 4 * (begin + task);
        const V3 offset = v3((float)a * 2.3283064365386963e-10f, (float)b * 2.3283064365386963e-10f,
                             (float)c * 2.3283064365386963e-10f);
        const int tri = res[0];
        V3 origin = add(v3(r[0], r[1], r[2]), scale_r(v3(r[4], r[5], r[6]), std::fmax(as_float(res[1]) - 1.0e-5f, 0.0f)));
        if (tri == -1) origin = v3(0.0f, 0.0f, 0.0f);
        const V3 L = v3(light[0], light[1], light[2]);
        for (int i = 0; i < ns; ++i) {
            float sx, sy;
            sobol2d(i, sx, sy);
            V3 pos = add(v3(sx, sy, hammersley(i, ns)), offset);
            if (pos.x >= 1.0f) pos.x -= 1.0f;
            if (pos.y >= 1.0f) pos.y -= 1.0f;
            if (pos.z >= 1.0f) pos.z -= 1.0f;
            pos = sub(scale(2.0f, pos), v3(1.0f, 1.0f, 1.0f));
            const V3 dir = sub(add(L, scale(light_radius, pos)), origin);
            put_ray(out_rays, (int64_t)ns * task + i, origin, 0.0f, tri == -1 ? v3(1.0f, 0.0f, 0.0f) : normalize(dir),
                    tri == -1 ? -1.0f : len(dir));
        }
    }
}

/* RayGenKernels.cu:232-349 (generateAORays) */
void hxo_ao_rays(int64_t begin, int64_t size, int32_t ns, const uint32_t* seeds, float max_dist, const float* in_rays,
                 const int32_t* in_results, const int32_t* tris, const float* verts, float* out_rays) {
#pragma omp parallel for schedule(static)
    for (int64_t task = 0; task < size; ++task) {
        const float* r = in_rays + 8 * (begin + task);
        const int32_t* res = in_results + 4 * (begin + task);
        const V3 din = v3(r[4], r[5], r[6]);
        const V3 origin = add(v3(r[0], r[1], r[2]), scale_r(din, std::fmax(as_float(res[1]) - 1.0e-4f, 0.0f)));
        V3 nrm = v3(1.0f, 0.0f, 0.0f);
        if (res[0] != -1) {
            const int32_t* t = tris + 3 * (int64_t)res[0];
            const V3 v0 = ld(verts, t[0]), v1 = ld(verts, t[1]), v2 = ld(verts, t[2]);
            nrm = normalize(cross(sub(v1, v0), sub(v2, v0)));
        }
        if (dot(nrm, din) > 0.0f) nrm = v3(-nrm.x, -nrm.y, -nrm.z);
        const V3 na = v3(std::fabs(nrm.x), std::fabs(nrm.y), std::fabs(nrm.z));
        const float nm = std::fmax(std::fmax(na.x, na.y), na.z);
        V3 perp = v3(nrm.y, -nrm.x, 0.0f);
        if (nm == na.z) perp = v3(0.0f, nrm.z, -nrm.y);
        else if (nm == na.x) perp = v3(-nrm.z, 0.0f, nrm.x);
        perp = normalize(perp);
        const V3 biperp = cross(nrm, perp);
        uint32_t a = seeds[task] + (uint32_t)task, b = 0x9e3779b9u, c = 0x9e3779b9u;
        jenkins(a, b, c);
        jenkins(a, b, c);
        const float angle = 2.0f * kPi * (float)c * 2.3283064365386963e-10f;
        const float ca = std::cos(angle), sa = std::sin(angle);
        const V3 t0 = add(scale_r(perp, ca), scale_r(biperp, sa));
        const V3 t1 = add(scale_r(perp, -sa), scale_r(biperp, ca));
        for (int i = 0; i < ns; ++i) {
            float x = 0.0f, xadd = 1.0f;
            for (uint32_t h2 = (uint32_t)i + 1; h2 != 0; h2 >>= 1) {
                xadd *= 0.5f;
                if (h2 & 1) x += xadd;
            }
            float y = 0.0f, yadd = 1.0f;
            for (int h3 = i + 1; h3 != 0; h3 /= 3) {
                yadd *= 1.0f / 3.0f;
                y += (float)(h3 % 3) * yadd;
            }
            const float ang = 2.0f * kPi * y;
            const float rr = std::sqrt(x);
            x = rr * std::cos(ang);
            y = rr * std::sin(ang);
            const float z = std::sqrt(1.0f - x * x - y * y);
            put_ray(out_rays, (int64_t)ns * task + i, origin, 0.0f,
                    normalize(add(add(scale(x, t0), scale(y, t1)), scale(z, nrm))), res[0] == -1 ? -1.0f : max_dist);
        }
    }
}

/* RayGenKernels.cu:354-519 (generatePathRays), with the outputs in input order.
 * Returns the number of output rays. */
int64_t hxo_path_rays(int32_t russian_roulette, int64_t n, uint32_t* seeds, const int32_t* in_pixel,
                      const int32_t* mat_indices, const int32_t* tris, const float* normals, const void* materials_v,
                      const float* in_rays, const int32_t* in_results, float* out_rays, int32_t* out_pixel,
                      float* decreases) {
    const Material* materials = static_cast<const Material*>(materials_v);
    int64_t out = 0;
    for (int64_t task = 0; task < n; ++task) {
        const float* r = in_rays + 8 * task;
        const int32_t* res = in_results + 4 * task;
        if (res[0] < 0) continue;
        uint32_t seed = seeds[task];
        const int32_t* t = tris + 3 * (int64_t)res[0];
        const float u = as_float(res[2]), v = as_float(res[3]);
        const V3 n0 = ld(normals, t[0]), n1 = ld(normals, t[1]), n2 = ld(normals, t[2]);
        V3 nrm = normalize(add(add(scale(u, n0), scale(v, n1)), scale(1.0f - u - v, n2)));
        const V3 din = v3(r[4], r[5], r[6]);
        float dp = dot(nrm, din);
        if (dp > 0.0f) {
            nrm = v3(-nrm.x, -nrm.y, -nrm.z);
            dp = -dp;
        }
        const Material& mat = materials[mat_indices[res[0]]];
        V3 kd = v3(mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]);
        V3 ks = v3(mat.specular[0], mat.specular[1], mat.specular[2]);
        float pd = std::fmax(std::fmax(kd.x, kd.y), kd.z);
        float ps = std::fmax(std::fmax(ks.x, ks.y), ks.z);
        const float h = mat.shininess;
        if (pd + ps > 1.0f) {
            const float s = 1.0f / (pd + ps);
            kd = scale_r(kd, s);
            ks = scale_r(ks, s);
            pd *= s;
            ps *= s;
        }
        const float r0 = randf(seed), r1 = randf(seed), r2 = randf(seed);
        if (!russian_roulette) {
            const float s = 1.0f / (pd + ps);
            pd *= s;
            ps *= s;
        }
        seeds[task] = seed;
        if (!(r2 <= pd + ps)) continue;
        const V3 refl = normalize(sub(din, scale(2.0f * dp, nrm)));
        const V3 hitp = add(v3(r[0], r[1], r[2]), scale_r(din, as_float(res[1])));
        float pdf = 0.0f, cos_theta = 0.0f, cos_alpha = 0.0f;
        V3 dir = v3(0, 0, 0), w = v3(0, 0, 0);
        if (r2 >= pd) {
            const V3 bv = arbitrary_normal(refl);
            const V3 bu = cross(bv, refl);
            const float sin_t = std::sqrt(1.0f - std::pow(r1, 2.0f / (float)(h + 1.0f)));
            const float cos_t = std::pow(r1, 1.0f / (float)(h + 1.0f));
            const float phi = 2.0f * kPi * r0;
            pdf = (h + 1.0f) / (2.0f * kPi) * std::pow(cos_t, h);
            const V3 hem = v3(std::cos(phi) * sin_t, std::sin(phi) * sin_t, cos_t);
            dir = normalize(add(add(scale(hem.x, bu), scale(hem.y, bv)), scale(hem.z, refl)));
            cos_theta = dot(nrm, dir);
            cos_alpha = dot(refl, dir);
            const V3 brdf = scale_r(scale((h + 2.0f) / (2.0f * kPi), ks), std::pow(cos_alpha, h));
            w = v3(brdf.x * cos_theta / pdf / (pd + ps), brdf.y * cos_theta / pdf / (pd + ps),
                   brdf.z * cos_theta / pdf / (pd + ps));
        }
        if (r2 < pd || (r2 >= pd && (cos_theta < 0.0f || pdf < 0.01f || cos_alpha < 0.0f))) {
            const float theta = 2.0f * kPi * r1;
            const float radius = std::sqrt(r0);
            const float x = radius * std::sin(theta);
            const float z = radius * std::cos(theta);
            const float y = std::sqrt(1.0f - x * x - z * z);
            const V3 bv = arbitrary_normal(nrm);
            const V3 bu = cross(bv, nrm);
            dir = add(add(scale(x, bu), scale(z, bv)), scale(y, nrm));
            cos_theta = dot(nrm, dir);
            pdf = cos_theta / kPi;
            const V3 brdf = v3(kd.x / kPi, kd.y / kPi, kd.z / kPi);
            w = v3(brdf.x * cos_theta / pdf / (pd + ps), brdf.y * cos_theta / pdf / (pd + ps),
                   brdf.z * cos_theta / pdf / (pd + ps));
        }
        const int pixel = in_pixel[task];
        put_ray(out_rays, out, add(hitp, scale_r(dir, 1.0e-4f)), 0.0f, dir, kMaxFloat);
        out_pixel[out] = pixel;
        float* dec = decreases + 3 * (int64_t)pixel;
        dec[0] *= w.x;
        dec[1] *= w.y;
        dec[2] *= w.z;
        ++out;
    }
    return out;
}

/* RendererKernels.cu:302-372 (reconstructShadow / reconstructAO) */
void hxo_reconstruct_occlusion(int64_t begin, int64_t size, int32_t ns, int32_t replace, const int32_t* results,
                               const int32_t* index_to_pixel, const int32_t* index_to_slot, const float* in_pixels,
                               float* out_pixels) {
    for (int64_t task = 0; task < size; ++task) {
        const int pixel = index_to_pixel[begin + task];
        const float* c = in_pixels + 4 * (int64_t)pixel;
        float f = 0.0f;
        for (int i = 0; i < ns; ++i) {
            const int64_t k = task * ns + i;
            const int64_t slot = index_to_slot ? index_to_slot[k] : k;
            if (results[4 * slot] == -1) f += 1.0f;
        }
        f /= (float)ns;
        float* o = out_pixels + 4 * (int64_t)pixel;
        const float v[4] = {c[0] * f, c[1] * f, c[2] * f, 1.0f};
        for (int k = 0; k < 4; ++k) o[k] = replace ? v[k] : o[k] + v[k];
    }
}

/* InterpolatorKernels.cu:10-46 */
void hxo_interpolate(int64_t n_static, int64_t n, float t, const float* a, const float* b, float* out,
                     int32_t normalize_out) {
    for (int64_t i = 0; i < n; ++i) {
        const float s = 1.0f - t;
        V3 r = add(scale_r(ld(a, i), s), scale_r(ld(b, i), t));
        if (normalize_out) r = normalize(r);
        float* o = out + 3 * (n_static + i);
        o[0] = r.x;
        o[1] = r.y;
        o[2] = r.z;
    }
}

}  // extern "C"
