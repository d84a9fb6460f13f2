// Device-side shading library for the gfx950 path tracer: vector math, RNG, encodings,
// frames, BSDF lobes and combinators, materials, environment lookup.
//
// Semantics follow the reference's Slang library (paths relative to the reference root,
// cited per function). Every file including this one is compiled with -ffp-contract=off so
// expressions round as written; fused multiply-adds appear only where written explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/merian_amd.h"

#define MRT_HD __host__ __device__ __forceinline__

#include "mrt_math.h"

namespace mrt {

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};

MRT_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MRT_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MRT_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
MRT_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
MRT_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
MRT_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
MRT_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
MRT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MRT_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
MRT_HD float length(f3 a) { return sqrtf(dot(a, a)); }
MRT_HD f3 normalize(f3 a) {
    const float inv = 1.0f / sqrtf(dot(a, a));
    return a * inv;
}
// SPIR-V FMix: x * (1 - a) + y * a
MRT_HD float lerp(float x, float y, float a) { return x * (1.0f - a) + y * a; }
MRT_HD f3 lerp(f3 x, f3 y, float a) { return {lerp(x.x, y.x, a), lerp(x.y, y.y, a), lerp(x.z, y.z, a)}; }
MRT_HD float pow2(float x) { return x * x; }
MRT_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
MRT_HD float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
MRT_HD f3 ld3(const float* p) { return {p[0], p[1], p[2]}; }
MRT_HD bool finite3(f3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }

constexpr float PI = 3.14159265358979323846f;
constexpr float TWO_PI = 6.283185307179586f;
constexpr float INV_PI = 0.3183098861837907f;
constexpr float INV_2PI = 0.159154943091895335769f;
constexpr float M_1_PI_F = 0.318309886183790671538f;

// ---------------------------------------------------------------------------------------
// RNG: PCG (utils/random.slang:108-128), pcg4d16 (utils/hash.slang:138-146),
// uint_to_zero_one_float_rust (utils/bit-twiddling.slang:23-27)

MRT_HD uint32_t pcg4d16(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    w = w * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x += y * w;
    return x;
}
// pcg3d16 (utils/hash.slang:83-90): the MCPG renderer's per-pixel RNG seed
MRT_HD uint32_t pcg3d16(uint32_t x, uint32_t y, uint32_t z) {
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    x += y * z;
    y += z * x;
    z += x * y;
    x += y * z;
    return x;
}
MRT_HD float unit_float(uint32_t u) { return float(u >> 8u) * 5.9604644775390625e-08f; }
struct PCG {
    uint32_t state;
    MRT_HD uint32_t next() {
        const uint32_t s = state;
        state = state * 747796405u + 2891336453u;
        const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
        return (word >> 22u) ^ word;
    }
    MRT_HD float next_float() { return unit_float(next()); }
    MRT_HD f2 next_float2() {
        f2 r;
        r.x = next_float();
        r.y = next_float();
        return r;
    }
};

// Per-path hit-sequence hash (parity aid): h' = (h ^ (id + c1)) * c2
MRT_HD uint32_t sig_step(uint32_t h, uint32_t id) { return (h ^ (id + 0x7F4A7C15u)) * 0x01000193u; }

// ---------------------------------------------------------------------------------------
// Encodings (utils/encoding.slang:6-55), half floats, ray offset (scene/hit.slang:14-23)

MRT_HD int16_t pack_snorm(float v) { return int16_t(rintf(clampf(v, -1.f, 1.f) * 32767.0f)); }
MRT_HD uint32_t encode_normal(f3 n) {
    const float inv = 1.0f / ((fabsf(n.x) + fabsf(n.y)) + fabsf(n.z));
    float px, py;
    if (n.z < 0.0f) {
        px = copysignf(1.0f - fabsf(n.y * inv), n.x);
        py = copysignf(1.0f - fabsf(n.x * inv), n.y);
    } else {
        px = n.x * inv;
        py = n.y * inv;
    }
    return uint32_t(uint16_t(pack_snorm(px))) | (uint32_t(uint16_t(pack_snorm(py))) << 16);
}
MRT_HD f3 decode_normal(uint32_t enc) {
    const float px = clampf(float(int16_t(enc & 0xFFFFu)) / 32767.0f, -1.f, 1.f);
    const float py = clampf(float(int16_t(enc >> 16)) / 32767.0f, -1.f, 1.f);
    f3 v{px, py, (1.0f - fabsf(px)) - fabsf(py)};
    if (v.z < 0.0f) {
        const float x = v.x, y = v.y;
        v.x = copysignf(1.0f - fabsf(y), x);
        v.y = copysignf(1.0f - fabsf(x), y);
    }
    return normalize(v);
}
MRT_HD float half_to_float(uint16_t h) { return float(__builtin_bit_cast(_Float16, h)); }
MRT_HD uint16_t float_to_half(float f) { return __builtin_bit_cast(uint16_t, _Float16(f)); }

MRT_HD float offset_component(float p, float n) {
    const int32_t io = int32_t(n * 256.f);
    const int32_t bits = __builtin_bit_cast(int32_t, p) + ((p < 0.f) ? -io : io);
    const float ip = __builtin_bit_cast(float, bits);
    return fabsf(p) < (1.f / 16.f) ? p + n * (1.f / 65536.f) : ip;
}
MRT_HD f3 offset_ray_origin(f3 p, f3 n) {
    return {offset_component(p.x, n.x), offset_component(p.y, n.y), offset_component(p.z, n.z)};
}

// ---------------------------------------------------------------------------------------
// Frames (utils/frame.slang:7-135)

struct Frame {
    f3 r0, r1, r2;
    MRT_HD f3 to_local(f3 v) const { return {dot(r0, v), dot(r1, v), dot(r2, v)}; }
    MRT_HD f3 to_world(f3 v) const {
        return {v.x * r0.x + v.y * r1.x + v.z * r2.x, v.x * r0.y + v.y * r1.y + v.z * r2.y,
                v.x * r0.z + v.y * r1.z + v.z * r2.z};
    }
};
MRT_HD Frame frame_from_normal(f3 z) {
    const float s = (z.z >= 0) ? 1.f : -1.f;
    const float a = -1.0f / (s + z.z);
    const float b = z.x * z.y * a;
    Frame f;
    f.r0 = {1.0f + s * z.x * z.x * a, s * b, -s * z.x};
    f.r1 = {b, s + z.y * z.y * a, -z.y};
    f.r2 = z;
    return f;
}
MRT_HD Frame frame_from_nt(f3 n, f3 t, float hs) {
    const f3 N = normalize(n);
    const f3 T = normalize(t - N * dot(t, N));
    const f3 B = cross(N, T) * hs;
    return Frame{T, B, N};
}
MRT_HD Frame apply_normal_map(const Frame& f, f3 nl) {
    const f3 new_n = f.to_world(nl);
    const f3 new_t = normalize(f.r0 - new_n * dot(f.r0, new_n));
    return Frame{new_t, cross(new_n, new_t), new_n};
}

// ---------------------------------------------------------------------------------------
// BSDFs (shading/bsdfs/*.slang). Local frame: n = +z, wi/wo point away from the surface.

constexpr float BSDF_MIN_COS_THETA = 1e-6f; // bsdf.slang:13

struct Sample {
    f3 wo;
    float pdf;
    f3 weight;
    bool valid;
};
MRT_HD Sample none_sample() { return Sample{{0.f, 0.f, 0.f}, 0.f, {0.f, 0.f, 0.f}, false}; }
MRT_HD Sample make_sample(f3 wo, float pdf, f3 w) { return Sample{wo, pdf, w, true}; }

MRT_HD f3 reflect2(f3 i, f3 n, float ndoti) { return ((2.0f * ndoti) * n) - i; }
MRT_HD f3 refract(f3 I, f3 N, float eta) {
    const float d = dot(N, I);
    const float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k < 0.0f) return {0.f, 0.f, 0.f};
    return (eta * I) - ((eta * d + sqrtf(k)) * N);
}
MRT_HD float schlick_factor(float c) { return pow5(fmaxf(1.0f - c, 0.0f)); }
MRT_HD float fresnel_schlick(float c, float f0) { return f0 + (1.0f - f0) * schlick_factor(c); }
MRT_HD f3 fresnel_schlick(float c, f3 f0) {
    const float p = schlick_factor(c);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
MRT_HD float fresnel_dielectric(float cos_i, float eta) {
    const float sin2_t = (eta * eta) * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return 1.0f;
    const float cos_t = sqrtf(1.0f - sin2_t);
    const float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (r_parl * r_parl + r_perp * r_perp);
}
MRT_HD f3 sample_cos(PCG& rng) { // utils/sampling.slang:24-29
    const f2 xi = rng.next_float2();
    const float su = sqrtf(xi.x);
    float sn, cs;
    msincos(TWO_PI * xi.y, sn, cs);
    return {su * cs, su * sn, sqrtf(1.0f - xi.x)};
}

// GGX helpers (brdf-ggx.slang:24-278)
MRT_HD float g1_over_widotn(float widotn, float a2) {
    return 2.0f / (widotn + sqrtf(a2 + ((1.0f - a2) * (widotn * widotn))));
}
MRT_HD float g2_over_widotn(float wodotn, float widotn, float a2) {
    const float g1 = widotn * sqrtf(a2 + ((1.0f - a2) * (wodotn * wodotn)));
    const float g2 = wodotn * sqrtf(a2 + ((1.0f - a2) * (widotn * widotn)));
    return (2.0f * wodotn) / (g1 + g2);
}
MRT_HD float g1_over_widotn(f3 wi, f2 a) { return 2.0f / (wi.z + length(f3{a.x * wi.x, a.y * wi.y, wi.z})); }
MRT_HD float g2_over_widotn(f3 wo, f3 wi, f2 a) {
    const float g1 = wi.z * length(f3{a.x * wo.x, a.y * wo.y, wo.z});
    const float g2 = wo.z * length(f3{a.x * wi.x, a.y * wi.y, wi.z});
    return (2.0f * wo.z) / (g1 + g2);
}
MRT_HD float ggx_ndf(float ndoth2, float a2) { return a2 / (PI * pow2((ndoth2 * a2) + (1.0f - ndoth2))); }
MRT_HD float ggx_ndf(f3 h, f2 a) {
    const float axy = a.x * a.y;
    const f3 v{a.y * h.x, a.x * h.y, axy * h.z};
    const float w2 = axy / dot(v, v);
    return axy * w2 * w2 / PI;
}
MRT_HD f3 vndf_sample_h(f3 wi, f2 a, f2 r) { // Dupuy & Benyoub 2023 (brdf-ggx.slang:137-152)
    const f3 ws = normalize(f3{wi.x * a.x, wi.y * a.y, wi.z});
    const float phi = ((2.0f * r.x) - 1.0f) * PI;
    const float z = __builtin_fmaf(1.0f - r.y, 1.0f + ws.z, -ws.z);
    const float st = sqrtf(clampf(1.0f - (z * z), 0.0f, 1.0f));
    float sn, cs;
    msincos(phi, sn, cs);
    const float x = st * cs;
    const float y = st * sn;
    const f3 wm{x + ws.x, y + ws.y, z + ws.z};
    return normalize(f3{wm.x * a.x, wm.y * a.y, wm.z});
}
MRT_HD float vndf_h_pdf(f3 wi, f3 h, float a2) {
    const float g1 = g1_over_widotn(wi.z, a2);
    const float d = ggx_ndf(pow2(h.z), a2);
    return d * g1 * fmaxf(0.f, dot(wi, h));
}
MRT_HD float vndf_pdf_wo(float widotn, float ndoth2, float a2) {
    const float g1 = g1_over_widotn(widotn, a2);
    const float d = ggx_ndf(ndoth2, a2);
    return (d * g1) / 4.0f;
}
MRT_HD float vndf_pdf_wo(f3 wi, f3 h, f2 a) { return (ggx_ndf(h, a) * g1_over_widotn(wi, a)) / 4.0f; }

// Directional albedo (the GBuffer's demodulation albedo, gbuffer_rt.slang:76-77, reached through
// each BSDF's get_albedo).
//
// Pre-integrated GGX reflectance for the height-correlated Smith G2: the rational fit of Ray
// Tracing Gems ch. 32 eq. 4 (brdf-ggx.slang:283-317), F0 * scale + bias, each a ratio of
// polynomials in c = |cos theta| and alpha. The bias is faded out for F0 near zero.
MRT_HD f3 ggx_albedo_fit(f3 F0, float alpha, float cos_theta) {
    const float c = fabsf(cos_theta);
    const float c2 = c * c, c3 = c2 * c;
    const float a3 = (alpha * alpha) * alpha;
    const float bias_num = (0.995367f + -1.38839f * c) + (-0.24751f + 1.97442f * c) * alpha;
    const float bias_den = ((1.0f + 2.68132f * c) + 52.366f * c3) + ((16.0932f + -3.98452f * c) + 59.3013f * c3) * alpha +
                           ((-5.18731f + 255.259f * c) + 2544.07f * c3) * a3;
    const float scale_num = (-0.0564526f + 3.82901f * c) + (16.91f + -11.0303f * c) * alpha;
    const float scale_den = ((1.0f + 4.11118f * c2) + -1.37886f * c3) + ((19.3254f + -28.9947f * c2) + 16.9514f * c3) * alpha +
                            ((0.545386f + 96.0994f * c2) + -79.4492f * c3) * a3;
    const float luma = ((F0.x + F0.y) + F0.z) * (1.0f / 3.0f);
    const float bias = fmaxf(0.0f, (bias_num / bias_den) * clampf(luma * 50.0f, 0.0f, 1.0f));
    const float scale = fmaxf(0.0f, scale_num / scale_den);
    return {F0.x * scale + bias, F0.y * scale + bias, F0.z * scale + bias};
}

// Hemispherical reflectance of the Estevez & Kulla sheen lobe (brdf-sheen.slang:92-118): per
// alpha row (0, 0.1, .., 1) a fit t(x) = e + x^d + a / (1 + b x^c) of the transmitted fraction in
// x = 1 + 20 cos theta, interpolated linearly between the two rows around alpha.
MRT_HD float sheen_albedo_row(int row, float x) {
    // (a, b, c, d, e) per row
    constexpr float k[11][5] = {
        {1.01108f, 75.2286f, -9.4671f, 0.000973445f, -1.01325f}, {8.44447f, 0.0135054f, 1.11637f, 0.479139f, -9.30873f},
        {5.67124f, 0.0194668f, 1.12489f, 0.453551f, -6.53104f},  {4.50086f, 0.0237758f, 1.09376f, 0.423665f, -5.36521f},
        {3.77781f, 0.0273286f, 1.0666f, 0.400267f, -4.65601f},   {3.35338f, 0.0307392f, 0.989189f, 0.371074f, -4.24089f},
        {3.39605f, 0.0308828f, 0.845455f, 0.333813f, -4.28662f}, {3.62583f, 0.0303531f, 0.79754f, 0.325588f, -4.51725f},
        {4.11524f, 0.0172797f, 0.701166f, 0.275242f, -5.04394f}, {4.4012f, 0.0159469f, 0.685892f, 0.270221f, -5.32997f},
        {5.8912f, 0.0171461f, 0.66366f, 0.288735f, -6.78867f}};
    const float* r = k[row];
    return (r[4] + mpow(x, r[3])) + r[0] / (1.0f + r[1] * mpow(x, r[2]));
}
MRT_HD float sheen_directional_albedo(float cos_theta, float alpha) {
    const float x = clampf(cos_theta, 0.0f, 1.0f) * 20.0f + 1.0f;
    const int lo = int(clampf(floorf(alpha * 10.0f), 0.0f, 10.0f));
    const int hi = lo + 1 < 10 ? lo + 1 : 10;
    const float w_lo = float(hi) - alpha * 10.0f;
    const float t = sheen_albedo_row(lo, x) * w_lo + sheen_albedo_row(hi, x) * (1.0f - w_lo);
    return 1.0f - clampf(t, 0.0f, 1.0f);
}

struct Lambert { // brdf-lambert-diffuse.slang:9-61
    f3 albedo;
    MRT_HD float alpha() const { return 1.0f; } // brdf-lambert-diffuse.slang:58-60
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        return (INV_PI * albedo) * wo.z;
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        return INV_PI * wo.z;
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        const f3 wo = sample_cos(rng);
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return none_sample();
        return make_sample(wo, INV_PI * wo.z, albedo);
    }
    MRT_HD f3 get_albedo(f3) const { return albedo; }
};

struct GGX { // brdf-ggx.slang:331-412
    f2 a;
    MRT_HD float alpha() const { return fminf(a.x, a.y); } // brdf-ggx.slang:409-411
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        const f3 h = normalize(wo + wi);
        const float v = (ggx_ndf(h, a) * g2_over_widotn(wo, wi, a)) / 4.0f;
        return {v, v, v};
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        return vndf_pdf_wo(wi, normalize(wi + wo), a);
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        if (wi.z < BSDF_MIN_COS_THETA) return none_sample();
        const f3 h = vndf_sample_h(wi, a, rng.next_float2());
        const f3 wo = reflect2(wi, h, dot(wi, h));
        if (wo.z < BSDF_MIN_COS_THETA) return none_sample();
        const float p = vndf_pdf_wo(wi, h, a);
        const float w = g2_over_widotn(wo, wi, a) / g1_over_widotn(wi, a);
        return make_sample(wo, p, {w, w, w});
    }
    // the fit is isotropic: sqrt(ax * ay) keeps the NDF peak (brdf-ggx.slang:399-406)
    MRT_HD f3 preintegrated(f3 F0, f3 wi) const { return ggx_albedo_fit(F0, sqrtf(a.x * a.y), wi.z); }
    MRT_HD f3 get_albedo(f3 wi) const { return preintegrated(f3{1.0f, 1.0f, 1.0f}, wi); }
};

struct GGXBTDF { // brdf-ggx.slang:416-503
    float alpha, eta;
    f3 tint;
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA || wo.z > -BSDF_MIN_COS_THETA) return {0, 0, 0};
        f3 h = normalize(wo + (wi * eta));
        h = h * signf(h.z);
        const float widoth = dot(wi, h), wodoth = dot(wo, h);
        const float a2 = pow2(alpha);
        const float g2 = g2_over_widotn(fabsf(wo.z), wi.z, a2);
        const float d = ggx_ndf(pow2(h.z), a2);
        const float denom = wodoth + (eta * widoth);
        const float jac = pow2(eta) * widoth * wodoth / pow2(denom);
        return tint * (d * g2 * fabsf(jac));
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA || wo.z > -BSDF_MIN_COS_THETA) return 0.f;
        f3 h = normalize(wo + (wi * eta));
        h = h * signf(h.z);
        const float wodoth = dot(wo, h);
        const float denom = wodoth + (eta * dot(wi, h));
        return vndf_h_pdf(wi, h, pow2(alpha)) * fabsf(wodoth) / pow2(denom);
    }
};

struct RoughDielectric { // bsdf-rough-dielectric.slang:13-120
    float alpha, eta;
    f3 tint;
    MRT_HD float lobe_alpha() const { return alpha; } // the `alpha` field is the property
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA) return {0, 0, 0};
        if (wo.z > BSDF_MIN_COS_THETA) {
            const f3 h = normalize(wi + wo);
            const float R = fresnel_dielectric(fabsf(dot(wi, h)), eta);
            return R * GGX{{alpha, alpha}}.eval(wi, wo);
        }
        if (wo.z < -BSDF_MIN_COS_THETA) {
            f3 h = normalize(wo + wi * eta);
            h = h * signf(h.z);
            const float R = fresnel_dielectric(fabsf(dot(wi, h)), eta);
            return (1.0f - R) * GGXBTDF{alpha, eta, tint}.eval(wi, wo);
        }
        return {0, 0, 0};
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA) return 0.f;
        const float a2 = pow2(alpha);
        if (wo.z > BSDF_MIN_COS_THETA) {
            const f3 h = normalize(wi + wo);
            const float R = fresnel_dielectric(fabsf(dot(wi, h)), eta);
            return R * vndf_pdf_wo(wi.z, pow2(fmaxf(0.0f, h.z)), a2);
        }
        if (wo.z < -BSDF_MIN_COS_THETA) {
            f3 h = normalize(wo + wi * eta);
            h = h * signf(h.z);
            const float R = fresnel_dielectric(fabsf(dot(wi, h)), eta);
            const float wodoth = dot(wo, h);
            const float denom = wodoth + eta * dot(wi, h);
            return (1.0f - R) * vndf_h_pdf(wi, h, a2) * fabsf(wodoth) / pow2(denom);
        }
        return 0.f;
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        if (wi.z < BSDF_MIN_COS_THETA) return none_sample();
        const float a2 = pow2(alpha);
        const f3 wm = vndf_sample_h(wi, f2{alpha, alpha}, rng.next_float2());
        const float widoth = dot(wi, wm);
        const float R = fresnel_dielectric(fabsf(widoth), eta);
        if (rng.next_float() < R) {
            const f3 wo = reflect2(wi, wm, widoth);
            if (wo.z < BSDF_MIN_COS_THETA) return none_sample();
            const float w = g2_over_widotn(wo.z, wi.z, a2) / g1_over_widotn(wi.z, a2);
            const float p = R * vndf_pdf_wo(wi.z, pow2(wm.z), a2);
            return make_sample(wo, p, {w, w, w});
        }
        const f3 wo = refract(-wi, wm, eta);
        if ((wo.x == 0.f && wo.y == 0.f && wo.z == 0.f) || wo.z > -BSDF_MIN_COS_THETA) return none_sample();
        const float wodoth = dot(wo, wm);
        const float denom = wodoth + eta * widoth;
        const float p = (1.0f - R) * vndf_h_pdf(wi, wm, a2) * fabsf(wodoth) / pow2(denom);
        const float w = pow2(eta) * g2_over_widotn(fabsf(wo.z), wi.z, a2) / g1_over_widotn(wi.z, a2);
        return make_sample(wo, p, tint * w);
    }
    // bsdf-rough-dielectric.slang:111-115: reflected part plus the transmitted eta^2 * tint
    MRT_HD f3 get_albedo(f3 wi) const {
        const float f0 = pow2((1.0f - eta) / (1.0f + eta));
        const f3 spec = GGX{{alpha, alpha}}.preintegrated(f3{f0, f0, f0}, wi);
        const f3 t = pow2(eta) * tint;
        return {spec.x + (1.0f - spec.x) * t.x, spec.y + (1.0f - spec.y) * t.y, spec.z + (1.0f - spec.z) * t.z};
    }
};

struct Sheen { // brdf-sheen.slang:9-90 (Estevez & Kulla 2017)
    f3 color;
    float alpha;
    MRT_HD float lobe_alpha() const { return alpha; } // max(alpha, 1e-6) from the constructor
    MRT_HD float D(float ndoth) const {
        const float inv_alpha = 1.f / alpha;
        const float cos2h = ndoth * ndoth;
        const float sin2h = fmaxf(1.f - cos2h, 0.0078125f);
        return (2.f + inv_alpha) * mpow(sin2h, inv_alpha * 0.5f) * (0.5f * INV_PI);
    }
    MRT_HD float L(float x) const {
        const float r = (1.f - alpha) * (1.f - alpha);
        const float inv_r = 1.f - r;
        const float a = 25.3245f * r + 21.5473f * inv_r;
        const float b = 3.32435f * r + 3.82987f * inv_r;
        const float c = 0.16801f * r + 0.19823f * inv_r;
        const float d = -1.27393f * r - 1.97760f * inv_r;
        const float e = -4.85967f * r - 4.32054f * inv_r;
        return a / (1.f + b * mpow(x, c)) + d * x + e;
    }
    MRT_HD float lambda(float c) const {
        if (c < 0.5f) return mexp(L(c));
        return mexp(2.f * L(0.5f) - L(1.f - c));
    }
    MRT_HD float G(float ndoto, float ndoti) const {
        return 1.f / (1.f + mpow(lambda(ndoto), 1.f + 2.f * mpow(1.f - ndoto, 8.f)) + lambda(ndoti));
    }
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        const f3 h = normalize(wi + wo);
        const float s = D(h.z) * G(wo.z, wi.z);
        return (s * color) / (4.f * wi.z);
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        return INV_PI * wo.z;
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        const f3 wo = sample_cos(rng);
        if (fminf(wi.z, wo.z) < BSDF_MIN_COS_THETA) return none_sample();
        const float p = pdf(wi, wo);
        return make_sample(wo, p, eval(wi, wo) / p);
    }
    MRT_HD f3 get_albedo(f3 wi) const { return color * sheen_directional_albedo(wi.z, alpha); }
};

// ---------------------------------------------------------------------------------------
// Combinators. Mix with a weight of exactly 0 (1) returns the first (second) operand's value:
// x * (1 - w) + y * w == x for w == 0 whenever y is finite, which holds for every lobe here
// (alpha >= 0.0316^2 keeps the GGX terms finite), so the other operand is never evaluated.

// Lobe width of a BSDF (the `alpha` property of bsdf.slang:26): the MCPG renderer guides only
// where it is at least its threshold. Combinators blend their operands' squared widths.
template <class B> MRT_HD float lobe_alpha_of(const B& b) { return b.alpha(); }
MRT_HD float lobe_alpha_of(const RoughDielectric& b) { return b.lobe_alpha(); }
MRT_HD float lobe_alpha_of(const Sheen& b) { return b.lobe_alpha(); }

template <class Layer, class Base> struct FresnelMix { // bsdf-fresnel-mix.slang:7-63
    Layer layer;
    Base base;
    float f0;
    MRT_HD float alpha() const { return sqrtf(lerp(pow2(lobe_alpha_of(base)), pow2(lobe_alpha_of(layer)), f0)); }
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        const f3 h = normalize(wi + wo);
        const float F = fresnel_schlick(dot(wi, h), f0);
        return lerp(base.eval(wi, wo), layer.eval(wi, wo), F);
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        const float F = fresnel_schlick(wi.z, f0);
        return lerp(base.pdf(wi, wo), layer.pdf(wi, wo), F);
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        const float F = fresnel_schlick(wi.z, f0);
        const Sample s = rng.next_float() < F ? layer.sample_eval(wi, rng) : base.sample_eval(wi, rng);
        if (!s.valid) return none_sample();
        const float p = pdf(wi, s.wo);
        return make_sample(s.wo, p, eval(wi, s.wo) / p);
    }
    // bsdf-fresnel-mix.slang:50-53
    MRT_HD f3 get_albedo(f3 wi) const {
        const f3 spec = layer.preintegrated(f3{f0, f0, f0}, wi);
        const f3 b = base.get_albedo(wi);
        return {spec.x + (1.0f - spec.x) * b.x, spec.y + (1.0f - spec.y) * b.y, spec.z + (1.0f - spec.z) * b.z};
    }
};

template <class B0, class B1> struct Mix { // bsdf-mix.slang:6-65
    B0 b0;
    B1 b1;
    float weight;
    MRT_HD float alpha() const { return sqrtf(lerp(pow2(lobe_alpha_of(b0)), pow2(lobe_alpha_of(b1)), weight)); }
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        if (weight == 0.0f) return b0.eval(wi, wo);
        if (weight == 1.0f) return b1.eval(wi, wo);
        return lerp(b0.eval(wi, wo), b1.eval(wi, wo), weight);
    }
    MRT_HD float pdf(f3 wi, f3 wo) const {
        if (weight == 0.0f) return b0.pdf(wi, wo);
        if (weight == 1.0f) return b1.pdf(wi, wo);
        return lerp(b0.pdf(wi, wo), b1.pdf(wi, wo), weight);
    }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        if (weight > 0.0f && rng.next_float() < weight) {
            const Sample s = b1.sample_eval(wi, rng);
            if (!s.valid) return none_sample();
            if (weight == 1.0f) return make_sample(s.wo, s.pdf, (s.weight * s.pdf) / s.pdf);
            const float mix_pdf = lerp(b0.pdf(wi, s.wo), s.pdf, weight);
            const f3 mix_eval = lerp(b0.eval(wi, s.wo), s.weight * s.pdf, weight);
            return make_sample(s.wo, mix_pdf, mix_eval / mix_pdf);
        }
        const Sample s = b0.sample_eval(wi, rng);
        if (!s.valid) return none_sample();
        if (weight == 0.0f) return make_sample(s.wo, s.pdf, (s.weight * s.pdf) / s.pdf);
        const float mix_pdf = lerp(s.pdf, b1.pdf(wi, s.wo), weight);
        const f3 mix_eval = lerp(s.weight * s.pdf, b1.eval(wi, s.wo), weight);
        return make_sample(s.wo, mix_pdf, mix_eval / mix_pdf);
    }
    MRT_HD f3 get_albedo(f3 wi) const { return lerp(b0.get_albedo(wi), b1.get_albedo(wi), weight); } // bsdf-mix.slang:51-54
};

template <class Inner> struct Conductor { // bsdf-conductor-fresnel.slang:6-50
    Inner inner;
    f3 f0;
    MRT_HD float alpha() const { return lobe_alpha_of(inner); }
    MRT_HD f3 eval(f3 wi, f3 wo) const {
        const f3 h = normalize(wi + wo);
        return fresnel_schlick(dot(wi, h), f0) * inner.eval(wi, wo);
    }
    MRT_HD float pdf(f3 wi, f3 wo) const { return inner.pdf(wi, wo); }
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        Sample s = inner.sample_eval(wi, rng);
        if (!s.valid) return none_sample();
        const f3 h = normalize(wi + s.wo);
        s.weight = s.weight * fresnel_schlick(dot(wi, h), f0);
        return s;
    }
    MRT_HD f3 get_albedo(f3 wi) const { return inner.preintegrated(f0, wi); } // bsdf-conductor-fresnel.slang:39-41
};

// OpenPBR tree (openpbr-material.slang:19-25)
using OpaqueDielectric = FresnelMix<GGX, Lambert>;
using Dielectric = Mix<OpaqueDielectric, RoughDielectric>;
using Metal = Conductor<GGX>;
using Substrate = Mix<Dielectric, Metal>;
using Coated = FresnelMix<GGX, Substrate>;
using WithCoat = Mix<Substrate, Coated>;
using OpenPBRBSDF = Mix<WithCoat, Sheen>;

constexpr float MIN_ROUGHNESS = 0.0316f; // materials/material.slang:23

struct ShadingData {
    Frame frame;
    float ior;
    bool front_facing;
};

// A material evaluated at a shading point (material-system.slang:64-72).
struct MaterialSample {
    bool diffuse;
    Lambert lambert;
    OpenPBRBSDF pbr;
    float eta;
    MRT_HD Sample sample_eval(f3 wi, PCG& rng) const {
        return diffuse ? lambert.sample_eval(wi, rng) : pbr.sample_eval(wi, rng);
    }
    MRT_HD f3 eval(f3 wi, f3 wo) const { return diffuse ? lambert.eval(wi, wo) : pbr.eval(wi, wo); }
    MRT_HD float pdf(f3 wi, f3 wo) const { return diffuse ? lambert.pdf(wi, wo) : pbr.pdf(wi, wo); }
    MRT_HD float alpha() const { return diffuse ? lambert.alpha() : pbr.alpha(); }
    MRT_HD f3 get_albedo(f3 wi) const { return diffuse ? lambert.get_albedo(wi) : pbr.get_albedo(wi); }
};

MRT_HD f3 material_emission(const mrt_material& m, bool front_facing) {
    if (m.model == MRT_MAT_DIFFUSE || !front_facing) return {0, 0, 0};
    return ld3(m.emission);
}

// DiffuseMaterial::get_sample (diffuse-material.slang:163-178) and
// OpenPBRMaterial::get_sample (openpbr-material.slang:106-209), untextured.
MRT_HD MaterialSample get_material_sample(const mrt_material& m, const ShadingData& sd, f3& emission,
                                          Frame& shading_frame) {
    MaterialSample s;
    if (m.model == MRT_MAT_DIFFUSE) {
        s.diffuse = true;
        emission = {0, 0, 0};
        shading_frame = apply_normal_map(sd.frame, f3{0, 0, 1});
        s.lambert = Lambert{ld3(m.base_color)};
        s.eta = 1.0f;
        return s;
    }
    s.diffuse = false;
    const bool has_tr = (m.features & MRT_PBR_TRANSMISSION) != 0;
    const bool has_coat = (m.features & MRT_PBR_CLEARCOAT) != 0;
    const bool has_sheen = (m.features & MRT_PBR_SHEEN) != 0;
    const f3 base = ld3(m.base_color);
    const float min_alpha = pow2(MIN_ROUGHNESS);
    const f2 alpha{fmaxf(m.specular_alpha[0], min_alpha), fmaxf(m.specular_alpha[1], min_alpha)};
    shading_frame = sd.frame;
    const float ior_t = (m.specular_ior - 1.0f) / (m.specular_ior + 1.0f);
    const float dielectric_f0 = clampf(m.specular_weight * ior_t * ior_t, 0.0f, 1.0f);
    float eta = 1.0f, tw = 0.0f;
    f3 tcol{1.0f, 1.0f, 1.0f};
    if (has_tr) {
        eta = sd.ior / (sd.front_facing ? m.specular_ior : 1.0f);
        tw = m.transmission_weight;
        tcol = ld3(m.transmission_color);
    }
    const float alpha_iso = has_tr ? sqrtf(alpha.x * alpha.y) : alpha.x;
    const Dielectric diel{OpaqueDielectric{GGX{alpha}, Lambert{base}, dielectric_f0},
                          RoughDielectric{alpha_iso, eta, tcol}, tw};
    const Substrate substrate{diel, Metal{GGX{alpha}, base}, m.metalness};
    float cw = 0.0f, coat_alpha = min_alpha, coat_f0 = 0.0f;
    if (has_coat) {
        cw = m.coat_weight;
        coat_alpha = fmaxf(m.coat_alpha, min_alpha);
        const float ct = (m.coat_ior - 1.0f) / (m.coat_ior + 1.0f);
        coat_f0 = ct * ct;
    }
    float sw = 0.0f, sheen_alpha = 0.0f;
    f3 sheen_col{0.0f, 0.0f, 0.0f};
    if (has_sheen) {
        sw = m.sheen_weight;
        sheen_col = ld3(m.sheen_color);
        sheen_alpha = m.sheen_alpha;
    }
    s.pbr = OpenPBRBSDF{WithCoat{substrate, Coated{GGX{{coat_alpha, coat_alpha}}, substrate, coat_f0}, cw},
                        Sheen{sheen_col, fmaxf(sheen_alpha, 1e-6f)}, sw};
    emission = material_emission(m, sd.front_facing);
    s.eta = eta;
    return s;
}

// yuv_luminance (colors/colorspace_yuv.slang:5-7)
MRT_HD float yuv_luminance(f3 c) { return dot(c, f3{0.299f, 0.587f, 0.114f}); }

} // namespace mrt
