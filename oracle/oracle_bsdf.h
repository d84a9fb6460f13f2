// CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle_math.h).
//
// BSDF lobes, combinators and the OpenPBR / Diffuse material trees, restated from the
// reference's Slang operation by operation. Combinators evaluate literally (no shortcut for
// zero mixing weights), so the product's shortcuts are checked against the literal result.
#pragma once

#include "oracle_math.h"

namespace orc {

// ---------------------------------------------------------------------------------------
// RNG: PCG (include/merian-shaders/utils/random.slang:108-128), pcg4d16
// (utils/hash.slang:138-146), uint_to_zero_one_float_rust (utils/bit-twiddling.slang:23-27).

inline uint32_t pcg4d16(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
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

inline float u32_to_unit_float(uint32_t u) { return float(u >> 8u) * 5.9604644775390625e-08f; }

struct PCG {
    uint32_t state;
    uint32_t next() {
        const uint32_t s = state;
        state = state * 747796405u + 2891336453u;
        const uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
        return (word >> 22u) ^ word;
    }
    float next_float() { return u32_to_unit_float(next()); }
    f2 next_float2() {
        f2 r;
        r.x = next_float();
        r.y = next_float();
        return r;
    }
};

// ---------------------------------------------------------------------------------------

constexpr float BSDF_MIN_COS_THETA = 1e-6f; // shading/bsdfs/bsdf.slang:13

enum Lobe { LOBE_DELTA = 0, LOBE_GLOSSY = 1, LOBE_DIFFUSE = 2 };

struct Sample {
    bool valid;
    f3 wo;
    float pdf;
    f3 weight;
    int lobe;
};
inline Sample none() { return Sample{false, {0, 0, 0}, 0.f, {0, 0, 0}, 0}; }
inline Sample some(f3 wo, float pdf, f3 w, int lobe) { return Sample{true, wo, pdf, w, lobe}; }

// reflect2(i, n, ndoti) = 2 * ndoti * n - i (utils/math.slang:107-109)
inline f3 reflect2(f3 i, f3 n, float ndoti) { return ((2.0f * ndoti) * n) - i; }

// HLSL/GLSL refract(I, N, eta)
inline f3 refract(f3 I, f3 N, float eta) {
    const float d = dot(N, I);
    const float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k < 0.0f) return {0.f, 0.f, 0.f};
    return (eta * I) - ((eta * d + std::sqrt(k)) * N);
}

// utils/fresnel.slang:14-26; pow(max(1 - c, 0), 5)
inline float schlick_factor(float c) { return o_pow5(std::fmax(1.0f - c, 0.0f)); }
inline float fresnel_schlick(float c, float f0) { return f0 + (1.0f - f0) * schlick_factor(c); }
inline f3 fresnel_schlick(float c, f3 f0) {
    const float p = schlick_factor(c);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
// utils/fresnel.slang:33-42
inline float fresnel_dielectric(float cos_i, float eta) {
    const float sin2_t = (eta * eta) * (1.0f - cos_i * cos_i);
    if (sin2_t >= 1.0f) return 1.0f;
    const float cos_t = std::sqrt(1.0f - sin2_t);
    const float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    const float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (r_parl * r_parl + r_perp * r_perp);
}

// utils/sampling.slang:24-29
inline f3 sample_cos(PCG& rng) {
    const f2 xi = rng.next_float2();
    const float su = std::sqrt(xi.x);
    float sn, cs;
    o_sincos(TWO_PI * xi.y, sn, cs);
    return {su * cs, su * sn, std::sqrt(1.0f - xi.x)};
}

// ---------------------------------------------------------------------------------------
// GGX helpers (shading/bsdfs/brdf-ggx.slang)

// :33-42 smith_g1_over_widotn(widotn, alpha_squared)
inline float g1_over_widotn(float widotn, float a2) {
    return 2.0f / (widotn + std::sqrt(a2 + ((1.0f - a2) * (widotn * widotn))));
}
// :49-53
inline float g2_over_widotn(float wodotn, float widotn, float a2) {
    const float g1 = widotn * std::sqrt(a2 + ((1.0f - a2) * (wodotn * wodotn)));
    const float g2 = wodotn * std::sqrt(a2 + ((1.0f - a2) * (widotn * widotn)));
    return (2.0f * wodotn) / (g1 + g2);
}
// :69-71
inline float g1_over_widotn(f3 wi, f2 a) {
    return 2.0f / (wi.z + length(f3{a.x * wi.x, a.y * wi.y, wi.z}));
}
// :76-80
inline float g2_over_widotn(f3 wo, f3 wi, f2 a) {
    const float g1 = wi.z * length(f3{a.x * wo.x, a.y * wo.y, wo.z});
    const float g2 = wo.z * length(f3{a.x * wi.x, a.y * wi.y, wi.z});
    return (2.0f * wo.z) / (g1 + g2);
}
// :86-89
inline float ggx_ndf(float ndoth2, float a2) {
    return a2 / (PI * pow2((ndoth2 * a2) + (1.0f - ndoth2)));
}
// :93-98
inline float ggx_ndf(f3 h, f2 a) {
    const float axy = a.x * a.y;
    const f3 v{a.y * h.x, a.x * h.y, axy * h.z};
    const float w2 = axy / dot(v, v);
    return axy * w2 * w2 / PI;
}
// :137-152 (Dupuy & Benyoub 2023), mad() taken as a fused multiply-add
inline f3 vndf_sample_h(f3 wi, f2 a, f2 r) {
    const f3 wi_std = normalize(f3{wi.x * a.x, wi.y * a.y, wi.z});
    const float phi = ((2.0f * r.x) - 1.0f) * PI;
    const float z = std::fma(1.0f - r.y, 1.0f + wi_std.z, -wi_std.z);
    const float sin_theta = std::sqrt(clampf(1.0f - (z * z), 0.0f, 1.0f));
    float sn, cs;
    o_sincos(phi, sn, cs);
    const float x = sin_theta * cs;
    const float y = sin_theta * sn;
    const f3 wm_std{x + wi_std.x, y + wi_std.y, z + wi_std.z};
    return normalize(f3{wm_std.x * a.x, wm_std.y * a.y, wm_std.z});
}
// :223-239 ggx_VNDF_H_pdf(wi, h, alpha_squared)
inline float vndf_h_pdf(f3 wi, f3 h, float a2) {
    const float g1 = g1_over_widotn(wi.z, a2);
    const float d = ggx_ndf(pow2(h.z), a2);
    return d * g1 * std::fmax(0.f, dot(wi, h));
}
// :246-251
inline float vndf_pdf_wo(float widotn, float ndoth2, float a2) {
    const float g1 = g1_over_widotn(widotn, a2);
    const float d = ggx_ndf(ndoth2, a2);
    return (d * g1) / 4.0f;
}
// :276-278
inline float vndf_pdf_wo(f3 wi, f3 h, f2 a) { return (ggx_ndf(h, a) * g1_over_widotn(wi, a)) / 4.0f; }

// ---------------------------------------------------------------------------------------
// Directional albedo (BSDF::get_albedo; the GBuffer albedo of gbuffer_rt.slang:76-77, floored and
// divided out by render_pt.slang:104-116 under "demodulate albedo")

// Pre-integrated GGX with height-correlated G2 (brdf-ggx.slang:283-317, Ray Tracing Gems ch. 32
// eq. 4): F0 * scale + bias. The matrices of the fit are written out as polynomials:
// bias = [M1 (1,c)].(1,a) / [M2 (1,c,c^3)].(1,a,a^3), scale = [M3 (1,c)].(1,a) / [M4 (1,c^2,c^3)].(1,a,a^3).
inline f3 ggx_albedo_fit(f3 F0, float alpha, float cos_theta) {
    const float c = std::fabs(cos_theta);
    const float c2 = c * c, c3 = c2 * c;
    const float a3 = (alpha * alpha) * alpha;
    const float bias_num = (0.995367f + -1.38839f * c) + (-0.24751f + 1.97442f * c) * alpha;
    const float bias_den = ((1.0f + 2.68132f * c) + 52.366f * c3) + ((16.0932f + -3.98452f * c) + 59.3013f * c3) * alpha +
                           ((-5.18731f + 255.259f * c) + 2544.07f * c3) * a3;
    const float scale_num = (-0.0564526f + 3.82901f * c) + (16.91f + -11.0303f * c) * alpha;
    const float scale_den = ((1.0f + 4.11118f * c2) + -1.37886f * c3) + ((19.3254f + -28.9947f * c2) + 16.9514f * c3) * alpha +
                            ((0.545386f + 96.0994f * c2) + -79.4492f * c3) * a3;
    const float luma = ((F0.x + F0.y) + F0.z) * (1.0f / 3.0f);
    const float bias = std::fmax(0.0f, (bias_num / bias_den) * clampf(luma * 50.0f, 0.0f, 1.0f));
    const float scale = std::fmax(0.0f, scale_num / scale_den);
    return {F0.x * scale + bias, F0.y * scale + bias, F0.z * scale + bias};
}

// Sheen hemispherical reflectance (brdf-sheen.slang:92-118): 1 - the fitted transmitted fraction
// e + x^d + a / (1 + b x^c), x = 1 + 20 cos theta, linear between the alpha rows 0, 0.1, .., 1.
inline float sheen_directional_albedo(float cos_theta, float alpha) {
    static const float k[11][5] = {
        {1.01108f, 75.2286f, -9.4671f, 0.000973445f, -1.01325f}, {8.44447f, 0.0135054f, 1.11637f, 0.479139f, -9.30873f},
        {5.67124f, 0.0194668f, 1.12489f, 0.453551f, -6.53104f},  {4.50086f, 0.0237758f, 1.09376f, 0.423665f, -5.36521f},
        {3.77781f, 0.0273286f, 1.0666f, 0.400267f, -4.65601f},   {3.35338f, 0.0307392f, 0.989189f, 0.371074f, -4.24089f},
        {3.39605f, 0.0308828f, 0.845455f, 0.333813f, -4.28662f}, {3.62583f, 0.0303531f, 0.79754f, 0.325588f, -4.51725f},
        {4.11524f, 0.0172797f, 0.701166f, 0.275242f, -5.04394f}, {4.4012f, 0.0159469f, 0.685892f, 0.270221f, -5.32997f},
        {5.8912f, 0.0171461f, 0.66366f, 0.288735f, -6.78867f}};
    const float x = clampf(cos_theta, 0.0f, 1.0f) * 20.0f + 1.0f;
    const int lo = int(clampf(std::floor(alpha * 10.0f), 0.0f, 10.0f));
    const int hi = std::min(lo + 1, 10);
    const float w_lo = float(hi) - alpha * 10.0f;
    auto row = [&](int r) { return (k[r][4] + o_pow(x, k[r][3])) + k[r][0] / (1.0f + k[r][1] * o_pow(x, k[r][2])); };
    const float t = row(lo) * w_lo + row(hi) * (1.0f - w_lo);
    return 1.0f - clampf(t, 0.0f, 1.0f);
}

// ---------------------------------------------------------------------------------------
// Lobes

// brdf-lambert-diffuse.slang:9-61
struct Lambert {
    f3 albedo;
    float alpha() const { return 1.0f; }
    f3 eval(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        return (INV_PI * albedo) * wo.z;
    }
    float pdf(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        return INV_PI * wo.z;
    }
    Sample sample_eval(f3 wi, PCG& rng) const {
        const f3 wo = sample_cos(rng);
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return none();
        return some(wo, INV_PI * wo.z, albedo, LOBE_DIFFUSE);
    }
    f3 get_albedo(f3) const { return albedo; } // brdf-lambert-diffuse.slang:50-52
};

// brdf-ggx.slang:331-412
struct GGX {
    f2 a;
    float alpha() const { return std::fmin(a.x, a.y); }
    f3 eval(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        const f3 h = normalize(wo + wi);
        const float g2 = g2_over_widotn(wo, wi, a);
        const float d = ggx_ndf(h, a);
        const float v = (d * g2) / 4.0f;
        return {v, v, v};
    }
    float pdf(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        const f3 h = normalize(wi + wo);
        return vndf_pdf_wo(wi, h, a);
    }
    Sample sample_eval(f3 wi, PCG& rng) const {
        if (wi.z < BSDF_MIN_COS_THETA) return none();
        const f3 h = vndf_sample_h(wi, a, rng.next_float2());
        const f3 wo = reflect2(wi, h, dot(wi, h));
        if (wo.z < BSDF_MIN_COS_THETA) return none();
        const float p = vndf_pdf_wo(wi, h, a);
        const float w = g2_over_widotn(wo, wi, a) / g1_over_widotn(wi, a);
        return some(wo, p, {w, w, w}, LOBE_GLOSSY);
    }
    // brdf-ggx.slang:399-406 (isotropic fit at sqrt(ax * ay))
    f3 preintegrated(f3 F0, f3 wi) const { return ggx_albedo_fit(F0, std::sqrt(a.x * a.y), wi.z); }
    f3 get_albedo(f3 wi) const { return preintegrated(f3{1.0f, 1.0f, 1.0f}, wi); }
};

// brdf-ggx.slang:416-503
struct GGXBTDF {
    float alpha, eta;
    f3 tint;
    f3 eval(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA || wo.z > -BSDF_MIN_COS_THETA) return {0, 0, 0};
        f3 h = normalize(wo + (wi * eta));
        h = h * signf(h.z);
        const float widoth = dot(wi, h);
        const float wodoth = dot(wo, h);
        const float a2 = pow2(alpha);
        const float g2 = g2_over_widotn(std::fabs(wo.z), wi.z, a2);
        const float d = ggx_ndf(pow2(h.z), a2);
        const float denom = wodoth + (eta * widoth);
        const float jac = pow2(eta) * widoth * wodoth / pow2(denom);
        return tint * (d * g2 * std::fabs(jac));
    }
    float pdf(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA || wo.z > -BSDF_MIN_COS_THETA) return 0.f;
        f3 h = normalize(wo + (wi * eta));
        h = h * signf(h.z);
        const float wodoth = dot(wo, h);
        const float denom = wodoth + (eta * dot(wi, h));
        return vndf_h_pdf(wi, h, pow2(alpha)) * std::fabs(wodoth) / pow2(denom);
    }
    Sample sample_eval(f3 wi, PCG& rng) const {
        if (wi.z < BSDF_MIN_COS_THETA) return none();
        const float a2 = pow2(alpha);
        const f3 h = vndf_sample_h(wi, f2{alpha, alpha}, rng.next_float2());
        const f3 wo = refract(-wi, h, eta);
        if ((wo.x == 0.f && wo.y == 0.f && wo.z == 0.f) || wo.z > -BSDF_MIN_COS_THETA) return none();
        const float wodoth = dot(wo, h);
        const float denom = wodoth + (eta * dot(wi, h));
        const float p = vndf_h_pdf(wi, h, a2) * std::fabs(wodoth) / pow2(denom);
        const float w = pow2(eta) * g2_over_widotn(std::fabs(wo.z), wi.z, a2) / g1_over_widotn(wi.z, a2);
        return some(wo, p, tint * w, LOBE_GLOSSY);
    }
    f3 get_albedo(f3) const { return pow2(eta) * tint; } // brdf-ggx.slang:496-498
};

// bsdf-rough-dielectric.slang:13-120
struct RoughDielectric {
    float alpha, eta;
    f3 tint;
    float width() const { return alpha; }
    f3 eval(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA) return {0, 0, 0};
        if (wo.z > BSDF_MIN_COS_THETA) {
            const f3 h = normalize(wi + wo);
            const float R = fresnel_dielectric(std::fabs(dot(wi, h)), eta);
            return R * GGX{{alpha, alpha}}.eval(wi, wo);
        }
        if (wo.z < -BSDF_MIN_COS_THETA) {
            f3 h = normalize(wo + wi * eta);
            h = h * signf(h.z);
            const float R = fresnel_dielectric(std::fabs(dot(wi, h)), eta);
            return (1.0f - R) * GGXBTDF{alpha, eta, tint}.eval(wi, wo);
        }
        return {0, 0, 0};
    }
    float pdf(f3 wi, f3 wo) const {
        if (wi.z < BSDF_MIN_COS_THETA) return 0.f;
        const float a2 = pow2(alpha);
        if (wo.z > BSDF_MIN_COS_THETA) {
            const f3 h = normalize(wi + wo);
            const float R = fresnel_dielectric(std::fabs(dot(wi, h)), eta);
            return R * vndf_pdf_wo(wi.z, pow2(std::fmax(0.0f, h.z)), a2);
        }
        if (wo.z < -BSDF_MIN_COS_THETA) {
            f3 h = normalize(wo + wi * eta);
            h = h * signf(h.z);
            const float R = fresnel_dielectric(std::fabs(dot(wi, h)), eta);
            const float wodoth = dot(wo, h);
            const float denom = wodoth + eta * dot(wi, h);
            return (1.0f - R) * vndf_h_pdf(wi, h, a2) * std::fabs(wodoth) / pow2(denom);
        }
        return 0.f;
    }
    Sample sample_eval(f3 wi, PCG& rng) const {
        if (wi.z < BSDF_MIN_COS_THETA) return none();
        const float a2 = pow2(alpha);
        const f3 wm = vndf_sample_h(wi, f2{alpha, alpha}, rng.next_float2());
        const float widoth = dot(wi, wm);
        const float R = fresnel_dielectric(std::fabs(widoth), eta);
        if (rng.next_float() < R) {
            const f3 wo = reflect2(wi, wm, widoth);
            if (wo.z < BSDF_MIN_COS_THETA) return none();
            const float w = g2_over_widotn(wo.z, wi.z, a2) / g1_over_widotn(wi.z, a2);
            const float p = R * vndf_pdf_wo(wi.z, pow2(wm.z), a2);
            return some(wo, p, {w, w, w}, LOBE_GLOSSY);
        }
        const f3 wo = refract(-wi, wm, eta);
        if ((wo.x == 0.f && wo.y == 0.f && wo.z == 0.f) || wo.z > -BSDF_MIN_COS_THETA) return none();
        const float wodoth = dot(wo, wm);
        const float denom = wodoth + eta * widoth;
        const float p = (1.0f - R) * vndf_h_pdf(wi, wm, a2) * std::fabs(wodoth) / pow2(denom);
        const float w = pow2(eta) * g2_over_widotn(std::fabs(wo.z), wi.z, a2) / g1_over_widotn(wi.z, a2);
        return some(wo, p, tint * w, LOBE_GLOSSY);
    }
    // bsdf-rough-dielectric.slang:111-115
    f3 get_albedo(f3 wi) const {
        const float f0 = pow2((1.0f - eta) / (1.0f + eta));
        const f3 spec = GGX{{alpha, alpha}}.preintegrated(f3{f0, f0, f0}, wi);
        const f3 t = pow2(eta) * tint;
        return {spec.x + (1.0f - spec.x) * t.x, spec.y + (1.0f - spec.y) * t.y, spec.z + (1.0f - spec.z) * t.z};
    }
};

// brdf-sheen.slang:9-90 (Estevez & Kulla 2017)
struct Sheen {
    f3 color;
    float alpha;
    float width() const { return alpha; }
    static Sheen make(f3 c, float a) { return Sheen{c, std::fmax(a, 1e-6f)}; }
    float D(float ndoth) const {
        const float inv_alpha = 1.f / alpha;
        const float cos2h = ndoth * ndoth;
        const float sin2h = std::fmax(1.f - cos2h, 0.0078125f);
        return (2.f + inv_alpha) * o_pow(sin2h, inv_alpha * 0.5f) * (0.5f * INV_PI);
    }
    float L(float x) const {
        const float r = (1.f - alpha) * (1.f - alpha);
        const float inv_r = 1.f - r;
        const float a = 25.3245f * r + 21.5473f * inv_r;
        const float b = 3.32435f * r + 3.82987f * inv_r;
        const float c = 0.16801f * r + 0.19823f * inv_r;
        const float d = -1.27393f * r - 1.97760f * inv_r;
        const float e = -4.85967f * r - 4.32054f * inv_r;
        return a / (1.f + b * o_pow(x, c)) + d * x + e;
    }
    float lambda(float c) const {
        if (c < 0.5f) return o_exp(L(c));
        return o_exp(2.f * L(0.5f) - L(1.f - c));
    }
    float G(float ndoto, float ndoti) const {
        return 1.f / (1.f + o_pow(lambda(ndoto), 1.f + 2.f * o_pow(1.f - ndoto, 8.f)) + lambda(ndoti));
    }
    f3 eval(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return {0, 0, 0};
        const f3 h = normalize(wi + wo);
        const float s = D(h.z) * G(wo.z, wi.z);
        return (s * color) / (4.f * wi.z);
    }
    float pdf(f3 wi, f3 wo) const {
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return 0.f;
        return INV_PI * wo.z;
    }
    // sample_eval_impl (shading-function.slang:134-143) over sample = cosine sampling
    Sample sample_eval(f3 wi, PCG& rng) const {
        const f3 wo = sample_cos(rng);
        if (std::fmin(wi.z, wo.z) < BSDF_MIN_COS_THETA) return none();
        const float p = pdf(wi, wo);
        return some(wo, p, eval(wi, wo) / p, LOBE_DIFFUSE);
    }
    f3 get_albedo(f3 wi) const { return color * sheen_directional_albedo(wi.z, alpha); } // brdf-sheen.slang:83-85
};

// The `alpha` property (bsdf.slang:26): lobe width; combinators interpolate squared widths.
template <class B> inline float width_of(const B& b) { return b.alpha(); }
inline float width_of(const RoughDielectric& b) { return b.width(); }
inline float width_of(const Sheen& b) { return b.width(); }

// bsdf-fresnel-mix.slang:7-63
template <class Layer, class Base> struct FresnelMix {
    Layer layer;
    Base base;
    float f0;
    float alpha() const { return std::sqrt(lerp(pow2(width_of(base)), pow2(width_of(layer)), f0)); }
    f3 eval(f3 wi, f3 wo) const {
        const f3 h = normalize(wi + wo);
        const float F = fresnel_schlick(dot(wi, h), f0);
        return lerp(base.eval(wi, wo), layer.eval(wi, wo), F);
    }
    float pdf(f3 wi, f3 wo) const {
        const float F = fresnel_schlick(wi.z, f0);
        return lerp(base.pdf(wi, wo), layer.pdf(wi, wo), F);
    }
    Sample sample_eval(f3 wi, PCG& rng) const {
        const float F = fresnel_schlick(wi.z, f0);
        const Sample s = rng.next_float() < F ? layer.sample_eval(wi, rng) : base.sample_eval(wi, rng);
        if (!s.valid) return none();
        const float p = pdf(wi, s.wo);
        return some(s.wo, p, eval(wi, s.wo) / p, s.lobe);
    }
    // bsdf-fresnel-mix.slang:50-53
    f3 get_albedo(f3 wi) const {
        const f3 spec = layer.preintegrated(f3{f0, f0, f0}, wi);
        const f3 b = base.get_albedo(wi);
        return {spec.x + (1.0f - spec.x) * b.x, spec.y + (1.0f - spec.y) * b.y, spec.z + (1.0f - spec.z) * b.z};
    }
};

// bsdf-mix.slang:6-65
template <class B0, class B1> struct Mix {
    B0 b0;
    B1 b1;
    float weight;
    float alpha() const { return std::sqrt(lerp(pow2(width_of(b0)), pow2(width_of(b1)), weight)); }
    f3 eval(f3 wi, f3 wo) const { return lerp(b0.eval(wi, wo), b1.eval(wi, wo), weight); }
    float pdf(f3 wi, f3 wo) const { return lerp(b0.pdf(wi, wo), b1.pdf(wi, wo), weight); }
    Sample sample_eval(f3 wi, PCG& rng) const {
        if (weight > 0.0f && rng.next_float() < weight) {
            const Sample s = b1.sample_eval(wi, rng);
            if (!s.valid) return none();
            const float mix_pdf = lerp(b0.pdf(wi, s.wo), s.pdf, weight);
            const f3 mix_eval = lerp(b0.eval(wi, s.wo), s.weight * s.pdf, weight);
            return some(s.wo, mix_pdf, mix_eval / mix_pdf, s.lobe);
        }
        const Sample s = b0.sample_eval(wi, rng);
        if (!s.valid) return none();
        const float mix_pdf = lerp(s.pdf, b1.pdf(wi, s.wo), weight);
        const f3 mix_eval = lerp(s.weight * s.pdf, b1.eval(wi, s.wo), weight);
        return some(s.wo, mix_pdf, mix_eval / mix_pdf, s.lobe);
    }
    f3 get_albedo(f3 wi) const { return lerp(b0.get_albedo(wi), b1.get_albedo(wi), weight); } // bsdf-mix.slang:51-54
};

// bsdf-conductor-fresnel.slang:6-50
template <class Inner> struct Conductor {
    Inner inner;
    f3 f0;
    float alpha() const { return width_of(inner); }
    f3 eval(f3 wi, f3 wo) const {
        const f3 h = normalize(wi + wo);
        return fresnel_schlick(dot(wi, h), f0) * inner.eval(wi, wo);
    }
    float pdf(f3 wi, f3 wo) const { return inner.pdf(wi, wo); }
    Sample sample_eval(f3 wi, PCG& rng) const {
        Sample s = inner.sample_eval(wi, rng);
        if (!s.valid) return none();
        const f3 h = normalize(wi + s.wo);
        s.weight = s.weight * fresnel_schlick(dot(wi, h), f0);
        return s;
    }
    f3 get_albedo(f3 wi) const { return inner.preintegrated(f0, wi); } // bsdf-conductor-fresnel.slang:39-41
};

// OpenPBR tree (openpbr-material.slang:19-25)
using OpaqueDielectric = FresnelMix<GGX, Lambert>;
using Dielectric = Mix<OpaqueDielectric, RoughDielectric>;
using Metal = Conductor<GGX>;
using Substrate = Mix<Dielectric, Metal>;
using Coated = FresnelMix<GGX, Substrate>;
using WithCoat = Mix<Substrate, Coated>;
using OpenPBRBSDF = Mix<WithCoat, Sheen>;

} // namespace orc
