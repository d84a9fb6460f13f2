// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Restatement of the reference's Slang path-tracing path in plain C++ (gcc, compiled with
// -ffp-contract=off so every expression rounds exactly as written). Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product never does.
//
// This header: scalar/vector helpers shared by the oracle's translation units.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

struct f2 {
    float x, y;
};
struct f3 {
    float x, y, z;
};

inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
inline f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline f3 operator/(f3 a, f3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
// dot(a, b) = (a.x*b.x + a.y*b.y) + a.z*b.z
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float length(f3 a) { return std::sqrt(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v)))
inline f3 normalize(f3 a) {
    const float inv = 1.0f / std::sqrt(dot(a, a));
    return a * inv;
}
// lerp = SPIR-V FMix: x * (1 - a) + y * a
inline float lerp(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline f3 lerp(f3 x, f3 y, float a) { return {lerp(x.x, y.x, a), lerp(x.y, y.y, a), lerp(x.z, y.z, a)}; }
inline float pow2(float x) { return x * x; }
inline float clampf(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }
inline float signf(float x) { return x > 0.f ? 1.f : (x < 0.f ? -1.f : 0.f); }
inline bool finite3(f3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

inline uint32_t asuint(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float asfloat(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Constants of include/merian-shaders/utils/math-constants.slangh, as float literals.
constexpr float PI = 3.14159265358979323846f;
constexpr float TWO_PI = 6.283185307179586f;
constexpr float INV_PI = 0.3183098861837907f;
constexpr float INV_2PI = 0.159154943091895335769f;
constexpr float M_1_PI_F = 0.318309886183790671538f;

// ---------------------------------------------------------------------------------------
// Elementary functions from +, -, *, /, sqrt, rint and bit operations only (the product's
// device library evaluates the same sequences, so both round identically). Accuracy against
// float64 is pinned in tests/test_oracle_math.py.
//   sin/cos : Cody-Waite reduction by pi/2 (3-part constant), Taylor degree 9 / 10
//   atan    : 1/x and pi/6 reductions to |a| <= tan(pi/12), odd Taylor series to degree 13
//   log2    : exponent split + atanh series; exp2: rint split + Taylor degree 7

inline void o_sincos(float x, float& s, float& c) {
    const float k = std::rint(x * 0.63661975f);
    const float r = ((x - k * 1.5703125f) - k * 4.837512969970703125e-4f) - k * 7.549790126404332e-08f;
    const float r2 = r * r;
    float ps = 1.0f / 362880.0f;
    ps = ps * r2 + (-1.0f / 5040.0f);
    ps = ps * r2 + (1.0f / 120.0f);
    ps = ps * r2 + (-1.0f / 6.0f);
    const float sr = r + (r * r2) * ps;
    float pc = -1.0f / 3628800.0f;
    pc = pc * r2 + (1.0f / 40320.0f);
    pc = pc * r2 + (-1.0f / 720.0f);
    pc = pc * r2 + (1.0f / 24.0f);
    pc = pc * r2 + (-0.5f);
    const float cr = 1.0f + r2 * pc;
    switch (int(k) & 3) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}
inline float o_sin(float x) { float s, c; o_sincos(x, s, c); return s; }
inline float o_cos(float x) { float s, c; o_sincos(x, s, c); return c; }

inline float o_atan_unit(float a) {
    const bool big = a > 0.2679491924311227f;
    if (big) a = (a * 1.7320508075688772f - 1.0f) / (a + 1.7320508075688772f);
    const float z = a * a;
    float p = 1.0f / 13.0f;
    p = p * z + (-1.0f / 11.0f);
    p = p * z + (1.0f / 9.0f);
    p = p * z + (-1.0f / 7.0f);
    p = p * z + (1.0f / 5.0f);
    p = p * z + (-1.0f / 3.0f);
    float r = a + (a * z) * p;
    if (big) r = r + 0.52359877559829887308f;
    return r;
}
inline float o_atan(float x) {
    const float a = std::fabs(x);
    const float r = a > 1.0f ? 1.57079632679489661923f - o_atan_unit(1.0f / a) : o_atan_unit(a);
    return std::copysign(r, x);
}
inline float o_atan2(float y, float x) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) r = 0.0f;
    else if (ay <= ax) r = o_atan_unit(ay / ax);
    else r = 1.57079632679489661923f - o_atan_unit(ax / ay);
    if (std::signbit(x)) r = 3.14159265358979323846f - r;
    if (x != x || y != y) r = x + y;
    return std::copysign(r, y);
}
inline float o_acos(float x) {
    if (!(std::fabs(x) <= 1.0f)) return NAN;
    return o_atan2(std::sqrt((1.0f - x) * (1.0f + x)), x);
}
inline float o_log2(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
    if (x == INFINITY) return x;
    int e = 0;
    uint32_t b;
    std::memcpy(&b, &x, 4);
    if (b < 0x00800000u) {
        x = x * 16777216.0f;
        std::memcpy(&b, &x, 4);
        e = -24;
    }
    e += int(b >> 23) - 127;
    const uint32_t mb = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    std::memcpy(&m, &mb, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e += 1;
    }
    const float t = (m - 1.0f) / (m + 1.0f);
    const float t2 = t * t;
    float p = 1.0f / 11.0f;
    p = p * t2 + (1.0f / 9.0f);
    p = p * t2 + (1.0f / 7.0f);
    p = p * t2 + (1.0f / 5.0f);
    p = p * t2 + (1.0f / 3.0f);
    const float l = t + (t * t2) * p;
    return float(e) + (2.0f * l) * 1.4426950408889634f;
}
inline float o_pow2i(int k) { // 2^k for k in [-126, 127]
    const uint32_t b = uint32_t(k + 127) << 23;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline float o_ldexp(float x, int k) {
    if (k > 127) {
        x = x * o_pow2i(127);
        k -= 127;
    }
    if (k < -126) {
        x = x * o_pow2i(k + 64);
        return x * o_pow2i(-64);
    }
    return x * o_pow2i(k);
}
inline float o_exp2(float z) {
    if (z != z) return z;
    if (z >= 128.0f) return INFINITY;
    if (z <= -150.0f) return 0.0f;
    const float k = std::rint(z);
    const float f = z - k;
    const float r = f * 0.6931471805599453f;
    float p = 1.0f / 5040.0f;
    p = p * r + (1.0f / 720.0f);
    p = p * r + (1.0f / 120.0f);
    p = p * r + (1.0f / 24.0f);
    p = p * r + (1.0f / 6.0f);
    p = p * r + 0.5f;
    const float e = 1.0f + (r + (r * r) * p);
    return o_ldexp(e, int(k));
}
inline float o_exp(float x) {
    if (x != x) return x;
    if (x > 88.8f) return INFINITY;
    if (x < -104.0f) return 0.0f;
    const float k = std::rint(x * 1.4426950408889634f);
    const float r = (x - k * 0.693145751953125f) - k * 1.4286068203094173e-06f;
    float p = 1.0f / 5040.0f;
    p = p * r + (1.0f / 720.0f);
    p = p * r + (1.0f / 120.0f);
    p = p * r + (1.0f / 24.0f);
    p = p * r + (1.0f / 6.0f);
    p = p * r + 0.5f;
    const float e = 1.0f + (r + (r * r) * p);
    return o_ldexp(e, int(k));
}
inline float o_pow(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 0.0f : INFINITY;
    if (x < 0.0f) return NAN;
    if (x == INFINITY) return y > 0.0f ? x : 0.0f;
    return o_exp2(y * o_log2(x));
}
// pow(x, 5) of the Schlick term, as ((x x)(x x)) x
inline float o_pow5(float x) { return ((x * x) * (x * x)) * x; }

// IEEE binary16 <-> binary32, round to nearest even (SPIR-V OpFConvert default).
inline uint16_t f32_to_f16(float f) {
    const uint32_t x = asuint(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7FFFFFFFu;
    if (absx >= 0x7F800000u) { // inf / nan
        return uint16_t(sign | 0x7C00u | (absx > 0x7F800000u ? 0x200u : 0u));
    }
    if (absx >= 0x477FF000u) { // rounds to >= 65520 -> inf
        return uint16_t(sign | 0x7C00u);
    }
    if (absx < 0x38800000u) { // subnormal half (or zero)
        if (absx < 0x33000000u) return uint16_t(sign); // < 2^-25 rounds to 0
        const uint32_t exp = absx >> 23;
        const uint32_t mant = (absx & 0x7FFFFFu) | 0x800000u;
        const uint32_t shift = 126u - exp; // 14 + (113 - exp) ... value = mant * 2^(exp-150)
        // half subnormal unit = 2^-24; value / 2^-24 = mant * 2^(exp-126)
        const uint32_t sh = shift; // right shift amount, >= 14
        uint32_t q = mant >> sh;
        const uint32_t rem = mant & ((1u << sh) - 1u);
        const uint32_t half = 1u << (sh - 1u);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return uint16_t(sign | q);
    }
    // normal
    uint32_t h = ((absx >> 13) - (112u << 10));
    const uint32_t rem = absx & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return uint16_t(sign | h);
}

inline float f16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t exp = (h >> 10) & 0x1Fu;
    const uint32_t mant = h & 0x3FFu;
    if (exp == 0) {
        // subnormal: mant * 2^-24 (exact in float)
        const float v = float(mant) * 5.9604644775390625e-08f;
        return sign ? -v : v;
    }
    if (exp == 31) return asfloat(sign | 0x7F800000u | (mant << 13));
    return asfloat(sign | ((exp + 112u) << 23) | (mant << 13));
}

} // namespace orc
