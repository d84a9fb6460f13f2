// Elementary functions of the path tracer, built only from IEEE +, -, *, /, sqrt, rint and bit
// operations in a fixed order, so the device and a CPU restatement round identically.
//
// The reference's shaders call the Slang/SPIR-V built-ins (sin, cos, pow, atan2, acos, atan,
// exp), whose precision the Vulkan driver chooses; these replacements are accurate to a few ulp:
//   sin/cos : Cody & Waite (1980) reduction by pi/2 with a 3-part constant, Taylor polynomials
//             (degree 9 / 10) on [-pi/4, pi/4]
//   atan    : reduction to [0, tan(pi/12)] by 1/x and the pi/6 addition formula, odd Taylor
//             series to degree 13
//   log2    : exponent split, atanh series in t = (m - 1) / (m + 1), m in [sqrt(1/2), sqrt(2))
//   exp2    : k = rint(z), Taylor series of e^(f ln 2) to degree 7, scale by 2^k
//   pow5    : ((x x)(x x)) x, the Schlick term's pow(x, 5)
#pragma once

#include <stdint.h>

#ifndef MRT_HD
#define MRT_HD __host__ __device__ __forceinline__
#endif

namespace mrt {

MRT_HD float fbits(uint32_t u) { return __builtin_bit_cast(float, u); }
MRT_HD uint32_t ubits(float f) { return __builtin_bit_cast(uint32_t, f); }

constexpr float PIO2_1 = 1.5703125f;                // pi/2, first 12 bits
constexpr float PIO2_2 = 4.837512969970703125e-4f;  // next 12 bits
constexpr float PIO2_3 = 7.549790126404332e-08f;    // remainder
constexpr float TWO_OVER_PI = 0.63661975f;
constexpr float M_PI_F = 3.14159265358979323846f;
constexpr float M_PI_2_F = 1.57079632679489661923f;
constexpr float M_PI_6_F = 0.52359877559829887308f;
constexpr float SQRT3_F = 1.7320508075688772f;
constexpr float TAN_PI_12_F = 0.2679491924311227f;
constexpr float LOG2E_F = 1.4426950408889634f;
constexpr float LN2_F = 0.6931471805599453f;

// sin and cos of x, |x| up to ~1e4 (the path tracer's arguments lie in [-2 pi, 2 pi]).
MRT_HD void msincos(float x, float& s, float& c) {
    const float k = rintf(x * TWO_OVER_PI);
    const float r = ((x - k * PIO2_1) - k * PIO2_2) - k * PIO2_3;
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
    const int q = int(k) & 3;
    s = q == 0 ? sr : q == 1 ? cr : q == 2 ? -sr : -cr;
    c = q == 0 ? cr : q == 1 ? -sr : q == 2 ? -cr : sr;
}
MRT_HD float msin(float x) {
    float s, c;
    msincos(x, s, c);
    return s;
}
MRT_HD float mcos(float x) {
    float s, c;
    msincos(x, s, c);
    return c;
}

// atan of a in [0, 1]
MRT_HD float matan_unit(float a) {
    const bool big = a > TAN_PI_12_F;
    if (big) a = (a * SQRT3_F - 1.0f) / (a + SQRT3_F);
    const float z = a * a;
    float p = 1.0f / 13.0f;
    p = p * z + (-1.0f / 11.0f);
    p = p * z + (1.0f / 9.0f);
    p = p * z + (-1.0f / 7.0f);
    p = p * z + (1.0f / 5.0f);
    p = p * z + (-1.0f / 3.0f);
    float r = a + (a * z) * p;
    if (big) r = r + M_PI_6_F;
    return r;
}
MRT_HD float matan(float x) {
    const float a = fabsf(x);
    const float r = a > 1.0f ? M_PI_2_F - matan_unit(1.0f / a) : matan_unit(a);
    return copysignf(r, x);
}
MRT_HD float matan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    float r;
    if (ax == 0.0f && ay == 0.0f) {
        r = 0.0f;
    } else if (ay <= ax) {
        r = matan_unit(ay / ax);
    } else {
        r = M_PI_2_F - matan_unit(ax / ay);
    }
    if (signbit(x)) r = M_PI_F - r;
    if (x != x || y != y) r = x + y;
    return copysignf(r, y);
}
MRT_HD float macos(float x) {
    if (!(fabsf(x) <= 1.0f)) return __builtin_nanf("");
    return matan2(sqrtf((1.0f - x) * (1.0f + x)), x);
}

MRT_HD float mlog2(float x) {
    if (!(x > 0.0f)) return x == 0.0f ? -__builtin_inff() : __builtin_nanf("");
    if (x == __builtin_inff()) return x;
    int e = 0;
    uint32_t b = ubits(x);
    if (b < 0x00800000u) {
        x = x * 16777216.0f;
        b = ubits(x);
        e = -24;
    }
    e += int(b >> 23) - 127;
    float m = fbits((b & 0x007FFFFFu) | 0x3F800000u);
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
    return float(e) + (2.0f * l) * LOG2E_F;
}
MRT_HD float mldexp(float x, int k) {
    if (k > 127) {
        x = x * fbits(0x7F000000u); // 2^127
        k -= 127;
    }
    if (k < -126) {
        x = x * fbits(uint32_t(k + 64 + 127) << 23);
        return x * fbits(0x1F800000u); // 2^-64
    }
    return x * fbits(uint32_t(k + 127) << 23);
}
MRT_HD float mexp2(float z) {
    if (z != z) return z;
    if (z >= 128.0f) return __builtin_inff();
    if (z <= -150.0f) return 0.0f;
    const float k = rintf(z);
    const float f = z - k;
    const float r = f * LN2_F;
    float p = 1.0f / 5040.0f;
    p = p * r + (1.0f / 720.0f);
    p = p * r + (1.0f / 120.0f);
    p = p * r + (1.0f / 24.0f);
    p = p * r + (1.0f / 6.0f);
    p = p * r + 0.5f;
    const float e = 1.0f + (r + (r * r) * p);
    return mldexp(e, int(k));
}
// e^x = 2^k e^r, r = x - k ln2 with ln2 split in two (the high part has 16 bits, so k * hi is exact)
MRT_HD float mexp(float x) {
    if (x != x) return x;
    if (x > 88.8f) return __builtin_inff();
    if (x < -104.0f) return 0.0f;
    const float k = rintf(x * LOG2E_F);
    const float r = (x - k * 0.693145751953125f) - k * 1.4286068203094173e-06f;
    float p = 1.0f / 5040.0f;
    p = p * r + (1.0f / 720.0f);
    p = p * r + (1.0f / 120.0f);
    p = p * r + (1.0f / 24.0f);
    p = p * r + (1.0f / 6.0f);
    p = p * r + 0.5f;
    const float e = 1.0f + (r + (r * r) * p);
    return mldexp(e, int(k));
}
MRT_HD float mpow(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    if (x < 0.0f) return __builtin_nanf("");
    if (x == __builtin_inff()) return y > 0.0f ? x : 0.0f;
    return mexp2(y * mlog2(x));
}
MRT_HD float pow5(float x) { return ((x * x) * (x * x)) * x; }

} // namespace mrt
