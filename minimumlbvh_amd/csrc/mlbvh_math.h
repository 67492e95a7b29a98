// Self-written sine, cosine and power for the shading kernels.
//
// The reference shades with the device cosf / sinf (main_gpu.cu:30-31) and
// powf (main_gpu.cu:174, :283).  Library implementations differ between the
// GPU and a host libm by an ulp here and there, which flips an AO ray or a
// pixel's 8-bit level now and then.  These functions use only IEEE float
// add, subtract, multiply, divide, round-to-nearest-even and bit moves, in a
// fixed order, so the same code compiled for the host (the CPU oracle
// restates it operation for operation) gives the same bits.  Build with
// -ffp-contract=off: a fused multiply-add would change the rounding.
//
// Methods:
//  * sin / cos: Cody-Waite argument reduction (W. J. Cody and W. Waite,
//    Software Manual for the Elementary Functions, 1980): k = rint(x * 2/pi),
//    r = x - k*P1 - k*P2 - k*P3 with pi/2 = P1 + P2 + P3, P1 and P2 holding
//    13 significant bits so k*P1 and k*P2 are exact for |k| < 2^11; then the
//    Taylor polynomials of sin (to r^9) and cos (to r^10) on |r| <= pi/4 in
//    Horner form and the quadrant's sign/swap.  Truncation error below
//    2e-9.  Measured against float64 sin/cos: at most 2.3 ulp on [0, 2pi).
//    Domain: |x| <= 3000 (|k| < 2^11).
//  * pow(x, y) for x >= 0, y > 0: 2^(y * log2 x).  log2 x = e + ln(m)/ln 2
//    with x = 2^e * m, m in [sqrt(1/2), sqrt 2], ln m = 2 atanh(s),
//    s = (m-1)/(m+1), by its odd series to s^9 (|s| <= 0.172, truncation
//    below 1e-9).  The exponent y*e is split exactly (y cut to 12
//    significant bits times the integer e, plus the remainder) so that its
//    integer part carries no rounding error; 2^f for |f| <= 1/2 by its
//    Taylor series in f to f^7 (coefficients (ln 2)^n / n!, truncation below
//    1e-8), scaled by 2^k in two exact steps.  pow(0, y) = 0,
//    pow(inf, y) = inf, NaN stays NaN.  Measured against float64 pow at
//    y = 1/2.2 on [2^-24, 128]: at most 1.6 ulp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlbvh {
namespace fm {

// pi/2 = kP1 + kP2 + kP3 (kP1, kP2: 13 significant bits)
constexpr float kTwoOverPi = 0x1.45f306p-1f;
constexpr float kP1 = 0x1.921p+0f;
constexpr float kP2 = 0x1.f6ap-13f;
constexpr float kP3 = 0x1.110b46p-26f;
// 1/n!, alternating signs
constexpr float kS3 = -0x1.555556p-3f, kS5 = 0x1.111112p-7f, kS7 = -0x1.a01a02p-13f, kS9 = 0x1.71de3ap-19f;
constexpr float kC2 = -0x1p-1f, kC4 = 0x1.555556p-5f, kC6 = -0x1.6c16c2p-10f, kC8 = 0x1.a01a02p-16f,
                kC10 = -0x1.27e4fcp-22f;
// 2/(2j+1) for ln m = s * (2 + s^2 (2/3 + s^2 (2/5 + ...)))
constexpr float kL3 = 0x1.555556p-1f, kL5 = 0x1.99999ap-2f, kL7 = 0x1.24924ap-2f, kL9 = 0x1.c71c72p-3f;
constexpr float kInvLn2 = 0x1.715476p+0f;
constexpr float kSqrt2 = 0x1.6a09e6p+0f;
// (ln 2)^n / n!, n = 1..7
constexpr float kE1 = 0x1.62e430p-1f, kE2 = 0x1.ebfbe0p-3f, kE3 = 0x1.c6b08ep-5f, kE4 = 0x1.3b2ab6p-7f,
                kE5 = 0x1.5d87fep-10f, kE6 = 0x1.430912p-13f, kE7 = 0x1.ffcbfcp-17f;

__device__ inline void sincos(float x, float* s, float* c)
{
    const float kf = __builtin_rintf(x * kTwoOverPi);
    const int k = (int)kf;
    const float r = ((x - kf * kP1) - kf * kP2) - kf * kP3;
    const float r2 = r * r;
    const float sp = r + (r * r2) * (kS3 + r2 * (kS5 + r2 * (kS7 + r2 * kS9)));
    const float cp = 1.0f + r2 * (kC2 + r2 * (kC4 + r2 * (kC6 + r2 * (kC8 + r2 * kC10))));
    switch (k & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

// log2 x = e + L for a positive finite x: returns L (|L| <= 1/2), sets e
__device__ inline float log2_parts(float x, int* eOut)
{
    uint32_t b = __float_as_uint(x);
    int e = 0;
    if (b < 0x00800000u) {  // subnormal: scale by 2^24 (exact)
        x = x * 0x1p24f;
        b = __float_as_uint(x);
        e = -24;
    }
    e += (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > kSqrt2) {
        m = m * 0.5f;
        e += 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    const float lnm = s * (2.0f + s2 * (kL3 + s2 * (kL5 + s2 * (kL7 + s2 * kL9))));
    *eOut = e;
    return lnm * kInvLn2;
}

// 2^(k + f) for an integer-valued kf and |f| <= 1/2
__device__ inline float exp2_parts(float kf, float f)
{
    if (kf > 128.0f) return __uint_as_float(0x7F800000u);
    if (kf < -151.0f) return 0.0f;
    const float p = 1.0f + f * (kE1 + f * (kE2 + f * (kE3 + f * (kE4 + f * (kE5 + f * (kE6 + f * kE7))))));
    const int k = (int)kf;
    const int k1 = k >> 1, k2 = k - k1;  // each in [-76, 64]: a normal power of two
    const float s1 = __uint_as_float((uint32_t)(k1 + 127) << 23);
    const float s2 = __uint_as_float((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

// x^y for y > 0.  y * log2 x = yh*e + (yl*e + y*L), with yh = y cut to 12
// significant bits so that yh*e is exact (|e| <= 152): the integer part of
// the exponent carries no rounding error.
__device__ inline float pow_pos(float x, float y)
{
    if (x != x) return x;
    if (!(x > 0.0f)) return 0.0f;
    if (x == __uint_as_float(0x7F800000u)) return x;
    int e;
    const float L = log2_parts(x, &e);
    const float yh = __uint_as_float(__float_as_uint(y) & 0xFFFFF000u);
    const float yl = y - yh;  // exact
    const float fe = (float)e;
    const float zh = yh * fe;  // exact
    float kf = __builtin_rintf(zh);
    float f = (zh - kf) + (yl * fe + y * L);
    const float k2 = __builtin_rintf(f);
    f = f - k2;
    kf = kf + k2;
    return exp2_parts(kf, f);
}

}  // namespace fm
}  // namespace mlbvh
