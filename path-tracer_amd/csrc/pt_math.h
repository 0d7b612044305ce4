// pt_math.h - deterministic FP32 math shared by the HIP kernels and the CPU side.
//
// Every function here is built only from IEEE-754 operations that are
// correctly rounded on both gfx950 and x86-64 (add, mul, fma, div, sqrt,
// floor/rint, ldexp/frexp), evaluated in a fixed order. Compiled with
// -ffp-contract=off on both sides, a given input produces the same bits on
// the GPU and on the host, so per-pixel radiance can be compared exactly.
//
// The trace kernel of the reference (src/render/trace.slang) calls the
// platform's sin/cos/acos/atan2/pow, whose last-ulp behaviour is left open by
// WGSL; these are our own implementations of those functions:
//   sin/cos   Cody-Waite reduction by pi/2 (three-part constant) + Taylor
//             polynomials on [-pi/4, pi/4] (truncation error < 1e-11).
//   atan      reduction to |u| <= tan(pi/8) via atan(t) = pi/4 + atan((t-1)/(t+1)),
//             then u + u^3 P(u^2), P fitted by gen_detmath_coeffs.py.
//   asin/acos half-angle reduction acos(x) = 2 asin(sqrt((1-x)/2)) for |x| > 1/2,
//             then z + z^3 Q(z^2), Q fitted by gen_detmath_coeffs.py.
//   log2/exp2 frexp + atanh series on [sqrt(1/2), sqrt(2)); exp2 by rint split
//             and the Taylor series of e^(f ln 2); pow(x, y) = exp2(y log2 x),
//             which is the accuracy WGSL itself specifies for pow.
// Accuracy against glibc is a few ulp (tested on the host).
//
// The vector helpers follow the WGSL/Slang definitions used by the reference:
// dot = (x*x' + y*y') + z*z', normalize = v * (1 / sqrt(dot(v, v))),
// lerp(a, b, t) = a * (1 - t) + b * t (WGSL mix), rsqrt = 1 / sqrt.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PT_HD __host__ __device__ __forceinline__
#else
#define PT_HD inline
#endif

namespace ptm {

constexpr float PI = 3.14159265359f;             // trace.slang:3
constexpr float INV_PI = 0.31830988618f;         // trace.slang:4
constexpr float INV_2PI = 0.15915494309189535f;  // trace.slang:5
constexpr float FLT_MAX_ = 3.40282347e+38f;      // 0x1.fffffep+127

PT_HD float bits_to_float(uint32_t u) {
    union { uint32_t u; float f; } c;
    c.u = u;
    return c.f;
}
PT_HD uint32_t float_to_bits(float f) {
    union { uint32_t u; float f; } c;
    c.f = f;
    return c.u;
}

PT_HD float fma_(float a, float b, float c) { return fmaf(a, b, c); }
PT_HD float sqrt_(float x) { return sqrtf(x); }
PT_HD float rsqrt_(float x) { return 1.0f / sqrtf(x); }

// sqrtf for an argument that is +-0, at least 2^-96, +inf or NaN; each caller
// states why its argument is one of these. For such arguments the correctly
// rounded root is the hardware square root (within 1 ulp) followed by one test
// of each neighbouring float s': x - s'*s <= 0 takes the one below, > 0 the one
// above. These are the steps of the device's correctly rounded sqrtf without
// its rescaling of tiny arguments and its special-value selects, which change
// nothing for these arguments: the result has sqrtf's bits (+-0, +inf and NaN
// come through the tests unchanged; negative normal arguments give NaN, as
// sqrtf does). The host calls sqrtf.
PT_HD float sqrt_nt_(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = __builtin_amdgcn_sqrtf(x);
    const float below = bits_to_float(float_to_bits(s) - 1u);
    const float above = bits_to_float(float_to_bits(s) + 1u);
    const float r = fmaf(-below, s, x) <= 0.0f ? below : s;
    return fmaf(-above, s, x) > 0.0f ? above : r;
#else
    return sqrtf(x);
#endif
}

// a / b when the division needs no operand rescaling and no special-case
// fixup: b normal with 2^-60 <= |b| <= 2^60, and a +0 (not -0), NaN or
// 2^-60 <= |a| <= 2^4 (each caller states why). The correctly rounded quotient
// is then the hardware reciprocal refined once by Newton's step and the
// quotient corrected twice by its fma residual, the device division's own
// steps without the scaling and fixup instructions that are identities on this
// domain; the bits equal a / b. The host divides.
PT_HD float div_nt_(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    float r = __builtin_amdgcn_rcpf(b);
    r = fmaf(fmaf(-b, r, 1.0f), r, r);
    float q = a * r;
    q = fmaf(fmaf(-b, q, a), r, q);
    return fmaf(fmaf(-b, q, a), r, q);
#else
    return a / b;
#endif
}
// min/max with the C fmin/fmax NaN rule (return the non-NaN operand)
PT_HD float min_(float a, float b) { return fminf(a, b); }
PT_HD float max_(float a, float b) { return fmaxf(a, b); }
PT_HD float clamp_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
PT_HD float saturate_(float x) { return clamp_(x, 0.0f, 1.0f); }
PT_HD float frac_(float x) { return x - floorf(x); }
PT_HD float lerp_(float a, float b, float t) { return a * (1.0f - t) + b * t; }
PT_HD float sign_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }

// ---------------------------------------------------------------- sin / cos
constexpr float TWO_OVER_PI = 0.636619772367581343f;
// pi/2 split in three float parts (each the float nearest to what is left);
// with fma each step x - k*part is computed with a single rounding
constexpr float PIO2_HI = 1.5707963705062866f;       // 0x3fc90fdb
constexpr float PIO2_MID = -4.371138828673793e-08f;  // 0xb33bbd2e
constexpr float PIO2_LO = -1.7151245100058819e-15f;  // 0xa6f72ced

PT_HD float sin_poly(float r) {
    float t = r * r;
    float p = fma_(t, -2.50521083854417187751e-08f, 2.75573192239858906526e-06f);  // -1/11!, 1/9!
    p = fma_(t, p, -1.98412698412698412698e-04f);                                   // -1/7!
    p = fma_(t, p, 8.33333333333333333333e-03f);                                    // 1/5!
    p = fma_(t, p, -1.66666666666666666667e-01f);                                   // -1/3!
    return fma_(r * t, p, r);
}
PT_HD float cos_poly(float r) {
    float t = r * r;
    float p = fma_(t, 2.08767569878680989792e-09f, -2.75573192239858906526e-07f);  // 1/12!, -1/10!
    p = fma_(t, p, 2.48015873015873015873e-05f);                                    // 1/8!
    p = fma_(t, p, -1.38888888888888888889e-03f);                                   // -1/6!
    p = fma_(t, p, 4.16666666666666666667e-02f);                                    // 1/4!
    p = fma_(t, p, -0.5f);
    return fma_(t, p, 1.0f);
}
// reduce x to r in [-pi/4, pi/4] with quadrant q; valid for |x| < 1e5
PT_HD float reduce_pio2(float x, int& q) {
    float k = rintf(x * TWO_OVER_PI);
    float r = fma_(-k, PIO2_HI, x);
    r = fma_(-k, PIO2_MID, r);
    r = fma_(-k, PIO2_LO, r);
    q = (int)k;
    return r;
}
PT_HD float sin_(float x) {
    int q;
    float r = reduce_pio2(x, q);
    float s = (q & 1) ? cos_poly(r) : sin_poly(r);
    return (q & 2) ? -s : s;
}
PT_HD float cos_(float x) {
    int q;
    float r = reduce_pio2(x, q);
    float c = (q & 1) ? sin_poly(r) : cos_poly(r);
    return ((q + 1) & 2) ? -c : c;
}
// sin_(x) and cos_(x) bit for bit from one reduction: both polynomials are
// evaluated and the quadrant only selects between them, so the GPU build has
// no lane-divergent branch here
PT_HD void sincos_(float x, float& s, float& c) {
    int q;
    const float r = reduce_pio2(x, q);
    const float sp = sin_poly(r), cp = cos_poly(r);
    const float s0 = (q & 1) ? cp : sp;
    const float c0 = (q & 1) ? sp : cp;
    s = (q & 2) ? -s0 : s0;
    c = ((q + 1) & 2) ? -c0 : c0;
}

// ---------------------------------------------------------------- atan / atan2
PT_HD float atan_poly(float u) {  // |u| <= tan(pi/8)
    float t = u * u;
    float p = fma_(t, 0.030966660007834435f, -0.058961812406778336f);
    p = fma_(t, p, 0.0757002905011177f);
    p = fma_(t, p, -0.09079363942146301f);
    p = fma_(t, p, 0.11110468953847885f);
    p = fma_(t, p, -0.14285694062709808f);
    p = fma_(t, p, 0.20000000298023224f);
    p = fma_(t, p, -0.3333333432674408f);
    return fma_(u * t, p, u);
}
constexpr float PIO4 = 0.785398163397448309616f;
constexpr float PIO2 = 1.57079632679489661923f;
constexpr float PI_F = 3.14159265358979323846f;
constexpr float TAN_PIO8 = 0.414213562373095048802f;
// atan of t in [0, 1]
PT_HD float atan01(float t) {
    if (t > TAN_PIO8) {
        return PIO4 + atan_poly((t - 1.0f) / (t + 1.0f));
    }
    return atan_poly(t);
}
PT_HD float atan2_(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float a;
    if (ay == 0.0f && ax == 0.0f) {
        a = 0.0f;
    } else if (ay > ax) {
        a = PIO2 - atan01(ax / ay);
    } else {
        a = atan01(ay / ax);
    }
    if (signbit(x)) a = PI_F - a;
    return copysignf(a, y);
}

// ---------------------------------------------------------------- asin / acos
PT_HD float asin_poly(float z) {  // 0 <= z <= 1/2
    float t = z * z;
    float p = fma_(t, 0.032535187900066376f, 0.0023562510032206774f);
    p = fma_(t, p, 0.020492924377322197f);
    p = fma_(t, p, 0.02189820632338524f);
    p = fma_(t, p, 0.030422721058130264f);
    p = fma_(t, p, 0.04464094713330269f);
    p = fma_(t, p, 0.07500004023313522f);
    p = fma_(t, p, 0.1666666716337204f);
    return fma_(z * t, p, z);
}
// 1 - |x| of a float x is 0 or at least 2^-24 in magnitude, so both roots'
// arguments (halved) are 0, NaN or at least 2^-25: sqrt_nt_'s domain
PT_HD float acos_(float x) {
    if (x > 0.5f) {
        return 2.0f * asin_poly(sqrt_nt_((1.0f - x) * 0.5f));
    }
    if (x < -0.5f) {
        return PI_F - 2.0f * asin_poly(sqrt_nt_((1.0f + x) * 0.5f));
    }
    float ax = fabsf(x);
    float s = asin_poly(ax);
    return x < 0.0f ? PIO2 + s : PIO2 - s;
}

// ---------------------------------------------------------------- log2 / exp2 / pow
constexpr float SQRT_HALF = 0.707106781186547524401f;
constexpr float TWO_OVER_LN2 = 2.88539008177792681472f;
PT_HD float log2_(float x) {  // x > 0, finite
    int e;
    float m = frexpf(x, &e);  // m in [0.5, 1)
    if (m < SQRT_HALF) {
        m = m * 2.0f;
        e -= 1;
    }
    float s = (m - 1.0f) / (m + 1.0f);
    float t = s * s;
    float p = fma_(t, 0.0909090909090909091f, 0.111111111111111111f);  // 1/11, 1/9
    p = fma_(t, p, 0.142857142857142857f);                              // 1/7
    p = fma_(t, p, 0.2f);                                               // 1/5
    p = fma_(t, p, 0.333333333333333333f);                              // 1/3
    float atanh_s = fma_(s * t, p, s);
    return fma_(atanh_s, TWO_OVER_LN2, (float)e);
}
constexpr float LN2 = 0.693147180559945309417f;
PT_HD float exp2_(float y) {  // |y| < 126
    float n = rintf(y);
    float f = (y - n) * LN2;  // |f| <= ln2 / 2
    float p = fma_(f, 1.98412698412698412698e-04f, 1.38888888888888888889e-03f);  // 1/7!, 1/6!
    p = fma_(f, p, 8.33333333333333333333e-03f);                                  // 1/5!
    p = fma_(f, p, 4.16666666666666666667e-02f);                                  // 1/4!
    p = fma_(f, p, 1.66666666666666666667e-01f);                                  // 1/3!
    p = fma_(f, p, 0.5f);
    p = fma_(f, p, 1.0f);
    p = fma_(f, p, 1.0f);
    return ldexpf(p, (int)n);
}
PT_HD float pow_(float x, float y) {  // x >= 0
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : INFINITY);
    return exp2_(y * log2_(x));
}

// ---------------------------------------------------------------- float3
struct f3 {
    float x, y, z;
};
PT_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
PT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
PT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
PT_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PT_HD f3 operator+(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
PT_HD f3 operator-(float s, f3 a) { return f3{s - a.x, s - a.y, s - a.z}; }
PT_HD f3 operator-(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
PT_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
PT_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PT_HD float length(f3 a) { return sqrtf(dot(a, a)); }
PT_HD f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
PT_HD f3 fma3(f3 a, f3 b, f3 c) { return f3{fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)}; }
PT_HD f3 min3(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PT_HD f3 max3(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PT_HD f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
PT_HD f3 lerp3(f3 a, f3 b, float t) { return a * (1.0f - t) + b * t; }
PT_HD f3 lerp3v(f3 a, f3 b, f3 t) {
    return f3{lerp_(a.x, b.x, t.x), lerp_(a.y, b.y, t.y), lerp_(a.z, b.z, t.z)};
}
PT_HD f3 saturate3(f3 a) { return f3{saturate_(a.x), saturate_(a.y), saturate_(a.z)}; }
PT_HD f3 reflect(f3 i, f3 n) { return i - n * (2.0f * dot(n, i)); }
PT_HD float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
PT_HD bool any_pos(f3 a) { return a.x > 0.0f || a.y > 0.0f || a.z > 0.0f; }

// Rec. 709 luma (trace.slang:7)
PT_HD float luma(f3 c) { return dot(c, f3{0.2126f, 0.7152f, 0.0722f}); }

// PCG-style hash step of the reference RNG (trace.slang:195-202)
PT_HD float rand_step(uint32_t& s) {
    s = s * 747796405u + 2891336453u;
    uint32_t word = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return bits_to_float(0x3f800000u | (((word >> 22u) ^ word) >> 9u)) - 1.0f;
}
// per (pixel, sample) seed (trace.slang:736-737); dispatch z is always 0
PT_HD uint32_t rng_seed(uint32_t px, uint32_t py, uint32_t sample_count) {
    uint32_t s = (px * 0x9E3779B9u) ^ (py * 0x85EBCA6Bu) ^ (0u * 0xC2B2AE35u) ^
                 (sample_count * 0x27D4EB2Du) ^ 0xA511E9B3u;
    return s | 1u;
}

}  // namespace ptm
