// Deterministic scalar math shared by the HIP kernels and the host side.
//
// The renderer's results must be bit-identical between the gfx950 kernels and
// the scalar host restatement used to check them.  IEEE add/sub/mul/div/sqrt
// are correctly rounded on both sides (the build uses -ffp-contract=off and
// correctly rounded fp32 divide/sqrt), but elementary functions are not: the
// device library and glibc differ in the last ulp.  Every transcendental the
// render path needs is therefore evaluated here in FP64 with only +,-,*,/ and
// bit manipulation, then rounded once to FP32.  The polynomials are plain
// Taylor series with Cody-Waite argument reduction; accuracy is within 1 ulp
// of the correctly rounded FP32 result over the ranges the renderer uses
// (checked against libm by the test suite).
#pragma once

#include <cstdint>
#include <cmath>
#include <cstring>

// WPT_HD: inlined everywhere.  WPT_CALL: a real call on the device, for a
// large body reached from many call sites of one kernel (each inlined copy
// would add its full length to the kernel's code).
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WPT_HD __host__ __device__ __forceinline__
#define WPT_CALL __host__ __device__ inline __attribute__((noinline))
#else
#define WPT_HD inline
#define WPT_CALL inline
#endif

namespace wpt {

WPT_HD uint64_t bitsOf(double x) {
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof b);
    return b;
}
WPT_HD double fromBits(uint64_t b) {
    double x;
    __builtin_memcpy(&x, &b, sizeof x);
    return x;
}
WPT_HD uint32_t bitsOf(float x) {
    uint32_t b;
    __builtin_memcpy(&b, &x, sizeof b);
    return b;
}
WPT_HD float fromBits32(uint32_t b) {
    float x;
    __builtin_memcpy(&x, &b, sizeof x);
    return x;
}

WPT_HD bool isFinite(float x) { return (bitsOf(x) & 0x7f800000u) != 0x7f800000u; }
WPT_HD bool isNan(float x) { return (bitsOf(x) & 0x7fffffffu) > 0x7f800000u; }
WPT_HD float absf(float x) { return fromBits32(bitsOf(x) & 0x7fffffffu); }


// nextafterf(x, +/-inf) by integer stepping of the representation.
WPT_HD float nextUp(float x) {
    if (isNan(x) || x == __builtin_inff()) return x;
    if (x == 0.0f) return fromBits32(1u);
    uint32_t b = bitsOf(x);
    b = (b & 0x80000000u) ? b - 1u : b + 1u;
    return fromBits32(b);
}
WPT_HD float nextDown(float x) {
    if (isNan(x) || x == -__builtin_inff()) return x;
    if (x == 0.0f) return fromBits32(0x80000001u);
    uint32_t b = bitsOf(x);
    b = (b & 0x80000000u) ? b + 1u : b - 1u;
    return fromBits32(b);
}

// floor for doubles of moderate magnitude (|x| < 2^52), exact.
WPT_HD double floorD(double x) {
    const double two52 = 4503599627370496.0;
    double ax = x < 0.0 ? -x : x;
    if (!(ax < two52)) return x;
    double t = (ax + two52) - two52;  // round-to-nearest integer of |x|
    double r = x < 0.0 ? -t : t;
    if (r > x) r -= 1.0;
    return r;
}

namespace detail {
constexpr double kLn2Hi = 0.6931471806019545;        // ln 2 rounded to 32 significant bits
constexpr double kLn2Lo = -4.2009150726810846e-11;   // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;
constexpr double kSqrt2 = 1.4142135623730951;
constexpr double kPio2Hi = 1.5707963267341256;       // pi/2 in three 33-bit pieces
constexpr double kPio2Mid = 6.077100506303966e-11;
constexpr double kPio2Lo = 2.0222662487959506e-21;
constexpr double kTwoOverPi = 0.6366197723675814;

// e^r for |r| <= ln2/2, Taylor series to degree 13 (Horner).
WPT_HD double expReduced(double r) {
    double p = 1.0 / 6227020800.0;           // 1/13!
    p = p * r + 1.0 / 479001600.0;            // 1/12!
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    return p;
}

WPT_HD double sinReduced(double r) {  // |r| <= pi/4
    double r2 = r * r;
    double p = -1.0 / 121645100408832000.0;  // -1/19!
    p = p * r2 + 1.0 / 355687428096000.0;     // 1/17!
    p = p * r2 - 1.0 / 1307674368000.0;       // -1/15!
    p = p * r2 + 1.0 / 6227020800.0;          // 1/13!
    p = p * r2 - 1.0 / 39916800.0;            // -1/11!
    p = p * r2 + 1.0 / 362880.0;              // 1/9!
    p = p * r2 - 1.0 / 5040.0;                // -1/7!
    p = p * r2 + 1.0 / 120.0;                 // 1/5!
    p = p * r2 - 1.0 / 6.0;                   // -1/3!
    return r + r * (r2 * p);
}

WPT_HD double cosReduced(double r) {  // |r| <= pi/4
    double r2 = r * r;
    double p = 1.0 / 6402373705728000.0;      // 1/18!
    p = p * r2 - 1.0 / 20922789888000.0;      // -1/16!
    p = p * r2 + 1.0 / 87178291200.0;         // 1/14!
    p = p * r2 - 1.0 / 479001600.0;           // -1/12!
    p = p * r2 + 1.0 / 3628800.0;             // 1/10!
    p = p * r2 - 1.0 / 40320.0;               // -1/8!
    p = p * r2 + 1.0 / 720.0;                 // 1/6!
    p = p * r2 - 1.0 / 24.0;                  // -1/4!
    p = p * r2 + 0.5;                         // 1/2!
    return 1.0 - r2 * p;
}

// Quadrant-reduced sin/cos: returns both for x (double), Cody-Waite reduction
// by pi/2 in three pieces; exact k*piece products for |x| < 2^20.
WPT_HD void sinCosD(double x, double& s, double& c) {
    double k = floorD(x * kTwoOverPi + 0.5);
    double r = ((x - k * kPio2Hi) - k * kPio2Mid) - k * kPio2Lo;
    double sr = sinReduced(r);
    double cr = cosReduced(r);
    int64_t q = static_cast<int64_t>(k);
    int quadrant = static_cast<int>(q & 3);
    switch (quadrant) {
    case 0: s = sr; c = cr; break;
    case 1: s = cr; c = -sr; break;
    case 2: s = -sr; c = -cr; break;
    default: s = -cr; c = sr; break;
    }
}
}  // namespace detail

// e^x in double for finite x; overflow to +inf above 709.7, underflow to 0.
WPT_HD double expD(double x) {
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -745.0) return 0.0;
    double k = floorD(x * detail::kInvLn2 + 0.5);
    double r = (x - k * detail::kLn2Hi) - k * detail::kLn2Lo;
    double p = detail::expReduced(r);
    int64_t ki = static_cast<int64_t>(k);
    // Scale by 2^k in two steps so that subnormal results stay exact enough.
    int64_t k1 = ki / 2;
    int64_t k2 = ki - k1;
    double s1 = fromBits(static_cast<uint64_t>(k1 + 1023) << 52);
    double s2 = fromBits(static_cast<uint64_t>(k2 + 1023) << 52);
    return (p * s1) * s2;
}

// Natural log in double for x > 0 (normal doubles; every float is one).
WPT_HD double logD(double x) {
    if (x != x) return x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t b = bitsOf(x);
    int64_t e = static_cast<int64_t>((b >> 52) & 0x7ffu) - 1023;
    double m = fromBits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1,2)
    if (m > detail::kSqrt2) {
        m = m * 0.5;
        e += 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 19.0;
    p = p * s2 + 1.0 / 17.0;
    p = p * s2 + 1.0 / 15.0;
    p = p * s2 + 1.0 / 13.0;
    p = p * s2 + 1.0 / 11.0;
    p = p * s2 + 1.0 / 9.0;
    p = p * s2 + 1.0 / 7.0;
    p = p * s2 + 1.0 / 5.0;
    p = p * s2 + 1.0 / 3.0;
    p = p * s2 + 1.0;
    double logm = 2.0 * s * p;
    return static_cast<double>(e) * detail::kLn2 + logm;
}

WPT_HD float expF(float x) { return static_cast<float>(expD(static_cast<double>(x))); }
WPT_HD float logF(float x) { return static_cast<float>(logD(static_cast<double>(x))); }

// x^y for x >= 0 (the renderer never raises negative bases).
WPT_HD float powF(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (isNan(x) || isNan(y)) return x + y;
    if (x == 0.0f) return y > 0.0f ? 0.0f : __builtin_inff();
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 1.0f) return 1.0f;
    return static_cast<float>(expD(static_cast<double>(y) * logD(static_cast<double>(x))));
}

WPT_HD float sinF(float x) {
    if (!isFinite(x)) return __builtin_nanf("");
    double s, c;
    detail::sinCosD(static_cast<double>(x), s, c);
    return static_cast<float>(s);
}
WPT_HD float cosF(float x) {
    if (!isFinite(x)) return __builtin_nanf("");
    double s, c;
    detail::sinCosD(static_cast<double>(x), s, c);
    return static_cast<float>(c);
}
WPT_HD void sinCosF(float x, float& s, float& c) {
    if (!isFinite(x)) {
        s = c = __builtin_nanf("");
        return;
    }
    double sd, cd;
    detail::sinCosD(static_cast<double>(x), sd, cd);
    s = static_cast<float>(sd);
    c = static_cast<float>(cd);
}
WPT_HD float tanF(float x) {
    if (!isFinite(x)) return __builtin_nanf("");
    double s, c;
    detail::sinCosD(static_cast<double>(x), s, c);
    return static_cast<float>(s / c);
}

// sin / cos in double for |x| < 2^20 (environment sampling math).
WPT_HD double sinD(double x) {
    double s, c;
    detail::sinCosD(x, s, c);
    return s;
}
WPT_HD double cosD(double x) {
    double s, c;
    detail::sinCosD(x, s, c);
    return c;
}

namespace detail {
constexpr double kPi = 3.141592653589793;
constexpr double kPio2 = 1.5707963267948966;
constexpr double kPio6 = 0.5235987755982989;
constexpr double kSqrt3 = 1.7320508075688772;
constexpr double kTanPi12 = 0.2679491924311228;  // tan(pi/12)

// atan(t) for 0 <= t <= tan(pi/12): odd Taylor series through t^31.
WPT_HD double atanReduced(double t) {
    double t2 = t * t;
    double p = 1.0 / 31.0;
    p = 1.0 / 29.0 - t2 * p;
    p = 1.0 / 27.0 - t2 * p;
    p = 1.0 / 25.0 - t2 * p;
    p = 1.0 / 23.0 - t2 * p;
    p = 1.0 / 21.0 - t2 * p;
    p = 1.0 / 19.0 - t2 * p;
    p = 1.0 / 17.0 - t2 * p;
    p = 1.0 / 15.0 - t2 * p;
    p = 1.0 / 13.0 - t2 * p;
    p = 1.0 / 11.0 - t2 * p;
    p = 1.0 / 9.0 - t2 * p;
    p = 1.0 / 7.0 - t2 * p;
    p = 1.0 / 5.0 - t2 * p;
    p = 1.0 / 3.0 - t2 * p;
    return t - t * (t2 * p);
}

// atan(a) for a >= 0: a > 1 reflects through pi/2 - atan(1/a), and
// a > tan(pi/12) shifts by pi/6 (atan a = pi/6 + atan((a sqrt3 - 1)/(a + sqrt3))).
WPT_HD double atanPositive(double a) {
    bool invert = a > 1.0;
    if (invert) a = 1.0 / a;
    double r;
    if (a > kTanPi12)
        r = kPio6 + atanReduced((a * kSqrt3 - 1.0) / (a + kSqrt3));
    else
        r = atanReduced(a);
    return invert ? kPio2 - r : r;
}
}  // namespace detail

// atan2 in double with IEEE signed-zero and infinity-free quadrant rules.
WPT_HD double atan2D(double y, double x) {
    if (y != y || x != x) return y + x;
    bool yNeg = bitsOf(y) >> 63;
    bool xNeg = bitsOf(x) >> 63;
    double ay = yNeg ? -y : y, ax = xNeg ? -x : x;
    double r;
    if (ay == 0.0) {
        r = xNeg ? detail::kPi : 0.0;
    } else if (ax == 0.0) {
        r = detail::kPio2;
    } else {
        r = detail::atanPositive(ay / ax);
        if (xNeg) r = detail::kPi - r;
    }
    return yNeg ? -r : r;
}
WPT_HD float atan2F(float y, float x) {
    return static_cast<float>(atan2D(static_cast<double>(y), static_cast<double>(x)));
}

}  // namespace wpt
