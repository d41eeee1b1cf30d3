// FP32 3-vector used by the host scene code and the kernels.
//
// The expression order of every operator is fixed (no FMA contraction is
// allowed by the build flags), because the kernels must produce the same bits
// as the host restatement of the reference's Vec3 algebra
// (include/core/vec3.cuh:6-99 of the reference): dot is ((x*x + y*y) + z*z),
// division by a scalar multiplies by the reciprocal, and normalize multiplies
// by 1/sqrt.
#pragma once

#include <cmath>

#include "wpt_math.h"

namespace wpt {

struct Vec2 {
    float x, y;
};

struct Vec3 {
    float x, y, z;

    WPT_HD Vec3() : x(0.0f), y(0.0f), z(0.0f) {}
    WPT_HD Vec3(float xx, float yy, float zz) : x(xx), y(yy), z(zz) {}

    WPT_HD Vec3 operator-() const { return Vec3(-x, -y, -z); }
    WPT_HD Vec3 operator+(const Vec3& v) const { return Vec3(x + v.x, y + v.y, z + v.z); }
    WPT_HD Vec3 operator-(const Vec3& v) const { return Vec3(x - v.x, y - v.y, z - v.z); }
    WPT_HD Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
    WPT_HD Vec3 operator/(float s) const {
        float inv = 1.0f / s;
        return *this * inv;
    }
    WPT_HD Vec3& operator+=(const Vec3& v) {
        x += v.x;
        y += v.y;
        z += v.z;
        return *this;
    }
    WPT_HD Vec3& operator*=(float s) {
        x *= s;
        y *= s;
        z *= s;
        return *this;
    }
};

WPT_HD Vec3 operator*(float s, const Vec3& v) { return v * s; }
WPT_HD float dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
WPT_HD Vec3 cross(const Vec3& a, const Vec3& b) {
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
WPT_HD float lengthSquared(const Vec3& v) { return dot(v, v); }
WPT_HD float length(const Vec3& v) { return sqrtf(lengthSquared(v)); }
WPT_HD Vec3 normalize(const Vec3& v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
WPT_HD Vec3 hadamard(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
WPT_HD float maxComponent(const Vec3& v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }

struct Ray {
    Vec3 origin;
    Vec3 direction;
    WPT_HD Ray() {}
    WPT_HD Ray(const Vec3& o, const Vec3& d) : origin(o), direction(d) {}
    WPT_HD Vec3 at(float t) const { return origin + direction * t; }
};

// Traversal cutoff; reference include/core/ray.cuh:8.
constexpr float kRayMinDistance = 1.0e-6f;
constexpr float kFltEpsilon = 1.1920928955078125e-7f;

// Precision-scaled origin push along the geometric normal followed by one ulp
// step per axis away from the surface; reference include/core/ray.cuh:10-21.
WPT_HD Vec3 offsetRayOrigin(const Vec3& position, const Vec3& geometricNormal, const Vec3& direction,
    float hitDistance) {
    float magnitude = fmaxf(fmaxf(1.0f, fabsf(hitDistance)),
        fmaxf(fabsf(position.x), fmaxf(fabsf(position.y), fabsf(position.z))));
    float sign = dot(direction, geometricNormal) >= 0.0f ? 1.0f : -1.0f;
    Vec3 offset = geometricNormal * (sign * 32.0f * 1.1920928955078125e-7f * magnitude);
    Vec3 result = position + offset;
    if (offset.x != 0.0f) result.x = offset.x > 0.0f ? nextUp(result.x) : nextDown(result.x);
    if (offset.y != 0.0f) result.y = offset.y > 0.0f ? nextUp(result.y) : nextDown(result.y);
    if (offset.z != 0.0f) result.z = offset.z > 0.0f ? nextUp(result.z) : nextDown(result.z);
    return result;
}

}  // namespace wpt
