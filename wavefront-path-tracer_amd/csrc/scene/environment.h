// Equirectangular environment map: lookup convention, solid angles,
// importance sampling and radiance.
//
// Reference: include/scene/environment.cuh:10-71 (rotation, uv convention,
// texel solid angle, pdf, alias sampling with a double candidate), and
// device_traversal.cuh:571-575 (radiance lookup).  The reference reads the
// radiance with hardware bilinear filtering (normalized coordinates, wrap in
// u, clamp in v, 1/256 filter weights); here the same bilinear footprint is
// evaluated in FP32 with exact weights.
#pragma once

#include "device_scene.h"

namespace wpt {

constexpr float kEnvPi = 3.14159265359f;
constexpr double kEnvPiD = 3.14159265358979323846;

WPT_HD Vec3 rotateEnvironment(const Vec3& d, float angle) {
    float c = cosF(angle), s = sinF(angle);
    return Vec3(c * d.x + s * d.z, d.y, -s * d.x + c * d.z);
}

// +Y at the top, +X at the center column, +Z at three quarters.
WPT_HD Vec2 environmentUv(const Vec3& world, float rotation) {
    Vec3 d = rotateEnvironment(world, -rotation);
    float u = atan2F(d.z, d.x) / (2.0f * kEnvPi) + 0.5f;
    u -= floorf(u);
    float v = atan2F(sqrtf(d.x * d.x + d.z * d.z), d.y) / kEnvPi;
    return Vec2{u, fminf(1.0f, fmaxf(0.0f, v))};
}

// Solid angle of a texel of row `row`: (4 pi / w) sin(pi (row + 1/2) / h)
// sin(pi / (2 h)), i.e. cos(theta0) - cos(theta1) without cancellation.
WPT_HD float environmentSolidAngle(uint32_t row, uint32_t width, uint32_t height) {
    return static_cast<float>((4.0 * kEnvPiD / width) * sinD(kEnvPiD * (row + 0.5) / height) *
        sinD(kEnvPiD / (2.0 * height)));
}

WPT_HD uint32_t environmentTexel(const EnvironmentMap& m, const Vec3& direction) {
    Vec2 uv = environmentUv(direction, m.rotation);
    uint32_t x = static_cast<uint32_t>(uv.x * static_cast<float>(m.width));
    uint32_t y = static_cast<uint32_t>(uv.y * static_cast<float>(m.height));
    return (y < m.height ? y : m.height - 1) * m.width + (x < m.width ? x : m.width - 1);
}

WPT_HD float environmentPdf(const EnvironmentMap& m, const Vec3& direction) {
    if (m.width == 0 || m.height == 0 || !m.probabilities) return 0.0f;
    uint32_t index = environmentTexel(m, direction);
    return m.probabilities[index] / environmentSolidAngle(index / m.width, m.width, m.height);
}

struct EnvironmentSample {
    Vec3 direction;
    float pdf;
};

// Alias candidate and coin, then a uniform longitude and cos(latitude)
// within the texel.  The candidate is a double so maps beyond 2^24 texels
// keep index support.
WPT_HD EnvironmentSample sampleEnvironment(const EnvironmentMap& m, double candidate, float coin, float longitude,
    float latitude) {
    uint32_t count = m.width * m.height;
    uint32_t index = static_cast<uint32_t>(candidate * count);
    if (index >= count) index = count - 1;
    if (coin >= m.alias[index].probability) index = m.alias[index].alias;
    uint32_t x = index % m.width, y = index / m.width;
    // The longitude offset is formed in FP32 (texel column plus clamped
    // jitter, over the width), as the reference's mixed expression does.
    double u = (static_cast<float>(x) + fminf(1.0f - 1.0e-7f, fmaxf(1.0e-7f, longitude))) / static_cast<float>(m.width);
    double jitter = static_cast<double>(fminf(1.0f - 1.0e-7f, fmaxf(1.0e-7f, latitude)));
    double cosine = (1.0 - jitter) * cosD(kEnvPiD * y / m.height) + jitter * cosD(kEnvPiD * (y + 1.0) / m.height);
    double sine = sqrt(fmax(0.0, 1.0 - cosine * cosine));
    double phi = (u - 0.5) * 2 * kEnvPiD;
    Vec3 local(static_cast<float>(sine * cosD(phi)), static_cast<float>(cosine), static_cast<float>(sine * sinD(phi)));
    EnvironmentSample s;
    s.direction = rotateEnvironment(local, m.rotation);
    s.pdf = m.probabilities[index] / environmentSolidAngle(y, m.width, m.height);
    return s;
}

// Bilinear radiance: texel centers at ((i + 1/2) / w, (j + 1/2) / h), wrap
// in u, clamp in v.
WPT_HD Vec3 environmentRadiance(const EnvironmentMap& m, const Vec3& direction) {
    Vec2 uv = environmentUv(direction, m.rotation);
    float x = uv.x * static_cast<float>(m.width) - 0.5f;
    float y = uv.y * static_cast<float>(m.height) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int w = static_cast<int>(m.width), h = static_cast<int>(m.height);
    int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
    int x0 = ((ix % w) + w) % w, x1 = (((ix + 1) % w) + w) % w;
    int y0 = iy < 0 ? 0 : (iy > h - 1 ? h - 1 : iy);
    int y1 = iy + 1 < 0 ? 0 : (iy + 1 > h - 1 ? h - 1 : iy + 1);
    float4 a = m.texels[y0 * w + x0], b = m.texels[y0 * w + x1];
    float4 c = m.texels[y1 * w + x0], d = m.texels[y1 * w + x1];
    float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty), w01 = (1.0f - tx) * ty, w11 = tx * ty;
    Vec3 r(a.x * w00 + b.x * w10 + c.x * w01 + d.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + d.y * w11,
        a.z * w00 + b.z * w10 + c.z * w01 + d.z * w11);
    return r * m.intensity;
}

}  // namespace wpt
