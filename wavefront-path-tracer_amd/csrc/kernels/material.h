// Materials at a hit: textures, the material lookups that read them, normal
// mapping, the GGX specular-glossiness lobes and the glTF metallic-roughness
// BSDF with its specular and clearcoat extensions.
//
// Reference: include/renderer/device_traversal.cuh:325-525 (texture uv,
// sampling, sRGB curves, material albedo/opacity/emission/metallic/
// roughness/thickness/normal, glTF surface), specular_glossiness.cuh:13-80,
// gltf_pbr.cuh:1-119, geometry.cuh:68-82 (Hit).
//
// Textures are RGBA8 in HBM, read as normalized floats (texel / 255) and
// filtered bilinearly with wrapping in both directions, as the reference's
// texture objects are configured (scene.cu:53-75).  The reference's texture
// unit quantizes the bilinear weights to 1/256; here the weights are exact
// FP32 products, so a filtered value can differ from the reference's by up
// to about 1/512 of the step between neighbouring texels.
#pragma once

#include "../common/rng.h"
#include "../scene/device_scene.h"

namespace wpt {

constexpr float kInvPi = 0.31830988618f;
constexpr float kTwoPi = 6.28318530718f;

// Reference Hit (geometry.cuh:68-82).  uv1 and the tangent frames are
// filled only for materials with texture bindings, the only readers.
struct SurfaceHit {
    float t;
    Vec3 position;
    Vec3 geometricNormal;
    Vec3 normal;
    Vec2 uv, uv1;
    Vec3 tangent, bitangent, tangent1, bitangent1;
    uint32_t material;
    uint32_t lightIndex;
};

// Barycentric texture coordinates of the first and second sets
// (getTriangleUv, device_traversal.cuh:120-132); (0, 0) when absent.
WPT_HD Vec2 triangleUv(const TriShade& tri, float u, float v) {
    if (!(tri.flags & kTriHasTexcoords)) return Vec2{0.0f, 0.0f};
    float w = 1.0f - u - v;
    return Vec2{w * tri.uv0.x + u * tri.uv1.x + v * tri.uv2.x, w * tri.uv0.y + u * tri.uv1.y + v * tri.uv2.y};
}

WPT_HD Vec2 triangleUv1(const DevScene& s, const TriShade& tri, float u, float v) {
    if (!(tri.flags & kTriHasTexcoords1) || tri.extra == kInvalid) return Vec2{0.0f, 0.0f};
    const TriExtra& e = s.triExtra[tri.extra];
    float w = 1.0f - u - v;
    return Vec2{w * e.uv1_0.x + u * e.uv1_1.x + v * e.uv1_2.x, w * e.uv1_0.y + u * e.uv1_1.y + v * e.uv1_2.y};
}

// ---- textures
WPT_HD float4 unpackTexel(uint32_t p) {
    return float4{static_cast<float>(p & 255u) / 255.0f, static_cast<float>((p >> 8) & 255u) / 255.0f,
        static_cast<float>((p >> 16) & 255u) / 255.0f, static_cast<float>(p >> 24) / 255.0f};
}

// Normalized coordinate wrapped into [0, 1]; non-finite coordinates read 0.
WPT_HD float wrapUnit(float u) {
    float w = u - floorf(u);
    return w >= 0.0f && w <= 1.0f ? w : 0.0f;
}

WPT_HD uint32_t wrapIndex(int i, int n) { return static_cast<uint32_t>(i < 0 ? i + n : (i >= n ? i - n : i)); }

// Bilinear lookup at normalized (u, v): texel centers at half-integers,
// both directions wrapping.
WPT_HD float4 textureBilinear(const DevTexture& t, float u, float v) {
    int w = static_cast<int>(t.width), h = static_cast<int>(t.height);
    float x = wrapUnit(u) * static_cast<float>(w) - 0.5f;
    float y = wrapUnit(v) * static_cast<float>(h) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int ix = static_cast<int>(fx), iy = static_cast<int>(fy);  // in [-1, size]
    uint32_t x0 = wrapIndex(ix, w), x1 = wrapIndex(ix + 1, w);
    uint32_t y0 = wrapIndex(iy, h), y1 = wrapIndex(iy + 1, h);
    float4 a = unpackTexel(t.texels[y0 * t.width + x0]), b = unpackTexel(t.texels[y0 * t.width + x1]);
    float4 c = unpackTexel(t.texels[y1 * t.width + x0]), d = unpackTexel(t.texels[y1 * t.width + x1]);
    float w00 = (1.0f - ax) * (1.0f - ay), w10 = ax * (1.0f - ay), w01 = (1.0f - ax) * ay, w11 = ax * ay;
    return float4{a.x * w00 + b.x * w10 + c.x * w01 + d.x * w11, a.y * w00 + b.y * w10 + c.y * w01 + d.y * w11,
        a.z * w00 + b.z * w10 + c.z * w01 + d.z * w11, a.w * w00 + b.w * w10 + c.w * w01 + d.w * w11};
}

// getTextureUv, device_traversal.cuh:325-338.
WPT_HD Vec2 textureUv(const TexBinding& b, const Vec2& uv0, const Vec2& uv1) {
    Vec2 uv = b.texCoord == 1 ? uv1 : uv0;
    if (!b.hasTransform) return uv;
    uv.x *= b.scale.x;
    uv.y *= b.scale.y;
    float sine = sinF(b.rotation), cosine = cosF(b.rotation);
    return Vec2{cosine * uv.x - sine * uv.y + b.offset.x, sine * uv.x + cosine * uv.y + b.offset.y};
}

// sampleTexture, device_traversal.cuh:340-347: the image is read at 1 - v.
WPT_HD float4 sampleTexture(const DevScene& s, const TexBinding& b, const Vec2& uv0, const Vec2& uv1,
    float4 fallback) {
    if (b.texture == kInvalid || b.texture >= s.textureCount) return fallback;
    Vec2 uv = textureUv(b, uv0, uv1);
    return textureBilinear(s.textures[b.texture], uv.x, 1.0f - uv.y);
}

WPT_HD float4 whiteTexel() { return float4{1.0f, 1.0f, 1.0f, 1.0f}; }

WPT_HD Vec3 srgbToLinear(const float4& t) { return Vec3(powF(t.x, 2.2f), powF(t.y, 2.2f), powF(t.z, 2.2f)); }

WPT_HD float gltfSrgbChannel(float value) {
    return value <= 0.04045f ? value / 12.92f : powF((value + 0.055f) / 1.055f, 2.4f);
}

WPT_HD Vec3 gltfSrgbToLinear(const float4& t) {
    return Vec3(gltfSrgbChannel(t.x), gltfSrgbChannel(t.y), gltfSrgbChannel(t.z));
}

// ---- material lookups (device_traversal.cuh:359-500).  An untextured
// material (ext == kInvalid) reads its constants: every texture term would be
// the white fallback, which multiplies exactly to 1.
template <bool Full = true>
WPT_HD Vec3 materialAlbedo(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return m.albedo;
    const MaterialExt& e = s.materialExt[m.ext];
    float4 texel = sampleTexture(s, e.albedo, hit.uv, hit.uv1, whiteTexel());
    return hadamard(m.albedo, e.specularGlossinessWorkflow ? gltfSrgbToLinear(texel) : srgbToLinear(texel));
}

WPT_HD float materialOpacity(const DevScene& s, const Material& m, const Vec2& uv, const Vec2& uv1) {
    if (m.ext == kInvalid) return m.opacity;
    const MaterialExt& e = s.materialExt[m.ext];
    if (e.albedo.texture == kInvalid || e.albedo.texture >= s.textureCount) return m.opacity;
    return m.opacity * sampleTexture(s, e.albedo, uv, uv1, whiteTexel()).w;
}

template <bool Full = true>
WPT_HD Vec3 materialEmission(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return m.emission;
    return hadamard(m.emission, srgbToLinear(sampleTexture(s, s.materialExt[m.ext].emissive, hit.uv, hit.uv1,
        whiteTexel())));
}

template <bool Full = true>
WPT_HD float materialMetallic(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return m.metallic;
    return m.metallic * sampleTexture(s, s.materialExt[m.ext].metallicRoughness, hit.uv, hit.uv1, whiteTexel()).z;
}

template <bool Full = true>
WPT_HD float materialRoughness(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return m.roughness;
    const MaterialExt& e = s.materialExt[m.ext];
    if (e.specularGlossinessWorkflow)
        return 1.0f - m.glossiness * sampleTexture(s, e.specularGlossiness, hit.uv, hit.uv1, whiteTexel()).w;
    return m.roughness * sampleTexture(s, e.metallicRoughness, hit.uv, hit.uv1, whiteTexel()).y;
}

template <bool Full = true>
WPT_HD float materialThickness(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return 1.0f;
    const TexBinding& b = s.materialExt[m.ext].thickness;
    if (b.texture == kInvalid || b.texture >= s.textureCount) return 1.0f;
    return sampleTexture(s, b, hit.uv, hit.uv1, whiteTexel()).x;
}

// Finite test that holds for host and device builds alike (false for NaN
// and both infinities).
WPT_HD bool finiteF(float x) { return fabsf(x) <= 3.402823466e38f; }
WPT_HD bool finite3(const Vec3& v) { return finiteF(v.x) && finiteF(v.y) && finiteF(v.z); }

// getMaterialNormal, device_traversal.cuh:452-500: the tangent-space normal
// of `binding` (scaled in x and y) in the hit's re-orthogonalized tangent
// frame of the binding's coordinate set, optionally clamped to a minimum
// tangent-space cosine.
WPT_HD Vec3 mappedNormal(const DevScene& s, const TexBinding& binding, float normalScale, const SurfaceHit& hit) {
    if (!s.useNormalMaps || binding.texture == kInvalid || binding.texture >= s.textureCount) return hit.normal;
    Vec3 frameTangent = binding.texCoord == 1 ? hit.tangent1 : hit.tangent;
    Vec3 frameBitangent = binding.texCoord == 1 ? hit.bitangent1 : hit.bitangent;
    if (!finite3(frameTangent) || !finite3(frameBitangent) || lengthSquared(frameTangent) < 0.00000001f ||
        lengthSquared(frameBitangent) < 0.00000001f)
        return hit.normal;
    Vec3 tangent = normalize(frameTangent - hit.normal * dot(hit.normal, frameTangent));
    Vec3 bitangent = normalize(cross(hit.normal, tangent));
    if (!finite3(tangent) || !finite3(bitangent)) return hit.normal;
    if (dot(bitangent, frameBitangent) < 0.0f) bitangent = -bitangent;
    float4 texel = sampleTexture(s, binding, hit.uv, hit.uv1, float4{0.5f, 0.5f, 1.0f, 1.0f});
    Vec3 local((texel.x * 2.0f - 1.0f) * normalScale, (texel.y * 2.0f - 1.0f) * normalScale, texel.z * 2.0f - 1.0f);
    local = normalize(local);
    if (s.normalMapMinimumCosine > 0.0f && local.z < s.normalMapMinimumCosine) {
        float minimumCosine = fminf(s.normalMapMinimumCosine, 1.0f);
        float tangentLength = sqrtf(fmaxf(0.0f, local.x * local.x + local.y * local.y));
        if (tangentLength > 0.0f) {
            float scale = sqrtf(fmaxf(0.0f, 1.0f - minimumCosine * minimumCosine)) / tangentLength;
            local.x *= scale;
            local.y *= scale;
        } else {
            local.x = 0.0f;
            local.y = 0.0f;
        }
        local.z = minimumCosine;
    }
    Vec3 normal = tangent * local.x + bitangent * local.y + hit.normal * local.z;
    if (!finite3(normal) || lengthSquared(normal) < 0.00000001f) return hit.normal;
    return normalize(normal);
}

template <bool Full = true>
WPT_HD Vec3 materialNormal(const DevScene& s, const Material& m, const SurfaceHit& hit) {
    if (!Full || m.ext == kInvalid) return hit.normal;
    const MaterialExt& e = s.materialExt[m.ext];
    return mappedNormal(s, e.normal, e.normalScale, hit);
}

// ---- GGX specular-glossiness lobes (specular_glossiness.cuh:13-80)
struct SpecGlossSurface {
    Vec3 diffuse;
    Vec3 f0;
    float roughness;
};

WPT_HD float specGlossMax(const Vec3& v) { return fmaxf(v.x, fmaxf(v.y, v.z)); }

WPT_HD float specGlossProbability(const SpecGlossSurface& s) {
    float specular = specGlossMax(s.f0);
    float diffuse = specGlossMax(s.diffuse) * (1.0f - specular);
    return fminf(0.95f, fmaxf(0.05f, specular / fmaxf(1.0e-6f, specular + diffuse)));
}

WPT_HD float specGlossAlpha(const SpecGlossSurface& s) {
    float roughness = fmaxf(0.02f, s.roughness);
    return roughness * roughness;
}

WPT_HD float specGlossD(float nh, float alpha) {
    const float pi = 3.14159265359f;
    float a2 = alpha * alpha;
    float denominator = (1.0f - nh * nh) + a2 * nh * nh;
    return a2 / (pi * denominator * denominator);
}

WPT_HD float specGlossG1(float cosine, float alpha) {
    return 2.0f * cosine / (cosine + sqrtf(alpha * alpha + (1.0f - alpha * alpha) * cosine * cosine));
}

WPT_HD Vec3 evaluateSpecGloss(const SpecGlossSurface& s, const Vec3& n, const Vec3& outgoing, const Vec3& incoming,
    float& pdf) {
    pdf = 0.0f;
    float nv = dot(n, outgoing), nl = dot(n, incoming);
    Vec3 sum = outgoing + incoming;
    if (nv <= 0.0f || nl <= 0.0f || lengthSquared(sum) < 1.0e-12f) return Vec3();
    Vec3 half = normalize(sum);
    float nh = fminf(1.0f, fmaxf(0.0f, dot(n, half)));
    float vh = fminf(1.0f, fmaxf(1.0e-8f, dot(outgoing, half)));
    float alpha = specGlossAlpha(s);
    float d = specGlossD(nh, alpha);
    float probability = specGlossProbability(s);
    pdf = (1.0f - probability) * nl * kInvPi + probability * d * nh / (4.0f * vh);
    float schlick = powF(1.0f - vh, 5.0f);
    Vec3 fresnel = s.f0 + (Vec3(1.0f, 1.0f, 1.0f) - s.f0) * schlick;
    Vec3 diffuse = s.diffuse * ((1.0f - specGlossMax(s.f0)) * kInvPi);
    float g = specGlossG1(nv, alpha) * specGlossG1(nl, alpha);
    return diffuse + fresnel * (d * g / (4.0f * nv * nl));
}

WPT_HD Vec3 specGlossBasis(const Vec3& n, float x, float y, float z) {
    Vec3 helper = fabsf(n.x) > 0.1f ? Vec3(0.0f, 1.0f, 0.0f) : Vec3(1.0f, 0.0f, 0.0f);
    Vec3 tangent = normalize(cross(helper, n));
    return normalize(x * tangent + y * cross(n, tangent) + z * n);
}

WPT_HD Vec3 sampleSpecGloss(const SpecGlossSurface& s, const Vec3& n, const Vec3& outgoing, float choice, float u,
    float v) {
    float phi = kTwoPi * v;
    float sinPhi, cosPhi;
    sinCosF(phi, sinPhi, cosPhi);
    if (choice < specGlossProbability(s)) {
        float alpha = specGlossAlpha(s);
        float cosine = sqrtf((1.0f - u) / (1.0f + (alpha * alpha - 1.0f) * u));
        float sine = sqrtf(fmaxf(0.0f, 1.0f - cosine * cosine));
        Vec3 half = specGlossBasis(n, sine * cosPhi, sine * sinPhi, cosine);
        return 2.0f * dot(outgoing, half) * half - outgoing;
    }
    float radius = sqrtf(u);
    return specGlossBasis(n, radius * cosPhi, radius * sinPhi, sqrtf(1.0f - u));
}

// getSpecGlossSurface, device_traversal.cuh:396-401, and the metal lobe of
// evaluateSurface (:432-436).
template <bool Full = true>
WPT_HD SpecGlossSurface specGlossSurface(const DevScene& s, const Material& m, const SurfaceHit& hit,
    const Vec3& albedo, float roughness) {
    Vec3 f0 = m.specular;
    if (Full && m.ext != kInvalid)
        f0 = hadamard(m.specular, gltfSrgbToLinear(sampleTexture(s, s.materialExt[m.ext].specularGlossiness, hit.uv,
            hit.uv1, whiteTexel())));
    return SpecGlossSurface{albedo, f0, roughness};
}

// ---- glTF metallic-roughness with specular and clearcoat (gltf_pbr.cuh)
struct GltfSurface {
    Vec3 color;
    Vec3 normal, coatNormal;
    float metallic, roughness;
    Vec3 dielectricF0;
    float dielectricF90;
    float coat, coatRoughness;
    Vec3 reflectionNormal;  // zero: use `normal`
};

struct GltfWeights {
    float diffuse, specular, coat;
};

// The reflective frame turned toward the geometric normal (24 bisection
// steps) until the ideal reflection clears a small horizon margin.
WPT_HD Vec3 gltfReflectionNormal(Vec3 normal, Vec3 geometric, Vec3 view) {
    float geometricView = dot(geometric, view);
    if (geometricView <= 0.0f) return geometric;
    float target = fminf(0.01f, 0.9f * geometricView);
    Vec3 reflected = 2.0f * dot(normal, view) * normal - view;
    if (dot(geometric, reflected) >= target) return normal;
    float lower = 0.0f, upper = 1.0f;
    for (int i = 0; i < 24; ++i) {
        float t = (lower + upper) * 0.5f;
        Vec3 candidate = normalize(normal * (1.0f - t) + geometric * t);
        reflected = 2.0f * dot(candidate, view) * candidate - view;
        if (dot(geometric, reflected) >= target)
            upper = t;
        else
            lower = t;
    }
    return normalize(normal * (1.0f - upper) + geometric * upper);
}

WPT_HD Vec3 pbrSpecularNormal(const GltfSurface& s) {
    return lengthSquared(s.reflectionNormal) > 0.0f ? s.reflectionNormal : s.normal;
}

WPT_HD void correctReflectionNormals(GltfSurface& s, const Vec3& geometric, const Vec3& view) {
    s.reflectionNormal = gltfReflectionNormal(s.normal, geometric, view);
    if (s.coat > 0.0f) s.coatNormal = gltfReflectionNormal(s.coatNormal, geometric, view);
}

WPT_HD float pbrSaturate(float x) { return fminf(1.0f, fmaxf(0.0f, x)); }

WPT_HD float pbrSchlick(float cosine) {
    float x = 1.0f - pbrSaturate(cosine);
    return x * x * x * x * x;
}

WPT_HD float pbrCoatWeight(const GltfSurface& s, const Vec3& view) {
    return s.coat * (0.04f + 0.96f * pbrSchlick(dot(s.coatNormal, view)));
}

WPT_HD Vec3 pbrDielectricFresnel(const GltfSurface& s, float cosine) {
    return s.dielectricF0 + (Vec3(s.dielectricF90, s.dielectricF90, s.dielectricF90) - s.dielectricF0) *
        pbrSchlick(cosine);
}

WPT_HD Vec3 pbrBaseFresnel(const GltfSurface& s, float cosine) {
    Vec3 metal = s.color + (Vec3(1.0f, 1.0f, 1.0f) - s.color) * pbrSchlick(cosine);
    return pbrDielectricFresnel(s, cosine) * (1.0f - s.metallic) + metal * s.metallic;
}

WPT_HD GltfWeights pbrWeights(const GltfSurface& s, const Vec3& view) {
    float coat = pbrCoatWeight(s, view);
    float specular = s.metallic + (1.0f - s.metallic) * specGlossMax(pbrDielectricFresnel(s,
        dot(pbrSpecularNormal(s), view)));
    return GltfWeights{(1.0f - coat) * (1.0f - specular), (1.0f - coat) * specular, coat};
}

WPT_HD float pbrAlpha(float roughness) { return fmaxf(1.0e-4f, roughness * roughness); }

WPT_HD float pbrReflectionPdf(const Vec3& normal, const Vec3& view, const Vec3& light, float roughness) {
    if (roughness <= 0.0f || dot(normal, view) <= 0.0f || dot(normal, light) <= 0.0f ||
        lengthSquared(view + light) < 1.0e-12f)
        return 0.0f;
    Vec3 half = normalize(view + light);
    float vh = dot(view, half), nh = dot(normal, half);
    return vh > 0.0f && nh > 0.0f ? specGlossD(pbrSaturate(nh), pbrAlpha(roughness)) * nh / (4.0f * vh) : 0.0f;
}

WPT_HD float pbrReflectionKernel(const Vec3& normal, const Vec3& view, const Vec3& light, float roughness) {
    float nv = dot(normal, view), nl = dot(normal, light);
    if (roughness <= 0.0f || nv <= 0.0f || nl <= 0.0f || lengthSquared(view + light) < 1.0e-12f) return 0.0f;
    float alpha = pbrAlpha(roughness), nh = pbrSaturate(dot(normal, normalize(view + light)));
    return specGlossD(nh, alpha) * specGlossG1(nv, alpha) * specGlossG1(nl, alpha) / (4.0f * nv * nl);
}

WPT_HD Vec3 pbrSampleReflection(const Vec3& normal, const Vec3& view, float roughness, float u, float v) {
    if (roughness <= 0.0f) return 2.0f * dot(normal, view) * normal - view;
    float alpha = pbrAlpha(roughness);
    float cosine = sqrtf((1.0f - u) / ((1.0f - u) + alpha * alpha * u));
    float sine = sqrtf(fmaxf(0.0f, 1.0f - cosine * cosine)), phi = kTwoPi * v;
    Vec3 half = specGlossBasis(normal, sine * cosF(phi), sine * sinF(phi), cosine);
    return 2.0f * dot(view, half) * half - view;
}

// f x cosine and the density of the lobe mixture; `diffusePdf` is the
// diffuse proposal's actual density.
WPT_HD Vec3 evaluateGltfCos(const GltfSurface& s, const Vec3& view, const Vec3& light, float diffusePdf, float& pdf) {
    GltfWeights w = pbrWeights(s, view);
    float nl = fmaxf(0.0f, dot(s.normal, light));
    Vec3 reflectionNormal = pbrSpecularNormal(s);
    pdf = w.diffuse * diffusePdf + w.specular * pbrReflectionPdf(reflectionNormal, view, light, s.roughness) +
        w.coat * pbrReflectionPdf(s.coatNormal, view, light, s.coatRoughness);
    Vec3 base;
    if (lengthSquared(view + light) > 1.0e-12f) {
        float vh = pbrSaturate(dot(view, normalize(view + light)));
        float diffuseWeight = (1.0f - s.metallic) * (1.0f - specGlossMax(pbrDielectricFresnel(s, vh)));
        if (dot(s.normal, view) > 0.0f) base = s.color * (diffuseWeight * nl * kInvPi);
        float specularCos = fmaxf(0.0f, dot(reflectionNormal, light));
        base += pbrBaseFresnel(s, vh) * (pbrReflectionKernel(reflectionNormal, view, light, s.roughness) * specularCos);
    }
    float coatCos = fmaxf(0.0f, dot(s.coatNormal, light));
    float coatValue = w.coat * pbrReflectionKernel(s.coatNormal, view, light, s.coatRoughness) * coatCos;
    return base * (1.0f - w.coat) + Vec3(coatValue, coatValue, coatValue);
}

// getGltfPbrSurface, device_traversal.cuh:502-517, with the reflective
// frames corrected for the view (correctGltfReflectionNormals).
WPT_HD GltfSurface gltfSurface(const DevScene& s, const Material& m, const SurfaceHit& hit, const Vec3& shadingNormal,
    const Vec3& albedo, float roughness, const Vec3& view) {
    const MaterialExt& e = s.materialExt[m.ext];
    float weight = e.specularWeightFactor * sampleTexture(s, e.specularWeight, hit.uv, hit.uv1, whiteTexel()).w;
    Vec3 color = hadamard(e.specularColorFactor, gltfSrgbToLinear(sampleTexture(s, e.specularColor, hit.uv, hit.uv1,
        whiteTexel())));
    float f0 = (m.ior - 1.0f) / (m.ior + 1.0f);
    f0 *= f0;
    GltfSurface g;
    g.dielectricF0 = Vec3(fminf(1.0f, f0 * color.x) * weight, fminf(1.0f, f0 * color.y) * weight,
        fminf(1.0f, f0 * color.z) * weight);
    Vec3 coatNormal = mappedNormal(s, e.clearcoatNormal, e.clearcoatNormalScale, hit);
    if (dot(coatNormal, hit.geometricNormal) < 0.0f) coatNormal = -coatNormal;
    g.color = albedo;
    g.normal = shadingNormal;
    g.coatNormal = coatNormal;
    g.metallic = materialMetallic(s, m, hit);
    g.roughness = roughness;
    g.dielectricF90 = weight;
    g.coat = e.clearcoatFactor * sampleTexture(s, e.clearcoat, hit.uv, hit.uv1, whiteTexel()).x;
    g.coatRoughness = e.clearcoatRoughnessFactor * sampleTexture(s, e.clearcoatRoughness, hit.uv, hit.uv1,
        whiteTexel()).y;
    g.reflectionNormal = Vec3();
    correctReflectionNormals(g, hit.geometricNormal, view);
    return g;
}

WPT_HD bool isGltf(const DevScene& s, const Material& m) {
    return m.ext != kInvalid && s.materialExt[m.ext].gltfPbrWorkflow;
}

}  // namespace wpt
