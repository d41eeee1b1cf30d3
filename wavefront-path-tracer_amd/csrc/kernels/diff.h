// Differentiable path tracing: a forward estimator whose random decisions do
// not depend on the parameters, and its path-replay adjoint.
//
// No reference counterpart.  The parameters are an RGB albedo texture (on
// the static triangles of one material), one GGX roughness per reflective
// static triangle and one emission scale per emissive static triangle.  The
// estimator is built so that, with the random numbers fixed, the path's
// radiance is a smooth function of the parameters:
//   * reflective static triangles use albedo/pi + F0 D G / (4 cos_i cos_o)
//     (GGX with alpha = roughness, Smith G), continued by cosine sampling;
//     other diffuse surfaces are Lambertian with their material albedo;
//   * light selection uses the scene's alias table (built from the base
//     emission), MIS weights use the cosine pdf: neither depends on the
//     parameters;
//   * no Russian roulette and no clamp; dielectrics continue as in the
//     render path (Fresnel coin), which no parameter touches.
// The adjoint replays each path with the same random numbers.  Explicit
// terms (next-event estimation, emitter hits) are differentiated directly;
// the continuation weight w = f cos / pdf scales everything collected after
// the vertex, L_after = L_path - L_so_far, so its derivative contributes
// L_after * (dw/dtheta) / w.  L_path comes from the forward pass.
#pragma once

#include "shade.h"

namespace wpt {

constexpr float kDiffSpecularF0 = 0.04f;
constexpr float kDiffPi = 3.14159265358979f;
// Roughness / emission slots accumulated in LDS per workgroup.
constexpr uint32_t kDiffLdsSlots = 512;

struct DiffParams {
    const float* texture;    // texHeight x texWidth x 3, row j at v = (j + 0.5) / texHeight
    const float* roughness;  // per reflective static triangle slot
    const float* emission;   // per emissive static triangle slot
};

struct DiffGrads {
    float* texture;
    float* roughness;  // global copies; the kernels accumulate slots < kDiffLdsSlots in LDS first
    float* emission;
};

struct DiffLayout {
    uint32_t texWidth, texHeight;
    uint32_t texturedMaterial;    // kInvalid: no texture
    const uint32_t* staticSlot;   // static triangle -> roughness slot (reflective) or emission slot (emissive)
    const uint32_t* lightSlot;    // light -> emission slot, or kInvalid
    uint32_t roughnessCount, emissionCount;
};

// Bilinear footprint of uv with clamp-to-edge addressing.
struct TexFootprint {
    uint32_t index[4];  // texel index (row * width + column)
    float weight[4];
};

WPT_HD TexFootprint texFootprint(const DiffLayout& l, Vec2 uv) {
    float x = uv.x * static_cast<float>(l.texWidth) - 0.5f;
    float y = uv.y * static_cast<float>(l.texHeight) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float tx = x - fx, ty = y - fy;
    int ix = static_cast<int>(fx), iy = static_cast<int>(fy);
    int maxX = static_cast<int>(l.texWidth) - 1, maxY = static_cast<int>(l.texHeight) - 1;
    int x0 = ix < 0 ? 0 : (ix > maxX ? maxX : ix);
    int x1 = ix + 1 < 0 ? 0 : (ix + 1 > maxX ? maxX : ix + 1);
    int y0 = iy < 0 ? 0 : (iy > maxY ? maxY : iy);
    int y1 = iy + 1 < 0 ? 0 : (iy + 1 > maxY ? maxY : iy + 1);
    TexFootprint f;
    f.index[0] = static_cast<uint32_t>(y0) * l.texWidth + static_cast<uint32_t>(x0);
    f.index[1] = static_cast<uint32_t>(y0) * l.texWidth + static_cast<uint32_t>(x1);
    f.index[2] = static_cast<uint32_t>(y1) * l.texWidth + static_cast<uint32_t>(x0);
    f.index[3] = static_cast<uint32_t>(y1) * l.texWidth + static_cast<uint32_t>(x1);
    f.weight[0] = (1.0f - tx) * (1.0f - ty);
    f.weight[1] = tx * (1.0f - ty);
    f.weight[2] = (1.0f - tx) * ty;
    f.weight[3] = tx * ty;
    return f;
}

WPT_HD Vec3 texFetch(const float* texture, const TexFootprint& f) {
    Vec3 a(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < 4; ++k) {
        const float* t = texture + 3u * f.index[k];
        a = a + Vec3(t[0], t[1], t[2]) * f.weight[k];
    }
    return a;
}

// Surface of a diffuse hit: albedo, and for reflective static triangles the
// roughness slot (GGX lobe on) and the texture footprint.
struct DiffSurface {
    Vec3 albedo;
    float roughness;
    uint32_t roughnessSlot;  // kInvalid: Lambertian, no parameters
    uint32_t textured;
    TexFootprint footprint;
};

// GGX specular lobe and its derivative in alpha.
struct SpecularTerm {
    float value;   // F0 D G / (4 cos_i cos_o)
    float dAlpha;  // d value / d alpha
};

WPT_HD float smithG1(float mu, float alpha, float& dAlpha) {
    float a2 = alpha * alpha;
    float s = sqrtf(a2 + (1.0f - a2) * mu * mu);
    float denom = mu + s;
    dAlpha = -2.0f * mu * alpha * (1.0f - mu * mu) / (s * denom * denom);
    return 2.0f * mu / denom;
}

WPT_HD SpecularTerm specularTerm(const Vec3& n, const Vec3& wo, const Vec3& wi, float alpha) {
    SpecularTerm r{0.0f, 0.0f};
    float ci = dot(n, wi), co = dot(n, wo);
    if (!(ci > 0.0f) || !(co > 0.0f)) return r;
    Vec3 h = normalize(wi + wo);
    float nh = fmaxf(0.0f, dot(n, h));
    float c = nh * nh;
    float a2 = alpha * alpha;
    float k = c * (a2 - 1.0f) + 1.0f;
    float d = a2 / (kDiffPi * k * k);
    float dD = 2.0f * alpha * (k - 2.0f * a2 * c) / (kDiffPi * k * k * k);
    float dGi, dGo;
    float gi = smithG1(ci, alpha, dGi);
    float go = smithG1(co, alpha, dGo);
    float scale = kDiffSpecularF0 / (4.0f * ci * co);
    r.value = scale * (d * (gi * go));
    r.dAlpha = scale * (dD * (gi * go) + d * (dGi * go + gi * dGo));
    return r;
}

WPT_HD DiffSurface diffSurface(const DevScene& s, const DiffLayout& l, const DiffParams& p, const HitRecord& h,
    const SurfaceHit& hit, const Material& m) {
    DiffSurface d;
    d.albedo = m.albedo;
    d.roughness = 0.0f;
    d.roughnessSlot = kInvalid;
    d.textured = 0;
    if (h.kind == kHitStatic) {
        d.roughnessSlot = l.staticSlot[h.prim];
        if (d.roughnessSlot != kInvalid) d.roughness = p.roughness[d.roughnessSlot];
        if (hit.material == l.texturedMaterial && l.texWidth > 0) {
            d.textured = 1;
            d.footprint = texFootprint(l, hit.uv);
            d.albedo = texFetch(p.texture, d.footprint);
        }
    }
    return d;
}

// BSDF value (without the cosine) and the specular term.
WPT_HD Vec3 diffBsdf(const DiffSurface& d, const Vec3& n, const Vec3& wo, const Vec3& wi, SpecularTerm& spec) {
    spec = SpecularTerm{0.0f, 0.0f};
    Vec3 f = d.albedo * kInvPi;
    if (d.roughnessSlot != kInvalid) {
        spec = specularTerm(n, wo, wi, d.roughness);
        f = f + Vec3(spec.value, spec.value, spec.value);
    }
    return f;
}

// Where the adjoint sends its gradients; a no-op sink for the forward pass.
struct NoGradients {
    WPT_HD void texture(const TexFootprint&, const Vec3&) {}
    WPT_HD void roughness(uint32_t, float) {}
    WPT_HD void emission(uint32_t, float) {}
};

WPT_HD float emissionScale(const DiffParams& p, uint32_t slot) { return slot == kInvalid ? 1.0f : p.emission[slot]; }

// Light sample of a diffuse vertex (four draws: candidate, alias coin,
// sqrt(u), v) and its shadow segment.  `valid` is false when the geometry
// rules out a contribution, without a shadow ray.
struct DiffLightSample {
    uint32_t lightIndex;
    Vec3 direction;          // toward the light
    float distanceSquared;
    float shadingCosine, lightCosine;
    Ray shadow;
    float shadowTmax;
    bool valid;
};

WPT_HD DiffLightSample diffLightSample(const DevScene& s, const SurfaceHit& hit, const Vec3& n, PathState& path,
    uint32_t strategy) {
    DiffLightSample ls;
    uint32_t candidate = static_cast<uint32_t>(rand01(path, strategy) * static_cast<float>(s.lightCount));
    if (candidate > s.lightCount - 1) candidate = s.lightCount - 1;
    LightAlias entry = s.lightAlias[candidate];
    ls.lightIndex = rand01(path, strategy) < entry.probability ? candidate : entry.alias;
    const TriangleLight& light = s.lights[ls.lightIndex];
    float rootU = sqrtf(rand01(path, strategy));
    float v = rand01(path, strategy);
    Vec3 lightPosition = (1.0f - rootU) * light.v0 + rootU * (1.0f - v) * light.v1 + rootU * v * light.v2;
    Vec3 toLight = lightPosition - hit.position;
    ls.distanceSquared = lengthSquared(toLight);
    float distance = sqrtf(ls.distanceSquared);
    ls.direction = toLight / distance;
    float geometricCosine = fmaxf(0.0f, dot(hit.geometricNormal, ls.direction));
    ls.shadingCosine = fmaxf(0.0f, dot(n, ls.direction));
    ls.lightCosine = fmaxf(0.0f, dot(light.normal, -ls.direction));
    ls.valid = false;
    ls.shadowTmax = 0.0f;
    if (geometricCosine != 0.0f && ls.shadingCosine != 0.0f && ls.lightCosine != 0.0f) {
        Vec3 shadowOrigin = offsetRayOrigin(hit.position, hit.geometricNormal, ls.direction, hit.t);
        Vec3 shadowTarget = offsetRayOrigin(lightPosition, light.normal, -ls.direction, distance);
        Vec3 segment = shadowTarget - shadowOrigin;
        float shadowDistance = length(segment);
        if (shadowDistance > kRayMinDistance) {
            ls.shadow = Ray(shadowOrigin, segment / shadowDistance);
            ls.shadowTmax = shadowDistance - kRayMinDistance;
            ls.valid = true;
        }
    }
    return ls;
}

// Emitter reached by the path: its MIS-weighted emission (and, in the
// adjoint, the gradient of its emission scale).
template <bool Adjoint, class Sink>
WPT_HD void diffEmitterHit(const DevScene& s, const DiffLayout& l, const DiffParams& p, const PathState& path,
    const SurfaceHit& hit, const Material& m, uint32_t depth, const Vec3& delta, Vec3& acc, Sink& sink) {
    float misWeight = 1.0f;
    uint32_t slot = kInvalid;
    if (hit.lightIndex != kInvalid) {
        slot = l.lightSlot[hit.lightIndex];
        const TriangleLight& light = s.lights[hit.lightIndex];
        float lightCosine = fmaxf(0.0f, dot(light.normal, -path.ray.direction));
        if (lightCosine == 0.0f) {
            misWeight = 0.0f;
        } else if (depth > 0 && !path.specularBounce) {
            float lightPdf = light.selectionPdf * hit.t * hit.t / (lightCosine * light.area);
            misWeight = powerHeuristic(path.previousBsdfPdf, lightPdf);
        }
    }
    Vec3 term = hadamard(path.throughput, m.emission) * misWeight;
    acc += term * emissionScale(p, slot);
    if (Adjoint && slot != kInvalid) sink.emission(slot, dot(delta, term));
}

// A diffuse vertex after its light sample `ls` (visible or not): the light
// sample's contribution, the cosine-sampled continuation (two draws) and, in
// the adjoint, the vertex's texture and roughness gradients.  Returns false
// when the path ends here; else `direction` is the continuation.
template <bool Adjoint, class Sink>
WPT_HD bool diffSurfaceVertex(const DevScene& s, const DiffLayout& l, const DiffParams& p, PathState& path,
    const HitRecord& h, const SurfaceHit& hit, const Material& m, const Vec3& n, const DiffLightSample& ls,
    bool lightVisible, uint32_t strategy, const Vec3& delta, const Vec3& total, Vec3& acc, Sink& sink,
    Vec3& direction) {
    Vec3 wo = -path.ray.direction;
    DiffSurface surf = diffSurface(s, l, p, h, hit, m);
    Vec3 texGrad(0.0f, 0.0f, 0.0f);  // dLoss/dalbedo at this vertex, per channel
    float roughGrad = 0.0f;
    if (ls.valid && lightVisible) {
        const TriangleLight& light = s.lights[ls.lightIndex];
        float geometryTerm = ls.lightCosine * light.area / (ls.distanceSquared * light.selectionPdf);
        float lightPdf = light.selectionPdf * ls.distanceSquared / (ls.lightCosine * light.area);
        float bsdfPdf = ls.shadingCosine * kInvPi;
        float g = geometryTerm * powerHeuristic(lightPdf, bsdfPdf);
        uint32_t slot = l.lightSlot[ls.lightIndex];
        float e = emissionScale(p, slot);
        SpecularTerm spec;
        Vec3 f = diffBsdf(surf, n, wo, ls.direction, spec);
        Vec3 le = s.materials[light.material].emission;
        // C = T (f cos) Le g e, with base = T Le g cos.
        Vec3 base = hadamard(path.throughput, le) * (g * ls.shadingCosine);
        Vec3 fb = hadamard(base, f);
        acc += fb * e;
        if (Adjoint) {
            if (slot != kInvalid) sink.emission(slot, dot(delta, fb));
            Vec3 db = hadamard(delta, base) * e;
            texGrad += db * kInvPi;
            roughGrad += (db.x + db.y + db.z) * spec.dAlpha;
        }
    }
    // Cosine-sampled continuation.
    direction = sampleCosineHemisphere(n, path, strategy);
    float cosine = dot(n, direction);
    float pdf = fmaxf(0.0f, cosine) * kInvPi;
    bool continues = pdf > 0.0f && cosine > 0.0f && dot(direction, hit.geometricNormal) > 0.0f;
    Vec3 weight(0.0f, 0.0f, 0.0f);
    if (continues) {
        SpecularTerm spec;
        Vec3 f = diffBsdf(surf, n, wo, direction, spec);
        float cosOverPdf = cosine / pdf;
        weight = f * cosOverPdf;
        if (Adjoint) {
            // Everything collected after this vertex scales with weight.
            Vec3 after = total - acc;
            Vec3 ratio(weight.x > 0.0f ? delta.x * after.x / weight.x : 0.0f,
                weight.y > 0.0f ? delta.y * after.y / weight.y : 0.0f,
                weight.z > 0.0f ? delta.z * after.z / weight.z : 0.0f);
            texGrad += ratio * (cosOverPdf * kInvPi);
            roughGrad += (ratio.x + ratio.y + ratio.z) * (spec.dAlpha * cosOverPdf);
        }
    }
    if (Adjoint) {
        if (surf.textured) sink.texture(surf.footprint, texGrad);
        if (surf.roughnessSlot != kInvalid) sink.roughness(surf.roughnessSlot, roughGrad);
    }
    if (!continues) return false;
    path.throughput = hadamard(path.throughput, weight);
    path.previousBsdfPdf = pdf;
    path.specularBounce = 0;
    return true;
}

WPT_HD Vec3 facingNormal(const SurfaceHit& hit) {
    Vec3 n = hit.normal;
    if (dot(n, hit.geometricNormal) < 0.0f) n = -n;
    return n;
}

// Dielectric interface: the render path's Fresnel coin.  The differentiable
// estimator has no environment map (rejected at creation), so the interface
// needs no escape tests.
WPT_HD Vec3 diffDielectric(const DevScene& s, PathState& path, const SurfaceHit& hit, const Vec3& n, const Material& m,
    uint32_t strategy, uint32_t maxDepth) {
    ShadeParams interfaceParams{maxDepth, maxDepth, strategy, 0.0f, 0u, 0u};
    DeferOcclusion noEscapeTests;
    return continueDielectric<false>(s, path, hit, n, m, m.albedo, interfaceParams, noEscapeTests);
}

// One path with each vertex's shadow ray traced where its light sample is
// drawn: the forward pass (Adjoint false, returns the path's radiance), or
// the adjoint with `delta` = dLoss/dL_path and `total` the path's forward
// radiance, gradients going to `sink`.  The adjoint replays the forward
// pass's random numbers, so it visits the same vertices.  (A variant that
// deferred each shadow ray to the next loop iteration, so that one
// traversal call site serves both queries, held fewer registers but ran
// slower: 12.0 ms against 9.8 ms for the 512x512x16 diff-Cornell adjoint.)
template <bool Adjoint, class Stack, class Sink>
WPT_HD Vec3 diffInlinePath(const DevScene& s, const DiffLayout& l, const DiffParams& p, const CameraParams& cam,
    uint32_t x, uint32_t y, uint32_t width, uint32_t height, uint32_t sample, uint32_t strategy, uint32_t maxDepth,
    Stack& stack, const Vec3& delta, const Vec3& total, Sink& sink) {
    PathState path = primaryPath(cam, x, y, width, height, sample, strategy);
    Vec3 acc(0.0f, 0.0f, 0.0f);
    for (uint32_t depth = 0; depth < maxDepth; ++depth) {
        HitRecord h = traceClosest<false>(s, path.ray, stack);
        if (h.kind == kHitNone) break;
        SurfaceHit hit = surfaceAttributes<false>(s, path.ray, h);
        const Material& m = s.materials[hit.material];
        if (m.type == kEmissive) {
            diffEmitterHit<Adjoint>(s, l, p, path, hit, m, depth, delta, acc, sink);
            break;
        }
        Vec3 n = facingNormal(hit);
        Vec3 direction;
        if (m.type == kDielectric) {
            direction = diffDielectric(s, path, hit, n, m, strategy, maxDepth);
        } else {
            DiffLightSample ls;
            ls.valid = false;
            bool visible = false;
            if (s.lightCount != 0) {
                ls = diffLightSample(s, hit, n, path, strategy);
                visible = ls.valid && !traceOccluded<false>(s, ls.shadow, ls.shadowTmax, stack);
            }
            if (!diffSurfaceVertex<Adjoint>(s, l, p, path, h, hit, m, n, ls, visible, strategy, delta, total, acc,
                    sink, direction))
                break;
        }
        path.ray = Ray(offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t), direction);
        path.depth = depth + 1;
    }
    return acc;
}

}  // namespace wpt
