// One caustic photon: emission from a triangle light and its path through
// mirrors and dielectrics to the first diffuse surface (emitPhotons and
// tracePhotons, src/renderer/photon_mapping.cu:4-74).  A photon is kept
// only if it arrives after at least one specular bounce.
#pragma once

#include "shade.h"

namespace wpt {

struct PhotonSettings {
    uint32_t count;
    uint32_t maxDepth;
    uint32_t seed;
    uint32_t strategy;
    uint32_t spectral;
};

template <class Stack>
WPT_HD Photon tracePhoton(const DevScene& scene, const PhotonSettings& ps, uint32_t i, Stack& stack) {
    Photon p;
    p.valid = 0;
    p.index = i;
    p.pad = 0;
    p.position = p.direction = p.flux = Vec3();
    if (scene.lightCount == 0) return p;
    uint64_t rng = rngSeed(photonRngKey(i, ps.seed, ps.strategy), ps.strategy);
    uint32_t candidate = static_cast<uint32_t>(rngNextFloat(rng, ps.strategy) * static_cast<float>(scene.lightCount));
    if (candidate > scene.lightCount - 1) candidate = scene.lightCount - 1;
    LightAlias entry = scene.lightAlias[candidate];
    uint32_t lightIndex = rngNextFloat(rng, ps.strategy) < entry.probability ? candidate : entry.alias;
    const TriangleLight& light = scene.lights[lightIndex];
    float rootU = sqrtf(rngNextFloat(rng, ps.strategy));
    float v = rngNextFloat(rng, ps.strategy);
    Vec3 point = (1.0f - rootU) * light.v0 + rootU * ((1.0f - v) * light.v1 + v * light.v2);
    Vec3 direction = cosineHemisphere(light.normal, rng, ps.strategy);
    const Material& emitter = scene.materials[light.material];
    uint32_t channel = 3;
    Vec3 flux = emitter.emission * (light.area * 3.14159265359f / static_cast<float>(ps.count));
    if (ps.spectral) {
        uint32_t c = static_cast<uint32_t>(rngNextFloat(rng, ps.strategy) * 3.0f);
        channel = c < 2u ? c : 2u;
        float channelFlux = (channel == 0 ? emitter.emission.x : channel == 1 ? emitter.emission.y : emitter.emission.z) *
            (light.area * 3.14159265359f * 3.0f / static_cast<float>(ps.count));
        flux = Vec3(channel == 0 ? channelFlux : 0.0f, channel == 1 ? channelFlux : 0.0f,
            channel == 2 ? channelFlux : 0.0f);
    }
    Ray ray(point + light.normal * 0.001f, direction);
    uint32_t specularBounces = 0;
    for (uint32_t depth = 0; depth < ps.maxDepth; ++depth) {
        HitRecord h = traceClosest<true>(scene, ray, stack);
        if (h.kind == kHitNone || h.material >= scene.materialCount) break;
        SurfaceHit hit = surfaceAttributes<true>(scene, ray, h);
        const Material& m = scene.materials[hit.material];
        Vec3 normal = dot(hit.normal, -ray.direction) >= 0.0f ? hit.normal : -hit.normal;
        if (m.type == kDiffuse) {
            if (specularBounces != 0) {
                p.position = hit.position;
                p.direction = -ray.direction;
                p.flux = flux;
                p.valid = 1;
            }
            break;
        }
        Vec3 next;
        if (m.type == kMetal) {
            next = normalize(reflect(normalize(ray.direction), normal));
        } else if (m.type == kDielectric) {
            float ior = m.ior;
            if (m.dispersion > 0.0f && channel < 3) ior += (static_cast<float>(channel) - 1.0f) * m.dispersion;
            DielectricInterface b = dielectricInterface(normalize(ray.direction), hit.geometricNormal, hit.normal, ior);
            bool reflected = b.reflectance >= 1.0f || b.reflectance > rngNextFloat(rng, ps.strategy);
            next = reflected ? b.reflection : b.transmission;
        } else {
            next = cosineHemisphere(normal, rng, ps.strategy);
        }
        if (m.type == kMetal || m.type == kDielectric) ++specularBounces;
        flux = hadamard(flux, m.albedo);
        ray = Ray(offsetRayOrigin(hit.position, hit.geometricNormal, next, hit.t), next);
    }
    return p;
}

}  // namespace wpt
