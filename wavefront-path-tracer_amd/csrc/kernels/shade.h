// Path shading: surface attributes of the winning hit, next-event estimation
// toward triangle, punctual and environment lights, BSDF continuation,
// Russian roulette and radiance accumulation.
//
// Reference: include/renderer/device_path_shading.cuh:6-173 (shadePaths),
// device_traversal.cuh:100-180 (hit attributes), :403-441 (environment
// mixture, surface evaluation), :529-569 (sampleDirectLight), :571-635 (sky,
// environment and dielectric continuation), :637-683 (environment and
// punctual NEE), :685-787 (diffuse, glossy and glTF continuations),
// device_sampling.cuh:6-93 (sampling, Fresnel), path_state.cuh:33-43
// (indirect clamp).  Materials and their BSDF lobes are in material.h.
//
// The random numbers are drawn in the reference's order; where the reference
// leaves the order unspecified (the three draws of randomInUnitSphere are
// constructor arguments) they are drawn x, y, z.
#pragma once

#include "../common/rng.h"
#include "../scene/environment.h"
#include "material.h"
#include "photon.h"
#include "traverse.h"

namespace wpt {

struct PathState {
    Ray ray;
    Vec3 throughput;
    Vec3 radiance;
    uint32_t pixel;
    uint32_t depth;
    uint64_t rng;
    uint32_t mediumMaterial;
    float mediumDensity;
    float previousBsdfPdf;
    uint32_t specularBounce;
    uint32_t active;
    float etaScale;
};

struct ShadeParams {
    uint32_t maxDepth;
    uint32_t rrStart;
    uint32_t rngStrategy;
    float clamp;
    uint32_t intersectionDebug;
    uint32_t shadingNormalDebug;
    const PhotonGridView* photons;  // caustic photon map gathered at diffuse vertices, or null
};

WPT_HD float rand01(PathState& p, uint32_t strategy) { return rngNextFloat(p.rng, strategy); }

// The indirect clamp of addPathRadiance: it depends on the contribution and
// the path's depth only, so a deferred addition can apply it up front.
// Scales the contribution down to mean magnitude `clamp` (> 0) if above it.
WPT_HD Vec3 clampMagnitude(Vec3 contribution, float clamp) {
    float magnitude = fabsf(contribution.x) / 3.0f + fabsf(contribution.y) / 3.0f + fabsf(contribution.z) / 3.0f;
    if (magnitude > clamp) contribution *= clamp / magnitude;
    return contribution;
}

WPT_HD Vec3 clampContribution(const PathState& path, Vec3 contribution, float clamp, bool emissionOrBackground) {
    bool indirect = path.depth > (emissionOrBackground ? 1u : 0u);
    if (indirect && clamp > 0.0f) contribution = clampMagnitude(contribution, clamp);
    return contribution;
}

WPT_HD void addPathRadiance(PathState& path, Vec3 contribution, float clamp, bool emissionOrBackground) {
    path.radiance += clampContribution(path, contribution, clamp, emissionOrBackground);
}

WPT_HD float powerHeuristic(float a, float b) {
    float a2 = a * a;
    float b2 = b * b;
    return a2 / (a2 + b2);
}

// ---- surface attributes of the winner (device_traversal.cuh:100-180, 1150-1183)
WPT_HD Vec3 triangleShadingNormal(const TriShade& tri, const Ray& ray, float u, float v, bool faceForward) {
    Vec3 normal;
    if (tri.flags & kTriHasNormals)
        normal = normalize((1.0f - u - v) * tri.n0 + u * tri.n1 + v * tri.n2);
    else
        normal = normalize(cross(tri.v1 - tri.v0, tri.v2 - tri.v0));
    return faceForward && dot(normal, ray.direction) > 0.0f ? -normal : normal;
}

WPT_HD Vec3 triangleGeometricNormal(const TriShade& tri, const Ray& ray, bool faceForward) {
    Vec3 normal = normalize(cross(tri.v1 - tri.v0, tri.v2 - tri.v0));
    return faceForward && dot(normal, ray.direction) > 0.0f ? -normal : normal;
}

// Texture frames of a textured triangle hit, in the triangle's space
// (intersectTriangleHit, device_traversal.cuh:159-176): interpolated vertex
// tangents made orthogonal to the shading normal, with the handedness sign
// giving the bitangent, else the face frame.
WPT_HD void triangleFrames(const DevScene& s, const TriShade& tri, float u, float v, const Vec3& normal,
    SurfaceHit& hit) {
    hit.uv1 = triangleUv1(s, tri, u, v);
    hit.tangent = tri.tangent;
    hit.bitangent = tri.bitangent;
    hit.tangent1 = Vec3();
    hit.bitangent1 = Vec3();
    if (tri.extra == kInvalid) return;
    const TriExtra& e = s.triExtra[tri.extra];
    if (tri.flags & kTriHasVertexTangents) {
        float w = 1.0f - u - v;
        Vec3 tangent = w * e.vertexTangent0 + u * e.vertexTangent1 + v * e.vertexTangent2;
        tangent = tangent - normal * dot(normal, tangent);
        if (lengthSquared(tangent) > 0.00000001f) {
            hit.tangent = normalize(tangent);
            float sign = w * e.sign0 + u * e.sign1 + v * e.sign2;
            hit.bitangent = (sign < 0.0f ? -1.0f : 1.0f) * normalize(cross(normal, hit.tangent));
        }
    }
    hit.tangent1 = e.tangent1;
    hit.bitangent1 = e.bitangent1;
}

template <bool Full = true>
WPT_HD SurfaceHit surfaceAttributes(const DevScene& s, const Ray& ray, const HitRecord& h) {
    SurfaceHit hit;
    hit.t = h.t;
    hit.material = h.material;
    hit.uv = Vec2{0.0f, 0.0f};
    hit.uv1 = Vec2{0.0f, 0.0f};
    hit.tangent = hit.bitangent = hit.tangent1 = hit.bitangent1 = Vec3();
    hit.lightIndex = kInvalid;
    if (h.kind == kHitSphere) {
        float4 sp = s.spheres[h.prim];
        hit.normal = normalize(ray.at(h.t) - xyz(sp));
        hit.geometricNormal = hit.normal;
    } else {
        const TriShade* tri;
        Ray local = ray;
        const Instance* inst = nullptr;
        if (h.kind == kHitInstance) {
            inst = &s.instances[h.instance];
            tri = &s.blases[inst->mesh].triangles[h.prim];
            local = toLocal(inst->transform, ray);
        } else if (h.kind == kHitStatic) {
            tri = &s.staticTriangles[h.prim];
        } else {
            tri = &s.flatTriangles[h.prim];
        }
        bool faceForward = s.materials[h.material].type != kDielectric;
        bool textured = Full && s.materials[h.material].ext != kInvalid;
        hit.uv = triangleUv(*tri, h.u, h.v);
        hit.geometricNormal = triangleGeometricNormal(*tri, local, faceForward);
        hit.normal = triangleShadingNormal(*tri, local, h.u, h.v, faceForward);
        if (textured) triangleFrames(s, *tri, h.u, h.v, hit.normal, hit);
        if (inst) {
            hit.lightIndex = inst->materialOverride == kInvalid ? tri->lightIndex : kInvalid;
            hit.geometricNormal = transformNormal(inst->transform, hit.geometricNormal);
            hit.normal = transformNormal(inst->transform, hit.normal);
            if (textured) {
                hit.tangent = normalize(transformVector(inst->transform, hit.tangent));
                hit.bitangent = normalize(transformVector(inst->transform, hit.bitangent));
                hit.tangent1 = normalize(transformVector(inst->transform, hit.tangent1));
                hit.bitangent1 = normalize(transformVector(inst->transform, hit.bitangent1));
            }
            if (hit.lightIndex != kInvalid) hit.lightIndex += inst->lightOffset;
        } else {
            hit.lightIndex = h.kind == kHitFlat ? kInvalid : tri->lightIndex;
        }
    }
    hit.position = ray.at(h.t);
    return hit;
}

// ---- sampling and BSDFs (device_sampling.cuh, specular_glossiness.cuh)
WPT_HD Vec3 reflect(const Vec3& incident, const Vec3& normal) {
    return incident - 2.0f * dot(incident, normal) * normal;
}

WPT_HD Vec3 refract(const Vec3& incident, const Vec3& normal, float eta) {
    float cosTheta = fminf(dot(-incident, normal), 1.0f);
    Vec3 perpendicular = eta * (incident + cosTheta * normal);
    Vec3 parallel = -sqrtf(fmaxf(0.0f, 1.0f - lengthSquared(perpendicular))) * normal;
    return perpendicular + parallel;
}

WPT_HD float dielectricReflectance(float cosine, float eta) {
    if (eta == 1.0f) return 0.0f;
    cosine = fminf(1.0f, fmaxf(0.0f, cosine));
    float sin2Transmitted = eta * eta * fmaxf(0.0f, 1.0f - cosine * cosine);
    if (sin2Transmitted >= 1.0f) return 1.0f;
    float transmittedCosine = sqrtf(1.0f - sin2Transmitted);
    float perpendicular = (eta * cosine - transmittedCosine) / (eta * cosine + transmittedCosine);
    float parallel = (cosine - eta * transmittedCosine) / (cosine + eta * transmittedCosine);
    return 0.5f * (perpendicular * perpendicular + parallel * parallel);
}

struct DielectricInterface {
    Vec3 normal;
    Vec3 reflection;
    Vec3 transmission;
    float eta;
    float reflectance;
    bool frontFace;
};

WPT_HD DielectricInterface dielectricInterface(const Vec3& incident, const Vec3& geometricNormal, Vec3 shadingNormal,
    float ior) {
    bool frontFace = dot(incident, geometricNormal) < 0.0f;
    Vec3 geometricFacing = frontFace ? geometricNormal : -geometricNormal;
    if (dot(shadingNormal, geometricNormal) < 0.0f) shadingNormal = -shadingNormal;
    Vec3 normal = frontFace ? shadingNormal : -shadingNormal;
    float eta = frontFace ? 1.0f / ior : ior;
    if (dot(incident, normal) >= 0.0f) normal = geometricFacing;
    float reflectance = dielectricReflectance(dot(-incident, normal), eta);
    Vec3 reflection = reflect(incident, normal);
    Vec3 transmission = refract(incident, normal, eta);
    if (dot(reflection, geometricFacing) <= 0.0f ||
        (reflectance < 1.0f && dot(transmission, geometricFacing) >= 0.0f)) {
        normal = geometricFacing;
        reflectance = dielectricReflectance(dot(-incident, normal), eta);
        reflection = reflect(incident, normal);
        transmission = refract(incident, normal, eta);
    }
    DielectricInterface r;
    r.normal = normal;
    r.reflection = reflection;
    r.transmission = transmission;
    r.eta = eta;
    r.reflectance = reflectance;
    r.frontFace = frontFace;
    return r;
}

WPT_HD Vec3 cosineHemisphere(const Vec3& normal, uint64_t& rng, uint32_t strategy) {
    float u1 = rngNextFloat(rng, strategy);
    float u2 = rngNextFloat(rng, strategy);
    float radius = sqrtf(u1);
    float phi = kTwoPi * u2;
    float sinPhi, cosPhi;
    sinCosF(phi, sinPhi, cosPhi);
    float x = radius * cosPhi;
    float y = radius * sinPhi;
    float z = sqrtf(1.0f - u1);
    Vec3 helper = fabsf(normal.x) > 0.1f ? Vec3(0.0f, 1.0f, 0.0f) : Vec3(1.0f, 0.0f, 0.0f);
    Vec3 tangent = normalize(cross(helper, normal));
    Vec3 bitangent = cross(normal, tangent);
    return normalize(tangent * x + bitangent * y + normal * z);
}

WPT_HD Vec3 sampleCosineHemisphere(const Vec3& normal, PathState& p, uint32_t strategy) {
    return cosineHemisphere(normal, p.rng, strategy);
}

WPT_HD Vec3 randomInUnitSphere(PathState& p, uint32_t strategy) {
    while (true) {
        float x = 2.0f * rand01(p, strategy) - 1.0f;
        float y = 2.0f * rand01(p, strategy) - 1.0f;
        float z = 2.0f * rand01(p, strategy) - 1.0f;
        Vec3 point(x, y, z);
        if (lengthSquared(point) < 1.0f) return point;
    }
}

// ---- occlusion policies
// Staged kernels defer area-light shadow rays to a later pass (the request
// carries the segment and the clamped contribution; NEE is the last radiance
// addition of a bounce, so resolving it later keeps the reference's order of
// floating-point additions).  Kernels that can traverse inline, and every
// scene with an environment map (whose environment NEE and dielectric
// escape tests happen mid-bounce), trace each shadow ray where the reference
// does.
struct ShadowRequest {
    Vec3 origin;
    Vec3 direction;
    float tmax;
    Vec3 contribution;
    uint32_t valid;
};

struct DeferOcclusion {
    static constexpr bool kInline = false;
    static constexpr bool kEvents = false;
    static constexpr bool kDeferGather = false;
    WPT_HD bool blocked(const DevScene&, const Ray&, float) { return false; }
    WPT_HD uint32_t record(const Ray&, float, const Vec3&, bool) { return 0; }
    WPT_HD void add(PathState& path, const Vec3& c) { path.radiance += c; }
    WPT_HD void terminateIfVisible(uint32_t) {}
};

// Ordered radiance events of one vertex, for staged shading that traces no
// ray itself.  Until the vertex's first occlusion-dependent addition, every
// addition goes to the path's radiance at once; from it on, each addition
// (the clamped contribution, with the ray and distance that must be
// unblocked for it, or none) is stored in order in the event buffer at the
// path's queue slot.  A later pass traces the rays and adds the visible
// contributions in the stored order, so the path's radiance sees the same
// floating-point additions as shading that traces inline.  An event can also
// end the path when its ray escapes (dielectric and glTF delta lobes toward
// the environment).
enum EventField : uint32_t {
    kEvCount, kEvTarget, kEvPixel, kEvRx, kEvRy, kEvRz,  // per slot: events, continuation, pixel, radiance before them
    kEvGx, kEvGy, kEvGz,                                 // the throughput at a deferred caustic gather
    kEventHeaderFields,
};
enum EventItemField : uint32_t {
    kEiOx, kEiOy, kEiOz, kEiDx, kEiDy, kEiDz, kEiTmax, kEiCx, kEiCy, kEiCz, kEiFlags,
    kEventItemFields,
};
constexpr uint32_t kEventTraced = 1u;      // the ray must be unblocked for the addition
constexpr uint32_t kEventTerminates = 2u;  // an unblocked ray ends the path
constexpr uint32_t kEventBlocked = 4u;     // set by the pass that traces the ray, when blocked

struct EventBuffer {
    uint32_t* words;      // (kEventHeaderFields + maxEvents * kEventItemFields) x capacity
    uint32_t capacity;
    uint32_t maxEvents;
    // Staged shading: the traced events of a bounce, appended per workgroup
    // segment of raySegment entries; entry k names event rayEvent[k] of
    // slot raySlot[k].  The deferred caustic gathers likewise, at most one
    // per slot, in segments of gatherSegment entries.
    uint32_t* raySlot;
    uint8_t* rayEvent;
    uint32_t raySegment;
    uint32_t* gatherSlot;
    uint8_t* gatherEvent;
    uint32_t gatherSegment;
};

WPT_HD uint32_t* eventWord(const EventBuffer& b, uint32_t field, uint32_t slot) {
    return b.words + static_cast<size_t>(field) * b.capacity + slot;
}
WPT_HD uint32_t* eventItemWord(const EventBuffer& b, uint32_t event, uint32_t field, uint32_t slot) {
    return eventWord(b, kEventHeaderFields + event * kEventItemFields + field, slot);
}

struct EventOcclusion {
    static constexpr bool kInline = false;
    static constexpr bool kEvents = true;
    static constexpr bool kDeferGather = false;
    EventBuffer buffer;
    uint32_t slot;
    uint32_t count;
    uint32_t traced;  // bit e: stored event e has a ray to trace
    WPT_HD bool blocked(const DevScene&, const Ray&, float) { return false; }
    // Stores event `count` (bounded by the buffer's maxEvents, which the
    // renderer sizes from the scene's lights and materials) and returns its
    // index.
    WPT_HD uint32_t record(const Ray& ray, float tmax, const Vec3& c, bool traced) {
        uint32_t e = count++;
        if (e >= buffer.maxEvents) return e;
        *eventItemWord(buffer, e, kEiOx, slot) = bitsOf(ray.origin.x);
        *eventItemWord(buffer, e, kEiOy, slot) = bitsOf(ray.origin.y);
        *eventItemWord(buffer, e, kEiOz, slot) = bitsOf(ray.origin.z);
        *eventItemWord(buffer, e, kEiDx, slot) = bitsOf(ray.direction.x);
        *eventItemWord(buffer, e, kEiDy, slot) = bitsOf(ray.direction.y);
        *eventItemWord(buffer, e, kEiDz, slot) = bitsOf(ray.direction.z);
        *eventItemWord(buffer, e, kEiTmax, slot) = bitsOf(tmax);
        *eventItemWord(buffer, e, kEiCx, slot) = bitsOf(c.x);
        *eventItemWord(buffer, e, kEiCy, slot) = bitsOf(c.y);
        *eventItemWord(buffer, e, kEiCz, slot) = bitsOf(c.z);
        *eventItemWord(buffer, e, kEiFlags, slot) = traced ? kEventTraced : 0u;
        if (traced) this->traced |= 1u << e;
        return e;
    }
    WPT_HD void add(PathState& path, const Vec3& c) {
        if (count == 0)
            path.radiance += c;
        else
            record(Ray(Vec3(0.0f, 0.0f, 0.0f), Vec3(0.0f, 0.0f, 1.0f)), 0.0f, c, false);
    }
    WPT_HD void terminateIfVisible(uint32_t e) {
        if (e < count && e < buffer.maxEvents) *eventItemWord(buffer, e, kEiFlags, slot) = kEventTraced | kEventTerminates;
    }
};

// Staged event shading that also leaves the caustic gather to a pass of
// its own: the gather becomes an event whose contribution that pass fills
// in (from the query point, normal, albedo and clamp stored in the event's
// ray, contribution and distance fields and the throughput in the slot's
// header), so the shading kernel carries no gather code or heap.
struct StagedEventOcclusion : EventOcclusion {
    static constexpr bool kDeferGather = true;
    uint32_t gather;  // event index + 1 of the vertex's gather, 0: none
    WPT_HD void recordGather(const PathState& path, const Vec3& position, const Vec3& normal, const Vec3& albedo,
        float clamp) {
        uint32_t e = count++;
        if (e >= buffer.maxEvents) return;
        *eventItemWord(buffer, e, kEiOx, slot) = bitsOf(position.x);
        *eventItemWord(buffer, e, kEiOy, slot) = bitsOf(position.y);
        *eventItemWord(buffer, e, kEiOz, slot) = bitsOf(position.z);
        *eventItemWord(buffer, e, kEiDx, slot) = bitsOf(normal.x);
        *eventItemWord(buffer, e, kEiDy, slot) = bitsOf(normal.y);
        *eventItemWord(buffer, e, kEiDz, slot) = bitsOf(normal.z);
        // clampContribution's test, made here where the depth is known.
        *eventItemWord(buffer, e, kEiTmax, slot) = bitsOf(path.depth > 0u && clamp > 0.0f ? clamp : 0.0f);
        *eventItemWord(buffer, e, kEiCx, slot) = bitsOf(albedo.x);
        *eventItemWord(buffer, e, kEiCy, slot) = bitsOf(albedo.y);
        *eventItemWord(buffer, e, kEiCz, slot) = bitsOf(albedo.z);
        *eventItemWord(buffer, e, kEiFlags, slot) = 0u;
        *eventWord(buffer, kEvGx, slot) = bitsOf(path.throughput.x);
        *eventWord(buffer, kEvGy, slot) = bitsOf(path.throughput.y);
        *eventWord(buffer, kEvGz, slot) = bitsOf(path.throughput.z);
        gather = e + 1u;
    }
};

// A vertex's radiance addition: at once, or (event shading) as the next
// event, traced unless `traced` is false.  With inline or deferred
// occlusion the caller has already tested the ray.
template <class Occluder>
WPT_HD void emitRadiance(Occluder& occ, PathState& path, const Vec3& contribution, float clamp, const Ray& ray,
    float tmax, bool traced) {
    if constexpr (Occluder::kEvents) {
        Vec3 c = clampContribution(path, contribution, clamp, false);
        if (traced)
            occ.record(ray, tmax, c, true);
        else
            occ.add(path, c);
    } else {
        addPathRadiance(path, contribution, clamp, false);
    }
}

// Shadow and escape rays traced where shading draws them.  The full-feature
// shading reaches this from about ten sites; as a call it exists once in the
// kernel's code instead of once per site (kBatch<true>: 1.08 MB of code
// inlined, 0.45 MB with the call; the terrain frame 928 -> 780 ms).
template <class Stack>
struct InlineOcclusion {
    static constexpr bool kInline = true;
    static constexpr bool kEvents = false;
    static constexpr bool kDeferGather = false;
    Stack& stack;
    uint32_t queries;
    WPT_CALL bool blocked(const DevScene& s, const Ray& ray, float tmax) {
        ++queries;
        return traceOccluded(s, ray, tmax, stack);
    }
    WPT_HD uint32_t record(const Ray&, float, const Vec3&, bool) { return 0; }
    WPT_HD void add(PathState& path, const Vec3& c) { path.radiance += c; }
    WPT_HD void terminateIfVisible(uint32_t) {}
};

// ---- environment sampling and MIS (device_traversal.cuh:403-425, 571-587,
// 637-651)
template <bool Full = true>
WPT_HD float diffuseEnvironmentProbability(const DevScene& s) {
    if (!Full) return 0.0f;
    const EnvironmentMap& m = s.environment;
    return m.texels && m.intensity > 0.0f && m.width && m.height && m.probabilities && m.alias ? 0.5f : 0.0f;
}

// Mixture of cosine sampling and environment sampling (half each when a
// usable map exists), the density the diffuse continuation samples.
template <bool Full = true>
WPT_HD float diffuseContinuationPdf(const DevScene& s, const Vec3& normal, const Vec3& direction) {
    float probability = diffuseEnvironmentProbability<Full>(s);
    float cosinePdf = fmaxf(0.0f, dot(normal, direction)) * 0.31830988618f;
    return probability > 0.0f ? (1.0f - probability) * cosinePdf +
        probability * environmentPdf(s.environment, direction) : cosinePdf;
}

// Five draws: alias candidate (two, combined in double), coin, longitude,
// latitude.
WPT_HD EnvironmentSample samplePathEnvironment(const DevScene& s, PathState& path, uint32_t strategy) {
    float coarse = rand01(path, strategy), fine = rand01(path, strategy);
    double candidate = static_cast<double>(coarse) + static_cast<double>(fine) * (1.0 / 16777216.0);
    float coin = rand01(path, strategy), u = rand01(path, strategy), v = rand01(path, strategy);
    return sampleEnvironment(s.environment, candidate, coin, u, v);
}

// ---- the BSDF at a vertex, as NEE evaluates it (evaluateSurfaceCos,
// device_traversal.cuh:427-441, 519-527): f x cosine toward a light and
// the density with which the continuation would have sampled it.
//
// Diffuse (Lambert with the environment-mixture density) and the GGX lobes
// of metals and specular-glossiness materials; NEE skips lights below the
// shading normal.
template <bool Full>
struct StandardBsdf {
    static constexpr bool kOneSided = true;
    static constexpr bool kFull = Full;
    bool glossy;
    Vec3 normal;
    SpecGlossSurface sg;  // diffuse: sg.diffuse is the albedo
    WPT_HD Vec3 evalCos(const DevScene& s, const Vec3& view, const Vec3& light, float& pdf) const {
        Vec3 value;
        if (glossy) {
            value = evaluateSpecGloss(sg, normal, view, light, pdf);
        } else {
            pdf = diffuseContinuationPdf<Full>(s, normal, light);
            value = sg.diffuse * kInvPi;
        }
        return value * fmaxf(0.0f, dot(normal, light));
    }
};

// glTF metallic-roughness with specular and clearcoat: evaluated for lights
// below the shading normal too (the coat and reflection frames differ).
struct GltfBsdf {
    static constexpr bool kOneSided = false;
    static constexpr bool kFull = true;
    GltfSurface g;
    WPT_HD Vec3 evalCos(const DevScene& s, const Vec3& view, const Vec3& light, float& pdf) const {
        return evaluateGltfCos(g, view, light, diffuseContinuationPdf(s, g.normal, light), pdf);
    }
};

// ---- next-event estimation toward triangle lights (device_traversal.cuh:529-569)
template <class Occluder, class Bsdf>
WPT_HD void sampleDirectLight(const DevScene& s, PathState& path, const SurfaceHit& hit, const Vec3& shadingNormal,
    const Bsdf& bsdfAt, const ShadeParams& sp, Occluder& occ, ShadowRequest& request) {
    if (s.lightCount == 0) return;
    uint32_t candidate = static_cast<uint32_t>(rand01(path, sp.rngStrategy) * static_cast<float>(s.lightCount));
    if (candidate > s.lightCount - 1) candidate = s.lightCount - 1;
    LightAlias entry = s.lightAlias[candidate];
    uint32_t lightIndex = rand01(path, sp.rngStrategy) < entry.probability ? candidate : entry.alias;
    const TriangleLight& light = s.lights[lightIndex];
    float rootU = sqrtf(rand01(path, sp.rngStrategy));
    float v = rand01(path, sp.rngStrategy);
    Vec3 lightPosition = (1.0f - rootU) * light.v0 + rootU * (1.0f - v) * light.v1 + rootU * v * light.v2;
    Vec3 toLight = lightPosition - hit.position;
    float distanceSquared = lengthSquared(toLight);
    float distance = sqrtf(distanceSquared);
    Vec3 direction = toLight / distance;
    float geometricCosine = fmaxf(0.0f, dot(hit.geometricNormal, direction));
    float shadingCosine = fmaxf(0.0f, dot(shadingNormal, direction));
    float lightCosine = fmaxf(0.0f, dot(light.normal, -direction));
    if (geometricCosine == 0.0f || (Bsdf::kOneSided && shadingCosine == 0.0f) || lightCosine == 0.0f) return;
    Vec3 shadowOrigin = offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t);
    Vec3 shadowTarget = offsetRayOrigin(lightPosition, light.normal, -direction, distance);
    Vec3 segment = shadowTarget - shadowOrigin;
    float shadowDistance = length(segment);
    if (!(shadowDistance > kRayMinDistance)) return;
    Ray shadow(shadowOrigin, segment / shadowDistance);
    if (Occluder::kInline && !s.ignoreDirectLightOcclusion && occ.blocked(s, shadow, shadowDistance - kRayMinDistance))
        return;
    const Material& lightMaterial = s.materials[light.material];
    float geometryTerm = lightCosine * light.area / (distanceSquared * light.selectionPdf);
    float lightPdf = light.selectionPdf * distanceSquared / (lightCosine * light.area);
    float bsdfPdf;
    Vec3 bsdf = bsdfAt.evalCos(s, -path.ray.direction, direction, bsdfPdf);
    float misWeight = powerHeuristic(lightPdf, bsdfPdf);
    Vec3 directLighting = hadamard(bsdf, lightMaterial.emission) * (geometryTerm * misWeight);
    Vec3 contribution = hadamard(path.throughput, directLighting);
    if constexpr (Occluder::kEvents) {
        emitRadiance(occ, path, contribution, sp.clamp, shadow, shadowDistance - kRayMinDistance,
            !s.ignoreDirectLightOcclusion);
        return;
    }
    if (Occluder::kInline || s.ignoreDirectLightOcclusion) {
        addPathRadiance(path, contribution, sp.clamp, false);
        return;
    }
    // addPathRadiance's clamp applied up front; the addition happens on resolve.
    if (path.depth > 0u && sp.clamp > 0.0f) {
        float magnitude = fabsf(contribution.x) / 3.0f + fabsf(contribution.y) / 3.0f + fabsf(contribution.z) / 3.0f;
        if (magnitude > sp.clamp) contribution *= sp.clamp / magnitude;
    }
    request.origin = shadowOrigin;
    request.direction = shadow.direction;
    request.tmax = shadowDistance - kRayMinDistance;
    request.contribution = contribution;
    request.valid = 1;
}

// ---- punctual lights (punctual_light.cuh:24-50, device_traversal.cuh:653-683)
struct PunctualSample {
    Vec3 direction;  // toward the light
    Vec3 incident;   // radiant intensity arriving at the point
    float distance;
};

WPT_HD PunctualSample samplePunctualLight(const PunctualLight& light, const Vec3& point) {
    PunctualSample r;
    r.distance = 1.0e30f;
    if (light.type == kLightDirectional) {
        r.direction = -light.direction;
        r.incident = light.intensity;
        return r;
    }
    Vec3 offset = light.position - point;
    float distanceSquared = lengthSquared(offset);
    if (!(distanceSquared > 0.0f)) {
        r.direction = Vec3();
        r.incident = Vec3();
        return r;
    }
    r.distance = sqrtf(distanceSquared);
    r.direction = offset / r.distance;
    float attenuation = 1.0f / distanceSquared;
    if (light.range > 0.0f) {
        float ratio = r.distance / light.range;
        attenuation *= fmaxf(0.0f, 1.0f - ratio * ratio * ratio * ratio);
    }
    if (light.type == kLightSpot) {
        float cosine = dot(light.direction, -r.direction);
        float cone = fminf(1.0f, fmaxf(0.0f, (cosine - light.outerConeCos) /
            fmaxf(1.0e-6f, light.innerConeCos - light.outerConeCos)));
        attenuation *= cone * cone;
    }
    r.incident = light.intensity * attenuation;
    return r;
}

// One punctual connection: a delta light, so no MIS weight.
template <class Occluder, class Bsdf>
WPT_HD void addPunctualLight(const DevScene& s, PathState& path, const SurfaceHit& hit, const Vec3& shadingNormal,
    const Bsdf& bsdfAt, const ShadeParams& sp, Occluder& occ, const PunctualLight& source,
    float inverseSelectionPdf) {
    PunctualSample light = samplePunctualLight(source, hit.position);
    if ((Bsdf::kOneSided && dot(shadingNormal, light.direction) <= 0.0f) ||
        dot(hit.geometricNormal, light.direction) <= 0.0f ||
        lengthSquared(light.incident) == 0.0f)
        return;
    Ray shadow(offsetRayOrigin(hit.position, hit.geometricNormal, light.direction, hit.t), light.direction);
    float shadowDistance = light.distance;
    if (source.type != kLightDirectional) {
        Vec3 segment = source.position - shadow.origin;
        shadowDistance = length(segment);
        if (!(shadowDistance > kRayMinDistance)) return;
        shadow.direction = segment / shadowDistance;
    }
    bool traced = !s.ignoreDirectLightOcclusion;
    if (traced && occ.blocked(s, shadow, shadowDistance - kRayMinDistance)) return;
    float ignoredPdf;
    Vec3 bsdf = bsdfAt.evalCos(s, -path.ray.direction, light.direction, ignoredPdf);
    Vec3 direct = hadamard(bsdf, light.incident) * inverseSelectionPdf;
    emitRadiance(occ, path, hadamard(path.throughput, direct), sp.clamp, shadow, shadowDistance - kRayMinDistance,
        traced);
}

// Every directional light, then one positional light chosen uniformly (one
// draw when positional lights exist).
template <class Occluder, class Bsdf>
WPT_HD void sampleDirectPunctualLight(const DevScene& s, PathState& path, const SurfaceHit& hit,
    const Vec3& shadingNormal, const Bsdf& bsdfAt, const ShadeParams& sp, Occluder& occ) {
    if (!Bsdf::kFull) return;
    for (uint32_t i = 0; i < s.directionalCount; ++i)
        addPunctualLight(s, path, hit, shadingNormal, bsdfAt, sp, occ, s.punctual[i], 1.0f);
    uint32_t count = s.punctualCount - s.directionalCount;
    if (count == 0) return;
    uint32_t pick = static_cast<uint32_t>(rand01(path, sp.rngStrategy) * static_cast<float>(count));
    uint32_t index = s.directionalCount + (pick < count - 1 ? pick : count - 1);
    addPunctualLight(s, path, hit, shadingNormal, bsdfAt, sp, occ, s.punctual[index], static_cast<float>(count));
}

// Next-event estimation toward the environment (device_traversal.cuh:637-651).
template <class Occluder, class Bsdf>
WPT_HD void sampleDirectEnvironment(const DevScene& s, PathState& path, const SurfaceHit& hit,
    const Vec3& shadingNormal, const Bsdf& bsdfAt, const ShadeParams& sp, Occluder& occ) {
    if (!Bsdf::kFull) return;
    const EnvironmentMap& m = s.environment;
    if (!m.texels || m.intensity == 0.0f) return;
    EnvironmentSample light = samplePathEnvironment(s, path, sp.rngStrategy);
    float cosine = dot(shadingNormal, light.direction);
    if ((Bsdf::kOneSided && cosine <= 0.0f) || dot(hit.geometricNormal, light.direction) <= 0.0f ||
        !(light.pdf > 0.0f))
        return;
    Ray shadow(offsetRayOrigin(hit.position, hit.geometricNormal, light.direction, hit.t), light.direction);
    bool traced = !s.ignoreDirectLightOcclusion;
    if (traced && occ.blocked(s, shadow, 1.0e30f)) return;
    float bsdfPdf;
    Vec3 bsdf = bsdfAt.evalCos(s, -path.ray.direction, light.direction, bsdfPdf);
    float weight = powerHeuristic(light.pdf, bsdfPdf);
    Vec3 direct = hadamard(bsdf, environmentRadiance(m, light.direction)) * (weight / light.pdf);
    emitRadiance(occ, path, hadamard(path.throughput, direct), sp.clamp, shadow, 1.0e30f, traced);
}

// ---- continuations
// Diffuse: cosine sampling, or half the time an environment sample when a
// map exists (device_traversal.cuh:685-704).
template <bool Full>
WPT_HD Vec3 continueDiffuse(const DevScene& s, PathState& path, const SurfaceHit& hit, const Vec3& shadingNormal,
    const Vec3& albedo, uint32_t strategy) {
    float probability = diffuseEnvironmentProbability<Full>(s);
    Vec3 direction = probability > 0.0f && rand01(path, strategy) < probability
        ? samplePathEnvironment(s, path, strategy).direction
        : sampleCosineHemisphere(shadingNormal, path, strategy);
    float cosine = dot(shadingNormal, direction);
    float pdf = diffuseContinuationPdf<Full>(s, shadingNormal, direction);
    if (!(pdf > 0.0f) || cosine <= 0.0f || dot(direction, hit.geometricNormal) <= 0.0f) {
        path.active = 0;
        return direction;
    }
    path.throughput = hadamard(path.throughput, albedo) * (cosine * kInvPi / pdf);
    path.previousBsdfPdf = pdf;
    path.specularBounce = 0;
    return direction;
}

WPT_HD Vec3 continueGlossy(PathState& path, const SurfaceHit& hit, const Vec3& shadingNormal,
    const SpecGlossSurface& surface, uint32_t strategy) {
    Vec3 outgoing = -path.ray.direction;
    float choice = rand01(path, strategy);
    float u = rand01(path, strategy);
    float v = rand01(path, strategy);
    Vec3 direction = sampleSpecGloss(surface, shadingNormal, outgoing, choice, u, v);
    float pdf;
    Vec3 bsdf = evaluateSpecGloss(surface, shadingNormal, outgoing, direction, pdf);
    if (!(pdf > 0.0f) || dot(direction, hit.geometricNormal) <= 0.0f) {
        path.active = 0;
        return direction;
    }
    path.throughput = hadamard(path.throughput, bsdf) * (fmaxf(0.0f, dot(shadingNormal, direction)) / pdf);
    path.previousBsdfPdf = pdf;
    path.specularBounce = 0;
    return direction;
}

// Dielectric: Fresnel coin between reflection and refraction
// (device_traversal.cuh:589-635).  With an environment map, both branches'
// escapes to the environment are integrated at the interface (each tested
// for occlusion), and a path whose chosen branch escapes ends here.
template <bool Full = true, class Occluder>
WPT_HD Vec3 continueDielectric(const DevScene& s, PathState& path, const SurfaceHit& hit, const Vec3& shadingNormal,
    const Material& m, const Vec3& albedo, const ShadeParams& sp, Occluder& occ) {
    path.throughput = hadamard(path.throughput, albedo);
    DielectricInterface boundary = dielectricInterface(normalize(path.ray.direction), hit.geometricNormal,
        shadingNormal, m.ior);
    float etaSquared = boundary.eta * boundary.eta;
    bool reflectionEscapes = false, transmissionEscapes = false;
    uint32_t reflectionEvent = kInvalid, transmissionEvent = kInvalid;  // event shading
    if (Full && s.environment.texels && path.depth + 1 < sp.maxDepth) {
        if (boundary.reflectance > 0.0f) {
            Vec3 direction = boundary.reflection;
            Ray ray(offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t), direction);
            if constexpr (Occluder::kEvents) {
                reflectionEvent = occ.record(ray, 1.0e30f, clampContribution(path, hadamard(path.throughput,
                    environmentRadiance(s.environment, direction)) * boundary.reflectance, sp.clamp, false), true);
            } else {
                reflectionEscapes = !occ.blocked(s, ray, 1.0e30f);
                if (reflectionEscapes)
                    addPathRadiance(path, hadamard(path.throughput, environmentRadiance(s.environment, direction)) *
                        boundary.reflectance, sp.clamp, false);
            }
        }
        if (boundary.reflectance < 1.0f) {
            Vec3 direction = boundary.transmission;
            Ray ray(offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t), direction);
            if constexpr (Occluder::kEvents) {
                transmissionEvent = occ.record(ray, 1.0e30f, clampContribution(path, hadamard(path.throughput,
                    environmentRadiance(s.environment, direction)) * ((1.0f - boundary.reflectance) * etaSquared),
                    sp.clamp, false), true);
            } else {
                transmissionEscapes = !occ.blocked(s, ray, 1.0e30f);
                if (transmissionEscapes)
                    addPathRadiance(path, hadamard(path.throughput, environmentRadiance(s.environment, direction)) *
                        ((1.0f - boundary.reflectance) * etaSquared), sp.clamp, false);
            }
        }
    }
    bool reflected = boundary.reflectance >= 1.0f || boundary.reflectance > rand01(path, sp.rngStrategy);
    if (!reflected) {
        path.throughput *= etaSquared;
        path.etaScale /= etaSquared;
        if (boundary.frontFace) {
            path.mediumMaterial = hit.material;
            path.mediumDensity = m.volumeDensity * materialThickness<Full>(s, m, hit);
        } else {
            path.mediumMaterial = kInvalid;
            path.mediumDensity = 1.0f;
        }
    }
    if constexpr (Occluder::kEvents)
        occ.terminateIfVisible(reflected ? reflectionEvent : transmissionEvent);
    else if (reflected ? reflectionEscapes : transmissionEscapes)
        path.active = 0;
    path.specularBounce = 1;
    path.previousBsdfPdf = 0.0f;
    // Per-component selects keep the interface in registers.
    return Vec3(reflected ? boundary.reflection.x : boundary.transmission.x,
        reflected ? boundary.reflection.y : boundary.transmission.y,
        reflected ? boundary.reflection.z : boundary.transmission.z);
}

// glTF metallic-roughness continuation (continueGltfPbrSurface,
// device_traversal.cuh:727-787): one draw picks coat, specular or diffuse
// by their weights; zero-roughness lobes are deltas.  With an environment
// map, visible delta reflections toward it are integrated at the vertex
// (each tested for occlusion), and a path whose sampled delta lobe escapes
// ends here.
template <class Occluder>
WPT_HD Vec3 continueGltf(const DevScene& s, PathState& path, const SurfaceHit& hit, const GltfSurface& g,
    const ShadeParams& sp, Occluder& occ) {
    Vec3 view = -path.ray.direction;
    GltfWeights weights = pbrWeights(g, view);
    float choice = rand01(path, sp.rngStrategy);
    bool coat = choice < weights.coat;
    bool specular = !coat && choice < weights.coat + weights.specular;
    bool delta = (coat && g.coatRoughness <= 0.0f) || (specular && g.roughness <= 0.0f);
    bool baseEscapes = false, coatEscapes = false;
    uint32_t baseEvent = kInvalid, coatEvent = kInvalid;  // event shading
    Vec3 baseCoefficient = pbrBaseFresnel(g, dot(pbrSpecularNormal(g), view)) * (1.0f - weights.coat);
    if (s.environment.texels && path.depth + 1 < sp.maxDepth) {
        for (int lobe = 0; lobe < 2; ++lobe) {
            bool isCoat = lobe == 1;
            float mass = isCoat ? weights.coat : weights.specular;
            float roughness = isCoat ? g.coatRoughness : g.roughness;
            if (!(mass > 0.0f) || roughness > 0.0f) continue;
            Vec3 normal = isCoat ? g.coatNormal : pbrSpecularNormal(g);
            Vec3 direction = pbrSampleReflection(normal, view, 0.0f, 0.0f, 0.0f);
            if (dot(normal, view) <= 0.0f || dot(direction, hit.geometricNormal) <= 0.0f) continue;
            Ray ray(offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t), direction);
            if constexpr (Occluder::kEvents) {
                Vec3 coefficient = isCoat ? Vec3(weights.coat, weights.coat, weights.coat) : baseCoefficient;
                uint32_t e = occ.record(ray, 1.0e30f, clampContribution(path, hadamard(path.throughput,
                    hadamard(coefficient, environmentRadiance(s.environment, direction))), sp.clamp, false), true);
                if (isCoat)
                    coatEvent = e;
                else
                    baseEvent = e;
                continue;
            }
            bool escapes = !occ.blocked(s, ray, 1.0e30f);
            if (isCoat)
                coatEscapes = escapes;
            else
                baseEscapes = escapes;
            if (escapes) {
                Vec3 coefficient = isCoat ? Vec3(weights.coat, weights.coat, weights.coat) : baseCoefficient;
                addPathRadiance(path, hadamard(path.throughput, hadamard(coefficient,
                    environmentRadiance(s.environment, direction))), sp.clamp, false);
            }
        }
    }
    Vec3 direction;
    if (coat || specular) {
        float u = rand01(path, sp.rngStrategy), v = rand01(path, sp.rngStrategy);
        Vec3 normal = coat ? g.coatNormal : pbrSpecularNormal(g);
        direction = pbrSampleReflection(normal, view, coat ? g.coatRoughness : g.roughness, u, v);
        if (dot(normal, view) <= 0.0f) {
            path.active = 0;
            return direction;
        }
    } else {
        float probability = diffuseEnvironmentProbability(s);
        direction = probability > 0.0f && rand01(path, sp.rngStrategy) < probability
            ? samplePathEnvironment(s, path, sp.rngStrategy).direction
            : sampleCosineHemisphere(g.normal, path, sp.rngStrategy);
    }
    if (dot(direction, hit.geometricNormal) <= 0.0f) {
        path.active = 0;
        return direction;
    }
    if (delta) {
        float mass = coat ? weights.coat : weights.specular;
        Vec3 coefficient = coat ? Vec3(weights.coat, weights.coat, weights.coat) : baseCoefficient;
        path.throughput = hadamard(path.throughput, coefficient) / mass;
        path.specularBounce = 1;
        path.previousBsdfPdf = 0.0f;
        if constexpr (Occluder::kEvents)
            occ.terminateIfVisible(coat ? coatEvent : baseEvent);
        else if (coat ? coatEscapes : baseEscapes)
            path.active = 0;
    } else {
        float pdf;
        Vec3 value = evaluateGltfCos(g, view, direction, diffuseContinuationPdf(s, g.normal, direction), pdf);
        if (!(pdf > 0.0f)) {
            path.active = 0;
            return direction;
        }
        path.throughput = hadamard(path.throughput, value) / pdf;
        path.previousBsdfPdf = pdf;
        path.specularBounce = 0;
    }
    return direction;
}

// One bounce of shading.  Returns true when the path continues with
// path.ray set to the continuation.  A terminated path adds `debugColor` to
// its pixel when `isDebug`, else its radiance after `shadow` is resolved
// (deferred occlusion only).
struct ShadeOutcome {
    bool alive;
    bool isDebug;
    Vec3 debugColor;
    ShadowRequest shadow;
};

template <bool Full, class Occluder>
WPT_HD bool shadeBounce(const DevScene& s, PathState& path, const HitRecord& h, const ShadeParams& sp,
    Occluder& occ, ShadeOutcome& out) {
    bool didHit = h.kind != kHitNone;
    if (sp.intersectionDebug || sp.shadingNormalDebug) {
        Vec3 color(0.05f, 0.10f, 0.25f);
        if (didHit) {
            SurfaceHit hit = surfaceAttributes<Full>(s, path.ray, h);
            Vec3 normal = sp.intersectionDebug ? hit.normal : materialNormal<Full>(s, s.materials[hit.material], hit);
            color = 0.5f * (normal + Vec3(1.0f, 1.0f, 1.0f));
        }
        path.active = 0;
        out.isDebug = true;
        out.debugColor = color;
        return false;
    }
    if (!didHit) {
        // addEnvironmentOrSky, device_traversal.cuh:577-587.
        if (Full && s.environment.texels) {
            float weight = path.depth > 0 && !path.specularBounce
                ? powerHeuristic(path.previousBsdfPdf, environmentPdf(s.environment, path.ray.direction))
                : 1.0f;
            addPathRadiance(path, hadamard(path.throughput, environmentRadiance(s.environment, path.ray.direction)) *
                weight, sp.clamp, true);
        } else if (!s.blackBackground) {
            float t = 0.5f * (path.ray.direction.y + 1.0f);
            Vec3 sky = (1.0f - t) * Vec3(1.0f, 1.0f, 1.0f) + t * Vec3(0.5f, 0.7f, 1.0f);
            addPathRadiance(path, hadamard(path.throughput, sky), sp.clamp, true);
        }
        path.active = 0;
        return false;
    }
    SurfaceHit hit = surfaceAttributes<Full>(s, path.ray, h);
    const Material& material = s.materials[hit.material];
    if (path.mediumMaterial != kInvalid) {
        const Material& medium = s.materials[path.mediumMaterial];
        if (medium.attenuationDistance > 0.0f) {
            float distance = hit.t * path.mediumDensity;
            Vec3 sigmaA(-logF(fmaxf(medium.attenuationColor.x, 0.0001f)) / medium.attenuationDistance,
                -logF(fmaxf(medium.attenuationColor.y, 0.0001f)) / medium.attenuationDistance,
                -logF(fmaxf(medium.attenuationColor.z, 0.0001f)) / medium.attenuationDistance);
            path.throughput = hadamard(path.throughput,
                Vec3(expF(-sigmaA.x * distance), expF(-sigmaA.y * distance), expF(-sigmaA.z * distance)));
        }
    }
    if (material.type == kEmissive) {
        float misWeight = 1.0f;
        if (hit.lightIndex != kInvalid) {
            const TriangleLight& light = s.lights[hit.lightIndex];
            float lightCosine = fmaxf(0.0f, dot(light.normal, -path.ray.direction));
            if (lightCosine == 0.0f) {
                misWeight = 0.0f;
            } else if (path.depth > 0 && !path.specularBounce) {
                float lightPdf = light.selectionPdf * hit.t * hit.t / (lightCosine * light.area);
                misWeight = powerHeuristic(path.previousBsdfPdf, lightPdf);
            }
        }
        addPathRadiance(path, hadamard(path.throughput, materialEmission<Full>(s, material, hit)) * misWeight, sp.clamp,
            true);
        path.active = 0;
        return false;
    }
    Vec3 shadingNormal = materialNormal<Full>(s, material, hit);
    if (dot(shadingNormal, hit.geometricNormal) < 0.0f) shadingNormal = -shadingNormal;
    Vec3 albedo = materialAlbedo<Full>(s, material, hit);
    uint32_t type =
        material.type == kDiffuse && materialMetallic<Full>(s, material, hit) > 0.5f ? kMetal : material.type;
    float roughness = materialRoughness<Full>(s, material, hit);
    Vec3 direction;
    if (Full && isGltf(s, material)) {
        GltfBsdf bsdf{gltfSurface(s, material, hit, shadingNormal, albedo, roughness, -path.ray.direction)};
        sampleDirectLight(s, path, hit, shadingNormal, bsdf, sp, occ, out.shadow);
        sampleDirectPunctualLight(s, path, hit, shadingNormal, bsdf, sp, occ);
        sampleDirectEnvironment(s, path, hit, shadingNormal, bsdf, sp, occ);
        direction = continueGltf(s, path, hit, bsdf.g, sp, occ);
    } else switch (type) {
    case kDiffuse: {
        StandardBsdf<Full> bsdf{false, shadingNormal, SpecGlossSurface{albedo, Vec3(), roughness}};
        sampleDirectLight(s, path, hit, shadingNormal, bsdf, sp, occ, out.shadow);
        sampleDirectPunctualLight(s, path, hit, shadingNormal, bsdf, sp, occ);
        sampleDirectEnvironment(s, path, hit, shadingNormal, bsdf, sp, occ);
        if (Full && sp.photons) {  // device_path_shading.cuh:99-102
            if constexpr (Occluder::kDeferGather) {
                occ.recordGather(path, hit.position, shadingNormal, albedo, sp.clamp);
            } else {
                Vec3 caustic = gatherCaustics(*sp.photons, hit.position, shadingNormal);
                emitRadiance(occ, path, hadamard(path.throughput, hadamard(albedo, caustic)), sp.clamp, path.ray,
                    0.0f, false);
            }
        }
        direction = continueDiffuse<Full>(s, path, hit, shadingNormal, albedo, sp.rngStrategy);
        break;
    }
    case kSpecularGlossiness: {
        StandardBsdf<Full> bsdf{true, shadingNormal, specGlossSurface<Full>(s, material, hit, albedo, roughness)};
        sampleDirectLight(s, path, hit, shadingNormal, bsdf, sp, occ, out.shadow);
        sampleDirectPunctualLight(s, path, hit, shadingNormal, bsdf, sp, occ);
        sampleDirectEnvironment(s, path, hit, shadingNormal, bsdf, sp, occ);
        direction = continueGlossy(path, hit, shadingNormal, bsdf.sg, sp.rngStrategy);
        break;
    }
    case kMetal:
        if (roughness > 0.0f) {
            StandardBsdf<Full> bsdf{true, shadingNormal, SpecGlossSurface{Vec3(), albedo, roughness}};
            sampleDirectLight(s, path, hit, shadingNormal, bsdf, sp, occ, out.shadow);
            sampleDirectPunctualLight(s, path, hit, shadingNormal, bsdf, sp, occ);
            sampleDirectEnvironment(s, path, hit, shadingNormal, bsdf, sp, occ);
            direction = continueGlossy(path, hit, shadingNormal, bsdf.sg, sp.rngStrategy);
            break;
        }
        {
            path.throughput = hadamard(path.throughput, albedo);
            Vec3 reflected = reflect(normalize(path.ray.direction), shadingNormal);
            Vec3 jitter = randomInUnitSphere(path, sp.rngStrategy);
            direction = normalize(reflected + roughness * jitter);
            if (dot(direction, shadingNormal) <= 0.0f || dot(direction, hit.geometricNormal) <= 0.0f)
                direction = reflect(normalize(path.ray.direction), hit.geometricNormal);
            path.specularBounce = 1;
            path.previousBsdfPdf = 0.0f;
        }
        break;
    case kDielectric:
        direction = continueDielectric<Full>(s, path, hit, shadingNormal, material, albedo, sp, occ);
        break;
    default:
        break;
    }
    if (path.active) {
        path.ray.origin = offsetRayOrigin(hit.position, hit.geometricNormal, direction, hit.t);
        path.ray.direction = direction;
        ++path.depth;
        if (path.depth >= sp.maxDepth) {
            path.active = 0;
            return false;
        }
        if (path.depth >= sp.rrStart) {
            float survival = fminf(0.95f, fmaxf(0.05f, path.etaScale * fmaxf(path.throughput.x,
                fmaxf(path.throughput.y, path.throughput.z))));
            if (rand01(path, sp.rngStrategy) > survival) {
                path.active = 0;
                return false;
            }
            path.throughput *= 1.0f / survival;
        }
        return true;
    }
    return false;
}

template <bool Full = true, class Occluder>
WPT_HD bool shadePath(const DevScene& s, PathState& path, const HitRecord& h, const ShadeParams& sp, Occluder& occ,
    ShadeOutcome& out) {
    out.isDebug = false;
    out.shadow.valid = 0;
    bool alive = shadeBounce<Full>(s, path, h, sp, occ, out);
    out.alive = alive;
    return alive;
}

// Primary ray and fresh path state for one pixel (renderer.cu:25-41,
// camera.cuh:17-27).
struct CameraParams {
    Vec3 position, forward, right, up;
    float verticalFov;
    float aspectRatio;
};

WPT_HD PathState primaryPath(const CameraParams& cam, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
    uint32_t sample, uint32_t strategy) {
    PathState p;
    p.pixel = y * width + x;
    p.rng = rngSeed(pixelRngKey(p.pixel, sample, strategy), strategy);
    float u = (static_cast<float>(x) + rand01(p, strategy)) / static_cast<float>(width);
    float v = (static_cast<float>(y) + rand01(p, strategy)) / static_cast<float>(height);
    float tanHalfFov = tanF(cam.verticalFov * 0.5f);
    float px = (2.0f * u - 1.0f) * cam.aspectRatio * tanHalfFov;
    float py = (1.0f - 2.0f * v) * tanHalfFov;
    p.ray = Ray(cam.position, normalize(cam.forward + cam.right * px + cam.up * py));
    p.throughput = Vec3(1.0f, 1.0f, 1.0f);
    p.radiance = Vec3(0.0f, 0.0f, 0.0f);
    p.depth = 0;
    p.mediumMaterial = kInvalid;
    p.mediumDensity = 1.0f;
    p.previousBsdfPdf = 0.0f;
    p.specularBounce = 1;
    p.active = 1;
    p.etaScale = 1.0f;
    return p;
}

}  // namespace wpt
