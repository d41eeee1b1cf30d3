// Caustic photon map: emission from the triangle lights, tracing through
// mirrors and dielectrics to the first diffuse surface, and the grid the
// gather reads.  Reference: src/renderer/photon_mapping.cu:4-94,
// render_driver.cu:180-234.
//
// One thread traces one photon (grid-stride over a resident grid, so the
// traversal stack's spill area is bounded).  The grid is a sort of
// (cell, photon index) keys, which makes each cell a contiguous, ordered
// range; the reference's per-cell linked lists are filled by atomic
// exchanges in a run-dependent order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "device_queue.h"
#include "wavefront.h"

namespace wpt {

namespace {

__global__ void __launch_bounds__(256) kPhotonTrace(DevScene scene, PhotonSettings ps, Photon* photons, uint2* spill,
    uint32_t spillThreads) {
    LdsStack stack = makeStack(spill, spillThreads);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ps.count; i += gridDim.x * blockDim.x) {
        Photon p;
        p.valid = 0;
        p.index = i;
        p.pad = 0;
        p.position = p.direction = p.flux = Vec3();
        if (scene.lightCount == 0) {
            photons[i] = p;
            continue;
        }
        // emitPhotons (photon_mapping.cu:4-30).
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
            float channelFlux = (channel == 0 ? emitter.emission.x : channel == 1 ? emitter.emission.y
                                                                                  : emitter.emission.z) *
                (light.area * 3.14159265359f * 3.0f / static_cast<float>(ps.count));
            flux = Vec3(channel == 0 ? channelFlux : 0.0f, channel == 1 ? channelFlux : 0.0f,
                channel == 2 ? channelFlux : 0.0f);
        }
        // tracePhotons (photon_mapping.cu:33-74).
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
                float dispersion = m.ext != kInvalid ? scene.materialExt[m.ext].dispersion : 0.0f;
                if (dispersion > 0.0f && channel < 3) ior += (static_cast<float>(channel) - 1.0f) * dispersion;
                DielectricInterface b = dielectricInterface(normalize(ray.direction), hit.geometricNormal, hit.normal,
                    ior);
                bool reflected = b.reflectance >= 1.0f || b.reflectance > rngNextFloat(rng, ps.strategy);
                next = reflected ? b.reflection : b.transmission;
            } else {
                next = cosineHemisphere(normal, rng, ps.strategy);
            }
            if (m.type == kMetal || m.type == kDielectric) ++specularBounces;
            flux = hadamard(flux, m.albedo);
            ray = Ray(offsetRayOrigin(hit.position, hit.geometricNormal, next, hit.t), next);
        }
        photons[i] = p;
    }
}

// Sort keys: (cell << 32) | photon, with photons outside the grid or not
// stored keyed past the last cell.
__global__ void kPhotonKeys(PhotonGridView g, PhotonSettings ps, const Photon* photons, uint64_t* keys,
    uint32_t* stored) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.count) return;
    const Photon& p = photons[i];
    uint32_t cells = g.resolution * g.resolution * g.resolution;
    uint32_t cell = cells;
    if (p.valid) {
        int x, y, z;
        bool inside;
        photonCell(g, p.position, x, y, z, inside);
        if (inside) {
            cell = (static_cast<uint32_t>(z) * g.resolution + static_cast<uint32_t>(y)) * g.resolution +
                static_cast<uint32_t>(x);
            atomicAdd(stored, 1u);
        }
    }
    keys[i] = (static_cast<uint64_t>(cell) << 32) | i;
}

__global__ void kFillWords(uint32_t* p, uint32_t n, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

// Cell ranges of the sorted keys, and the photons in sorted order.
__global__ void kPhotonRanges(PhotonSettings ps, uint32_t cells, const uint64_t* sortedKeys, const Photon* photons,
    Photon* sorted, uint32_t* cellStart, uint32_t* cellEnd) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= ps.count) return;
    uint64_t key = sortedKeys[p];
    uint32_t cell = static_cast<uint32_t>(key >> 32);
    sorted[p] = photons[static_cast<uint32_t>(key)];
    if (cell >= cells) return;
    if (p == 0 || static_cast<uint32_t>(sortedKeys[p - 1] >> 32) != cell) cellStart[cell] = p;
    if (p + 1 == ps.count || static_cast<uint32_t>(sortedKeys[p + 1] >> 32) != cell) cellEnd[cell] = p + 1;
}

uint32_t keyBits(uint32_t cells) {
    uint32_t bits = 0;
    while ((1ull << bits) <= cells) ++bits;  // cells itself marks "not stored"
    return 32 + bits;
}

}  // namespace

hipError_t photonSortBytes(uint32_t count, uint32_t cells, size_t* bytes) {
    *bytes = 0;
    return hipcub::DeviceRadixSort::SortKeys(nullptr, *bytes, static_cast<const uint64_t*>(nullptr),
        static_cast<uint64_t*>(nullptr), static_cast<int>(count), 0, static_cast<int>(keyBits(cells)));
}

hipError_t enqueuePhotonMap(const DevScene& scene, const PhotonSettings& ps, const PhotonMapBuffers& b,
    const PhotonGridView& view, uint2* spill, uint32_t spillThreads, hipStream_t stream) {
    uint32_t cells = view.resolution * view.resolution * view.resolution;
    uint32_t blocks = (ps.count + 255) / 256;
    uint32_t traceBlocks = blocks < spillThreads / 256 ? blocks : spillThreads / 256;
    if (traceBlocks == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kPhotonTrace, dim3(traceBlocks), dim3(256), stackLdsBytes(256), stream, scene, ps, b.photons,
        spill, spillThreads);
    hipLaunchKernelGGL(kFillWords, dim3(1), dim3(64), 0, stream, b.stored, 1u, 0u);
    hipLaunchKernelGGL(kPhotonKeys, dim3(blocks), dim3(256), 0, stream, view, ps, b.photons, b.keys, b.stored);
    size_t bytes = b.tempBytes;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(b.temp, bytes, b.keys, b.sortedKeys, static_cast<int>(ps.count),
        0, static_cast<int>(keyBits(cells)), stream);
    if (e != hipSuccess) return e;
    uint32_t fillBlocks = (cells + 255) / 256 < 4096 ? (cells + 255) / 256 : 4096;
    hipLaunchKernelGGL(kFillWords, dim3(fillBlocks), dim3(256), 0, stream, b.cellStart, cells, 0u);
    hipLaunchKernelGGL(kFillWords, dim3(fillBlocks), dim3(256), 0, stream, b.cellEnd, cells, 0u);
    hipLaunchKernelGGL(kPhotonRanges, dim3(blocks), dim3(256), 0, stream, ps, cells, b.sortedKeys, b.photons,
        b.sorted, b.cellStart, b.cellEnd);
    return hipGetLastError();
}

}  // namespace wpt
