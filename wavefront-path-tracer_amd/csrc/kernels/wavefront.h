// Wavefront queues and kernel launch interface (host side of the kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "shade.h"

namespace wpt {

// Path queue in SoA layout: every field is an array of `capacity` words, so a
// wave's 64 consecutive slots are one coalesced 256 B access per field.
enum QueueField : uint32_t {
    kOx, kOy, kOz, kDx, kDy, kDz,  // ray
    kTx, kTy, kTz,                 // throughput
    kRx, kRy, kRz,                 // radiance
    kPixel,
    kDepthFlags,                   // depth | specularBounce << 31
    kRngLo, kRngHi,
    kMedium, kMediumDensity, kPrevPdf, kEtaScale,
    kQueueFieldCount
};

struct PathQueue {
    uint32_t* words;  // kQueueFieldCount * capacity
    uint32_t capacity;
};

// Closest-hit records of a traced queue (SoA, staged mode only).
struct HitBuffer {
    float* t;
    float* u;
    float* v;
    uint32_t* kind;
    uint32_t* instance;
    uint32_t* prim;
    uint32_t* material;
};

struct RenderTile {
    uint32_t x, y, width, height, advanceSample;
};

// Device counters shared by the launches of one renderer.
struct WavefrontCounters {
    uint32_t count[2];       // live paths in queue 0 / 1
    uint32_t ticket;         // ordered chunk tickets of the running shade pass
    uint32_t sampleIndex;
    unsigned long long closestQueries;
    unsigned long long shadowQueries;
    uint32_t queueCounts[64];
};

struct LaunchContext {
    DevScene scene;
    CameraParams camera;
    ShadeParams shade;
    PathQueue queue[2];
    HitBuffer hits;
    float* framebuffer;      // width*height*3
    uint32_t width, height;
    uint32_t* chunkFlags;    // look-back state per chunk
    WavefrontCounters* counters;
    const RenderTile* tile;  // device copy of the current tile
    uint32_t blockSize;
    uint32_t gridBlocks;     // staged launches: ceil(capacity / blockSize)
    uint32_t persistentBlocks;
    uint32_t profileQueues;
};

// Staged wavefront: records primary + maxDepth x (reset, trace, shade) +
// advance on `stream` (captured once into a hipGraph by the caller).
hipError_t enqueueStagedSample(const LaunchContext& ctx, hipStream_t stream);
// Persistent cooperative kernel for one tile sample, then advance.
hipError_t enqueuePersistentSample(const LaunchContext& ctx, hipStream_t stream);
hipError_t persistentOccupancy(uint32_t blockSize, int* blocksPerCu);
// One sample of the current tile with every launch bracketed by events:
// ms/launches accumulate per kernel (0 primary, 1 reset, 2 trace, 3 shade,
// 4 persistent).  Synchronizes the stream.
hipError_t profileSample(const LaunchContext& ctx, hipStream_t stream, bool persistent, double ms[5],
    uint32_t launches[5]);

// Standalone traversal for the C ABI.
hipError_t launchTraceClosest(const DevScene& scene, const float* rays6, uint32_t n, void* hitsOut,
    hipStream_t stream);
hipError_t launchTraceOccluded(const DevScene& scene, const float* rays6, const float* tmax, uint32_t n,
    uint8_t* out, hipStream_t stream);
// Order-preserving compaction for the C ABI; scratch holds ticket + flags.
hipError_t launchCompact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity,
    uint32_t* count, uint32_t* scratch, uint32_t scratchWords, hipStream_t stream);
uint32_t compactScratchWords(uint32_t n);

}  // namespace wpt
