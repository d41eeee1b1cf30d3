// Wavefront queues and kernel launch interface (host side of the kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "photon_trace.h"
#include "shade.h"

namespace wpt {

// Path queue in SoA layout: every field is an array of `capacity` words, so a
// wave's 64 consecutive slots are one coalesced 256 B access per field.
enum QueueField : uint32_t {
    kOx, kOy, kOz, kDx, kDy, kDz,  // ray
    kTx, kTy, kTz,                 // throughput
    kRx, kRy, kRz,                 // radiance
    kPixel,
    kDepthFlags,                   // depth | ended << 30 | specularBounce << 31
    kRngLo, kRngHi,
    kMedium, kMediumDensity, kPrevPdf, kEtaScale,
    kQueueFieldCount
};

// Deferred shadow rays of one bounce (SoA).  target is the continuing path's
// slot in the next queue, or kShadowTerminated for a path that ended this
// bounce, whose pre-NEE radiance and pixel travel with the request.
enum ShadowField : uint32_t {
    kSOx, kSOy, kSOz, kSDx, kSDy, kSDz, kSTmax,
    kSCx, kSCy, kSCz,
    kSTarget, kSPixel, kSRx, kSRy, kSRz,
    kShadowFieldCount
};
constexpr uint32_t kShadowTerminated = 0xffffffffu;

// A queue is split into kSegments segments of equal capacity.  Workgroup b
// appends to segment b % kSegments, which is the XCD it runs on under the
// round-robin dispatch of MI355X, so concurrent appends from different XCDs
// hit different counters.  Consumers enumerate the segments' items in order.
constexpr uint32_t kSegments = 8;

struct PathQueue {
    uint32_t* words;           // kQueueFieldCount * capacity
    uint32_t capacity;         // kSegments * segmentCapacity
    uint32_t segmentCapacity;
};

struct ShadowQueue {
    uint32_t* words;           // kShadowFieldCount * capacity
    uint32_t capacity;
    uint32_t segmentCapacity;
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

// A tile and, for the staged and persistent schedulers, the number of
// consecutive samples one pass renders together (1 .. LaunchContext.slots).
struct RenderTile {
    uint32_t x, y, width, height, advanceSample;
    uint32_t samples;
};

// Paths of one pass carry their sample slot in the pixel word's top bits.
constexpr uint32_t kSlotShift = 26;
constexpr uint32_t kPixelMask = (1u << kSlotShift) - 1u;
constexpr uint32_t kMaxSlots = 64;  // slot < 64: six bits above the 26 pixel bits

// Part of a batch task rendered by one staged batch pass: `samples`
// consecutive samples from `sampleStart` of every pixel of the task's tile.
// The pass's paths are enumerated piece by piece (pathBase), and a path's
// radiance lands at its index in the pass; the resolve then adds each tile
// pixel's samples, in order, to the task's output.
struct PassPiece {
    RenderTile tile;
    uint32_t sampleStart, samples;
    uint32_t pathBase;        // first path of the piece in the pass
    uint32_t pixelBase;       // first tile pixel of the piece among the pass's pieces
    uint32_t outStride;       // pixels per output row
    uint32_t pad;
    unsigned long long out;   // float offset of the tile's first pixel in the output
};

// Device counters shared by the launches of one renderer.
struct WavefrontCounters {
    uint32_t count[2][kSegments];    // staged mode: items per segment of queue 0 / 1
    uint32_t shadowCount[kSegments];  // shadow requests (event rays) per segment of the current bounce
    uint32_t gatherCount[kSegments];  // deferred caustic gathers per segment of the current bounce
    uint32_t pcount[3][kSegments];   // persistent kernel: queues rotate through three count sets
    uint32_t sampleIndex;
    uint32_t pad;
    unsigned long long closestQueries;
    unsigned long long shadowQueries;
    uint32_t queueCounts[64];
};

// Traversal stack entries kept in LDS per lane; deeper entries spill to a
// per-thread column of `spill`.
#ifndef WPT_LDS_STACK_ENTRIES
#define WPT_LDS_STACK_ENTRIES 16
#endif
constexpr uint32_t kLdsStackEntries = WPT_LDS_STACK_ENTRIES;
constexpr uint32_t kSpillEntriesTotal = kMaxStackDepth - kLdsStackEntries;  // per thread

struct LaunchContext {
    DevScene scene;
    CameraParams camera;
    ShadeParams shade;
    PathQueue queue[2];
    ShadowQueue shadow;
    HitBuffer hits;
    // Staged scenes whose shading adds radiance after an occlusion test
    // (environment, punctual lights, caustics): each shaded path's ordered
    // radiance events, indexed by its slot in the queue being shaded
    // (words null when the scene does not need them).
    EventBuffer events;
    // Staged batch pass (pieces non-null): the paths of `pieceCount` pieces
    // of one scene's tasks (same camera, frame and path parameters), `passPaths`
    // in all, `passPixels` tile pixels; radiance per path in slotRadiance,
    // resolved into `batchOut`.
    const PassPiece* pieces;
    uint32_t pieceCount, passPaths, passPixels;
    float* batchOut;
    float* framebuffer;      // width*height*3
    // Staged and persistent passes render `slots` samples of a tile at once:
    // a finished path stores its radiance in its sample's plane of
    // slotRadiance (slots x width*height*3), and the pass ends by adding the
    // planes to the framebuffer in sample order.
    float* slotRadiance;
    uint32_t slots;
    uint32_t width, height;
    WavefrontCounters* counters;
    const RenderTile* tile;  // device copy of the current tile
    uint2* spill;            // (kMaxStackDepth - kLdsStackEntries) x spillThreads, one column per thread index
    uint32_t spillThreads;
    uint32_t capacity;         // tile pixels x slots: the most items any queue holds
    uint32_t blockSize;
    uint32_t gridBlocks;       // staged kernels' grid (grid stride over their items)
    uint32_t resolveBlocks;    // ceil(tile pixels / blockSize)
    uint32_t persistentBlocks;
    uint32_t profileQueues;
};

// LDS bytes of the traversal stack for a workgroup of `blockSize` lanes.
inline size_t stackLdsBytes(uint32_t blockSize) { return sizeof(uint2) * kLdsStackEntries * blockSize; }

// Most waves per SIMD of the staged traversal kernels (0: as many as their
// registers and LDS stacks allow), so that their occupancy does not drift
// with their register count: a fifth wave never made them faster, and cost
// the closest-hit pass 10-14% where registers were squeezed for it.  The
// cap is held by padding each workgroup's LDS request, given the CU's
// 160 KiB.
#ifndef WPT_TRACE_MAX_WAVES
#define WPT_TRACE_MAX_WAVES 4
#endif
constexpr size_t kLdsPerCu = 160u * 1024u;
inline size_t traversalLdsBytes(uint32_t blockSize) {
    size_t stack = stackLdsBytes(blockSize);
    if (WPT_TRACE_MAX_WAVES <= 0) return stack;
    size_t waves = blockSize / 64u;                    // per workgroup
    size_t allowed = 4u * WPT_TRACE_MAX_WAVES / waves;  // workgroups per CU
    if (allowed == 0 || (kLdsPerCu / stack) <= allowed) return stack;
    return kLdsPerCu / allowed;
}

// Staged wavefront: records primary + maxDepth x (reset, trace, shade,
// shadow) + resolve + advance for one pass of the current tile (its
// `samples` samples) on `stream` (captured once into a hipGraph by the
// caller).
hipError_t enqueueStagedSample(const LaunchContext& ctx, hipStream_t stream);
// One staged batch pass (ctx.pieces set): primary, maxDepth x (reset, trace,
// shade, shadow), then the pieces' resolve into ctx.batchOut.
hipError_t enqueueStagedBatchPass(const LaunchContext& ctx, hipStream_t stream);
// Persistent cooperative kernel for one pass of the current tile, then
// resolve and advance.
hipError_t enqueuePersistentSample(const LaunchContext& ctx, hipStream_t stream);
// Resident workgroups per CU of each kernel at `blockSize`.
struct Occupancy {
    int trace, shade, shadow, persistent;  // resident workgroups per CU
};
// Occupancy of the kernels serving scenes with (full) or without
// environment maps, punctual lights and texture bindings.
// `events`: the full-feature cooperative kernel resolves radiance events
// (an event buffer exists), else it traces inline.
hipError_t kernelOccupancy(uint32_t blockSize, bool full, Occupancy* out, bool events = true);
// One pass of the current tile with every launch bracketed by events:
// ms/launches accumulate per kernel (0 primary, 1 reset, 2 trace, 3 shade,
// 4 persistent, 5 shadow, 6 resolve).  Synchronizes the stream.
constexpr int kProfileSlots = 7;
hipError_t profileSample(const LaunchContext& ctx, hipStream_t stream, bool persistent, double ms[kProfileSlots],
    uint32_t launches[kProfileSlots]);

// ---- batch consumer ---------------------------------------------------
// One task of a batch: a tile of one scene's frame and a sample range, with
// its own camera and path parameters.
struct BatchTask {
    CameraParams camera;
    ShadeParams shade;
    RenderTile tile;
    uint32_t scene;
    uint32_t width, height;  // frame: pixel ids and RNG keys
    uint32_t sampleStart, sampleCount;
    uint32_t pixelCount;     // tile pixels, 0 when the task renders nothing
    uint32_t outStride;      // pixels per output row
    unsigned long long out;  // float offset of the tile's first pixel in the output
    uint32_t accumulate;     // 1: add to the output's current sums, 0: overwrite
    uint32_t pad;
};

// Per-task progress of a running batch.
struct BatchTaskState {
    uint32_t cursor;  // next unclaimed tile pixel
    uint32_t done;    // finished pixels
    unsigned long long finishClock;
};

struct BatchCounters {
    unsigned long long startClock;
    unsigned long long closestQueries;
    unsigned long long shadowQueries;
};

struct BatchContext {
    const DevScene* scenes;
    const BatchTask* tasks;
    BatchTaskState* state;
    BatchCounters* counters;
    uint32_t taskCount;
    float* out;
    uint2* spill;
    uint32_t spillThreads;
    uint32_t full;  // some scene has an environment map, punctual lights or texture bindings
    // Full scenes: radiance events of each lane's last shading step, one
    // column per thread index of the launch (words null: trace inline).
    EventBuffer events;
};

// Caustic photon map (photon.hip, photon_trace.h): the renderer's buffers, and
// the build: trace, key, sort, cell starts.  `view` names the buffers the
// gather reads (sorted photons and cell starts) and the grid box.
struct PhotonMapBuffers {
    Photon* photons;  // emission order
    Photon* sorted;   // cell order
    uint64_t* keys;
    uint64_t* sortedKeys;
    void* temp;
    size_t tempBytes;
    uint32_t* cellStart;  // resolution^3 + 1 entries
    uint32_t* stored;    // photons kept inside the grid
    uint32_t* occupied;  // min x, y, z then max x, y, z of the non-empty cells
};

hipError_t photonSortBytes(uint32_t count, uint32_t cells, size_t* bytes);
hipError_t enqueuePhotonMap(const DevScene& scene, const PhotonSettings& settings, const PhotonMapBuffers& buffers,
    const PhotonGridView& view, uint2* spill, uint32_t spillThreads, hipStream_t stream);

// Resets the per-task state and counters, then runs the persistent batch
// consumer on `blocks` x `blockSize` threads.
hipError_t enqueueBatch(const BatchContext& ctx, uint32_t blocks, uint32_t blockSize, hipStream_t stream);
hipError_t batchOccupancy(uint32_t blockSize, bool full, int* blocksPerCu);
// Compiled resources (registers, scratch, static LDS) of the batch kernel and
// of the staged trace or cooperative persistent kernel.
hipError_t batchAttributes(bool full, hipFuncAttributes* out);
hipError_t kernelAttributes(bool persistent, bool full, hipFuncAttributes* out, bool events = true);

// Standalone traversal for the C ABI; `spill` holds
// (kMaxStackDepth - kLdsStackEntries) x abiTraceThreads() entries.
uint32_t abiTraceThreads();
hipError_t launchTraceClosest(const DevScene& scene, const float* rays6, uint32_t n, void* hitsOut, uint2* spill,
    hipStream_t stream);
hipError_t launchTraceOccluded(const DevScene& scene, const float* rays6, const float* tmax, uint32_t n,
    uint8_t* out, uint2* spill, hipStream_t stream);
// Order-preserving compaction for the C ABI; scratch holds ticket + flags.
hipError_t launchCompact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity,
    uint32_t* count, uint32_t* scratch, uint32_t scratchWords, hipStream_t stream);
uint32_t compactScratchWords(uint32_t n);

}  // namespace wpt
