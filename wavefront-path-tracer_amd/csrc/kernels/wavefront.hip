// gfx950 wavefront kernels.
//
// Reference kernels and their MI355X counterparts:
//   generatePrimaryRays (src/renderer/renderer.cu:10-42) ...... kPrimary
//   intersectScene (device_traversal.cuh:1186-1194) ........... kTrace
//   shadePaths (device_path_shading.cuh:6-173) ................ kShade
//   isOccluded calls made inside shadePaths ................... kShadow
//   appendWarpCompacted (queue_compaction.cuh:8-29) ........... blockAppend in
//                                                               kShade / kPersistent,
//                                                               kCompactAbi (C ABI)
//   advanceSampleIndex (renderer.cu:44-48) .................... kAdvance
//   persistentWavefrontTrace (renderer.cu:55-125) ............. kPersistent
//
// Traversal stacks live in LDS (kLdsStackEntries per lane, lane-interleaved
// so a wave's accesses at equal depth hit distinct banks); deeper entries
// spill to a per-thread column in global memory.
//
// Work distribution uses no shared cursor: trace, shade and shadow kernels
// take one item per thread over a grid sized to the queue's capacity, and
// the hardware dispatcher hands out workgroups as earlier ones finish.
// (Device-scope atomics on one address serialize across the 8 XCDs at tens
// of nanoseconds each; a claim per wave cost more than the traversal.)
// Continuations are appended with one atomic per workgroup on the counter of
// the workgroup's queue segment (blockIdx % 8, its XCD): a 64-bit ballot
// ranks a wave's survivors, mbcnt gives each lane its offset and the
// workgroup adds per-wave offsets from LDS.  Queue order only affects
// coherence, never results: a pixel receives one path per sample and each
// path's arithmetic depends on its own state alone.
//
// The C ABI's compaction (kCompactAbi) is order preserving instead, so its
// output is a stable filter of the input: workgroups take chunk tickets in
// order and chain their totals with a decoupled look-back (one flag word per
// chunk).  A workgroup only waits for chunks whose tickets were taken earlier
// by workgroups that are already running, so every wait ends.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <vector>

#include "../../../include/wpt.h"
#include "device_queue.h"
#include "wavefront.h"

namespace wpt {

namespace cg = cooperative_groups;

namespace {

constexpr uint32_t kValueMask = (1u << 30) - 1u;
constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kAbiBlock = 128;
constexpr uint32_t kAbiBlocks = 2048;

struct CompactShared {
    uint32_t waveOffset[kMaxWaves];
    uint32_t base;
    uint32_t total;
    uint32_t chunk;
    uint32_t shadow;
};

// Slot of each surviving lane in the ordered output; every thread of the
// workgroup must call it.  Returns the workgroup's inclusive prefix.
__device__ uint32_t orderedSlots(bool alive, uint32_t chunk, uint32_t* flags, CompactShared& sh, uint32_t& slot) {
    uint32_t lane = __lane_id();
    uint32_t wave = threadIdx.x >> 6;
    unsigned long long mask = __ballot(alive);
    uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
    if (lane == 0) sh.waveOffset[wave] = static_cast<uint32_t>(__popcll(mask));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t waves = blockDim.x >> 6;
        uint32_t sum = 0;
        for (uint32_t w = 0; w < waves; ++w) {
            uint32_t c = sh.waveOffset[w];
            sh.waveOffset[w] = sum;
            sum += c;
        }
        uint32_t exclusive = 0;
        if (chunk == 0) {
            storeAgent(&flags[0], kFlagPrefix | sum);
        } else {
            storeAgent(&flags[chunk], kFlagAggregate | sum);
            uint32_t p = chunk - 1;
            while (true) {
                uint32_t f = loadAgent(&flags[p]);
                uint32_t state = f & ~kValueMask;
                if (state == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                exclusive += f & kValueMask;
                if (state == kFlagPrefix) break;
                --p;
            }
            storeAgent(&flags[chunk], kFlagPrefix | (exclusive + sum));
        }
        sh.base = exclusive;
        sh.total = sum;
    }
    __syncthreads();
    slot = alive ? sh.base + sh.waveOffset[wave] + rank : kInvalid;
    return sh.base + sh.total;
}

// A queued path whose continuation an escaping event ended after it was
// appended (event shading): traced with its bounce, then skipped.
constexpr uint32_t kEndedBit = 0x40000000u;
constexpr uint32_t kDepthMask = 0x3fffffffu;

__device__ __forceinline__ uint32_t* field(const PathQueue& q, uint32_t f) {
    return q.words + static_cast<size_t>(f) * q.capacity;
}
__device__ __forceinline__ float qf(const PathQueue& q, uint32_t f, uint32_t i) { return __uint_as_float(field(q, f)[i]); }
__device__ __forceinline__ uint32_t qu(const PathQueue& q, uint32_t f, uint32_t i) { return field(q, f)[i]; }
__device__ __forceinline__ void sf(const PathQueue& q, uint32_t f, uint32_t i, float v) { field(q, f)[i] = __float_as_uint(v); }
__device__ __forceinline__ void su(const PathQueue& q, uint32_t f, uint32_t i, uint32_t v) { field(q, f)[i] = v; }

__device__ __forceinline__ uint32_t* sfield(const ShadowQueue& q, uint32_t f) {
    return q.words + static_cast<size_t>(f) * q.capacity;
}

__device__ __forceinline__ Ray loadRay(const PathQueue& q, uint32_t i) {
    return Ray(Vec3(qf(q, kOx, i), qf(q, kOy, i), qf(q, kOz, i)), Vec3(qf(q, kDx, i), qf(q, kDy, i), qf(q, kDz, i)));
}

__device__ __forceinline__ PathState loadPath(const PathQueue& q, uint32_t i) {
    PathState p;
    p.ray = loadRay(q, i);
    p.throughput = Vec3(qf(q, kTx, i), qf(q, kTy, i), qf(q, kTz, i));
    p.radiance = Vec3(qf(q, kRx, i), qf(q, kRy, i), qf(q, kRz, i));
    p.pixel = qu(q, kPixel, i);
    uint32_t df = qu(q, kDepthFlags, i);
    p.depth = df & kDepthMask;
    p.specularBounce = df >> 31;
    p.rng = static_cast<uint64_t>(qu(q, kRngLo, i)) | (static_cast<uint64_t>(qu(q, kRngHi, i)) << 32);
    p.mediumMaterial = qu(q, kMedium, i);
    p.mediumDensity = qf(q, kMediumDensity, i);
    p.previousBsdfPdf = qf(q, kPrevPdf, i);
    p.etaScale = qf(q, kEtaScale, i);
    p.active = (df & kEndedBit) ? 0u : 1u;
    return p;
}

__device__ __forceinline__ void storePath(const PathQueue& q, uint32_t i, const PathState& p) {
    sf(q, kOx, i, p.ray.origin.x);
    sf(q, kOy, i, p.ray.origin.y);
    sf(q, kOz, i, p.ray.origin.z);
    sf(q, kDx, i, p.ray.direction.x);
    sf(q, kDy, i, p.ray.direction.y);
    sf(q, kDz, i, p.ray.direction.z);
    sf(q, kTx, i, p.throughput.x);
    sf(q, kTy, i, p.throughput.y);
    sf(q, kTz, i, p.throughput.z);
    sf(q, kRx, i, p.radiance.x);
    sf(q, kRy, i, p.radiance.y);
    sf(q, kRz, i, p.radiance.z);
    su(q, kPixel, i, p.pixel);
    su(q, kDepthFlags, i, (p.depth & kDepthMask) | (p.specularBounce ? 0x80000000u : 0u));
    su(q, kRngLo, i, static_cast<uint32_t>(p.rng));
    su(q, kRngHi, i, static_cast<uint32_t>(p.rng >> 32));
    su(q, kMedium, i, p.mediumMaterial);
    sf(q, kMediumDensity, i, p.mediumDensity);
    sf(q, kPrevPdf, i, p.previousBsdfPdf);
    sf(q, kEtaScale, i, p.etaScale);
}

// Paths of a pass: tile.samples consecutive samples of every tile pixel (or
// the batch pass's paths).
__device__ __forceinline__ uint32_t passPaths(const LaunchContext& ctx, const RenderTile& tile) {
    uint32_t count = ctx.pieces ? ctx.passPaths : tile.width * tile.height * tile.samples;
    return count < ctx.capacity ? count : ctx.capacity;
}

// The piece whose range [base, next base) holds `index` (base = pathBase or
// pixelBase of consecutive pieces, the first at 0).
template <bool Pixels>
__device__ __forceinline__ uint32_t findPiece(const LaunchContext& ctx, uint32_t index) {
    uint32_t lo = 0, hi = ctx.pieceCount;  // invariant: base[lo] <= index < base[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t base = Pixels ? ctx.pieces[mid].pixelBase : ctx.pieces[mid].pathBase;
        if (base <= index)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Fresh path g of a batch pass: sample (g - pathBase) / pixels of its piece's
// tile pixel; its pixel word is g, the radiance slot of the pass.
__device__ __forceinline__ PathState batchPassPath(const LaunchContext& ctx, uint32_t g) {
    const PassPiece& piece = ctx.pieces[findPiece<false>(ctx, g)];
    uint32_t pixels = piece.tile.width * piece.tile.height;
    uint32_t local = g - piece.pathBase;
    uint32_t slot = local / pixels;
    uint32_t x, y;
    tilePixel(piece.tile, local - slot * pixels, x, y);
    PathState p = primaryPath(ctx.camera, x, y, ctx.width, ctx.height, piece.sampleStart + slot, ctx.shade.rngStrategy);
    p.pixel = g;
    return p;
}

// Fresh path g of a pass: sample slot g / pixels of tile pixel g % pixels.
__device__ __forceinline__ PathState passPath(const LaunchContext& ctx, const RenderTile& tile, uint32_t sample,
    uint32_t g) {
    uint32_t pixels = tile.width * tile.height;
    uint32_t slot = g / pixels;
    uint32_t x, y;
    tilePixel(tile, g - slot * pixels, x, y);
    PathState p = primaryPath(ctx.camera, x, y, ctx.width, ctx.height, sample + slot, ctx.shade.rngStrategy);
    p.pixel |= slot << kSlotShift;
    return p;
}

// A finished path's radiance, stored in its sample's plane (batch pass: at
// the path's index).
__device__ __forceinline__ void finishPath(const LaunchContext& ctx, uint32_t tagged, const Vec3& v) {
    size_t i = ctx.pieces ? 3 * static_cast<size_t>(tagged)
                          : 3 * (static_cast<size_t>(tagged >> kSlotShift) * ctx.width * ctx.height + (tagged & kPixelMask));
    ctx.slotRadiance[i] = v.x;
    ctx.slotRadiance[i + 1] = v.y;
    ctx.slotRadiance[i + 2] = v.z;
}

__global__ void __launch_bounds__(512) kPrimary(LaunchContext ctx) {
    RenderTile tile = ctx.pieces ? RenderTile{} : *ctx.tile;
    uint32_t count = passPaths(ctx, tile);
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < kSegments) ctx.counters->count[0][gid] = primarySegmentCount(count, ctx.blockSize, gid);
    uint32_t sample = ctx.counters->sampleIndex;
    for (uint32_t g = gid; g < count; g += gridDim.x * blockDim.x)
        storePath(ctx.queue[0], primarySlot(g, ctx.blockSize, ctx.queue[0].segmentCapacity),
            ctx.pieces ? batchPassPath(ctx, g) : passPath(ctx, tile, sample, g));
}

// End of a batch pass: each piece's tile pixel adds its samples' radiance,
// in sample order, to the task's output.
__global__ void __launch_bounds__(512) kResolvePieces(LaunchContext ctx) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ctx.passPixels) return;
    const PassPiece& piece = ctx.pieces[findPiece<true>(ctx, t)];
    uint32_t pixels = piece.tile.width * piece.tile.height;
    uint32_t p = t - piece.pixelBase;
    uint32_t x, y;
    tilePixel(piece.tile, p, x, y);
    float* o = ctx.batchOut + piece.out + 3ull * ((y - piece.tile.y) * piece.outStride + (x - piece.tile.x));
    Vec3 sum(o[0], o[1], o[2]);
    for (uint32_t k = 0; k < piece.samples; ++k) {
        const float* r = ctx.slotRadiance + 3 * (static_cast<size_t>(piece.pathBase) + static_cast<size_t>(k) * pixels + p);
        sum += Vec3(r[0], r[1], r[2]);
    }
    o[0] = sum.x;
    o[1] = sum.y;
    o[2] = sum.z;
}

// End of a pass: each tile pixel adds its samples' radiance in sample order
// (the additions a one-sample-per-pass schedule makes), then the sample
// counter advances past the pass.
__global__ void __launch_bounds__(512) kResolve(LaunchContext ctx) {
    RenderTile tile = *ctx.tile;
    uint32_t pixels = tile.width * tile.height;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= pixels) return;
    uint32_t x, y;
    tilePixel(tile, g, x, y);
    size_t pixel = static_cast<size_t>(y) * ctx.width + x;
    size_t plane = 3 * static_cast<size_t>(ctx.width) * ctx.height;
    float* fb = ctx.framebuffer + 3 * pixel;
    Vec3 sum(fb[0], fb[1], fb[2]);
    for (uint32_t k = 0; k < tile.samples; ++k) {
        const float* r = ctx.slotRadiance + k * plane + 3 * pixel;
        sum += Vec3(r[0], r[1], r[2]);
    }
    fb[0] = sum.x;
    fb[1] = sum.y;
    fb[2] = sum.z;
}

__global__ void kReset(LaunchContext ctx, uint32_t nextSel) {
    uint32_t s = threadIdx.x;
    if (s < kSegments) {
        ctx.counters->count[nextSel][s] = 0;
        ctx.counters->shadowCount[s] = 0;
        ctx.counters->gatherCount[s] = 0;
    }
}

// Lanes of a wave that wait before they start new rays together (1: each
// starts its next ray as soon as the last one ends; 64: the wave's rays
// start and end together).
#ifndef WPT_TRACE_REFILL
#define WPT_TRACE_REFILL 16
#endif
constexpr int kTraceRefill = WPT_TRACE_REFILL;

// Rays in grid stride: the grid (at most a few times the resident
// workgroups) bounds the traversal spill area, and the dispatcher balances
// the workgroups.
template <bool Full>
__global__ void __launch_bounds__(512) WPT_TRACE_OCCUPANCY kTrace(LaunchContext ctx, uint32_t bounce, uint32_t sel) {
    SegmentView v = loadSegments(ctx.counters->count[sel]);
    uint32_t count = v.start[kSegments];
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) {
        atomicAdd(&ctx.counters->closestQueries, static_cast<unsigned long long>(count));
        if (ctx.profileQueues && bounce < 64) ctx.counters->queueCounts[bounce] = count;
    }
    const PathQueue& q = ctx.queue[sel];
    const uint32_t stride = gridDim.x * blockDim.x;
    LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
    if (bounce == 0) {  // camera rays: coherent, their walks end nearly together
        for (uint32_t g = gid; g < count; g += stride) {
            uint32_t i = segmentSlot(v, q.segmentCapacity, g);
            HitRecord h = traceClosest<Full>(ctx.scene, loadRay(q, i), stack);
            ctx.hits.t[i] = h.t;
            ctx.hits.u[i] = h.u;
            ctx.hits.v[i] = h.v;
            ctx.hits.kind[i] = h.kind;
            ctx.hits.instance[i] = h.instance;
            ctx.hits.prim[i] = h.prim;
            ctx.hits.material[i] = h.material;
        }
        return;
    }
    // One loop over node visits, not one per ray: lanes whose rays ended
    // store their hits and start their next rays (prefetched) while the
    // other lanes of the wave walk on, once at least kTraceRefill lanes (or
    // all that are left) wait, instead of idling until every ray of the
    // wave has ended.
    TlasWalk<Full, Query::kClosest> w;
    Walk state = Walk::kDone;
    bool have = false;  // the lane holds a ray
    uint32_t g = gid, i = 0;
    uint32_t nextSlot = g < count ? segmentSlot(v, q.segmentCapacity, g) : 0u;
    Ray next = g < count ? loadRay(q, nextSlot) : Ray();
    while (true) {
        bool walking = state == Walk::kTlas;
        unsigned long long waiting = __ballot(!walking), left = __ballot(true);
        if (walking || (__popcll(waiting) < kTraceRefill && waiting != left)) {
            if (walking) state = w.step(ctx.scene, stack);
            continue;
        }
        if (have) {  // the ray ended: its hit
            if (state == Walk::kAfter) w.finish(ctx.scene, stack);
            ctx.hits.t[i] = w.best.t;
            ctx.hits.u[i] = w.best.u;
            ctx.hits.v[i] = w.best.v;
            ctx.hits.kind[i] = w.best.kind;
            ctx.hits.instance[i] = w.best.instance;
            ctx.hits.prim[i] = w.best.prim;
            ctx.hits.material[i] = w.best.material;
            g += stride;
        }
        if (g >= count) break;
        i = nextSlot;
        Ray ray = next;
        if (g + stride < count) {
            nextSlot = segmentSlot(v, q.segmentCapacity, g + stride);
            next = loadRay(q, nextSlot);
        }
        state = w.start(ctx.scene, ray, kNoHit, false, stack);
        have = true;
    }
}

__device__ __forceinline__ HitRecord loadHit(const HitBuffer& hits, uint32_t i) {
    HitRecord h;
    h.t = hits.t[i];
    h.u = hits.u[i];
    h.v = hits.v[i];
    h.kind = hits.kind[i];
    h.instance = hits.instance[i];
    h.prim = hits.prim[i];
    h.material = hits.material[i];
    return h;
}

__device__ __forceinline__ void storeShadow(const ShadowQueue& q, uint32_t s, const ShadowRequest& r, uint32_t target,
    const PathState& path) {
    sfield(q, kSOx)[s] = __float_as_uint(r.origin.x);
    sfield(q, kSOy)[s] = __float_as_uint(r.origin.y);
    sfield(q, kSOz)[s] = __float_as_uint(r.origin.z);
    sfield(q, kSDx)[s] = __float_as_uint(r.direction.x);
    sfield(q, kSDy)[s] = __float_as_uint(r.direction.y);
    sfield(q, kSDz)[s] = __float_as_uint(r.direction.z);
    sfield(q, kSTmax)[s] = __float_as_uint(r.tmax);
    sfield(q, kSCx)[s] = __float_as_uint(r.contribution.x);
    sfield(q, kSCy)[s] = __float_as_uint(r.contribution.y);
    sfield(q, kSCz)[s] = __float_as_uint(r.contribution.z);
    sfield(q, kSTarget)[s] = target;
    sfield(q, kSPixel)[s] = path.pixel;
    sfield(q, kSRx)[s] = __float_as_uint(path.radiance.x);
    sfield(q, kSRy)[s] = __float_as_uint(path.radiance.y);
    sfield(q, kSRz)[s] = __float_as_uint(path.radiance.z);
}

// Staged shading in workgroup-uniform grid stride: continuing paths are
// appended to the workgroup's segment of the next queue, shadow requests to
// its segment of the shadow queue; finished paths without a pending shadow
// ray reach the framebuffer.  With `Inline` (scenes with an environment map
// or punctual lights, whose shading adds radiance after the area-light NEE)
// every shadow ray is traced here and the shadow queue stays empty.
template <bool Inline, bool Full>
__global__ void __launch_bounds__(512) kShade(LaunchContext ctx, uint32_t sel) {
    __shared__ AppendShared sh;
    SegmentView v = loadSegments(ctx.counters->count[sel]);
    uint32_t count = v.start[kSegments];
    // Queue descriptors by value: a dynamic index into the kernel argument
    // makes the compiler copy the whole argument to scratch.
    const PathQueue q = sel ? ctx.queue[1] : ctx.queue[0];
    const PathQueue next = sel ? ctx.queue[0] : ctx.queue[1];
    const uint32_t seg = blockIdx.x % kSegments;
    for (uint32_t first = blockIdx.x * blockDim.x; first < count; first += gridDim.x * blockDim.x) {
        uint32_t g = first + threadIdx.x;
        bool valid = g < count;
        PathState path;
        ShadeOutcome outcome;
        outcome.alive = false;
        outcome.shadow.valid = 0;
        if (valid) {
            uint32_t i = segmentSlot(v, q.segmentCapacity, g);
            path = loadPath(q, i);
            if constexpr (Inline) {
                LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
                InlineOcclusion<LdsStack> occ{stack, 0u};
                shadePath<Full>(ctx.scene, path, loadHit(ctx.hits, i), ctx.shade, occ, outcome);
                if (occ.queries) atomicAdd(&ctx.counters->shadowQueries, static_cast<unsigned long long>(occ.queries));
            } else {
                DeferOcclusion occ;
                shadePath<Full>(ctx.scene, path, loadHit(ctx.hits, i), ctx.shade, occ, outcome);
            }
        }
        bool alive = valid && outcome.alive;
        bool wantsShadow = valid && outcome.shadow.valid;
        uint32_t a, s;
        blockAppend(alive, wantsShadow, &ctx.counters->count[sel ^ 1u][seg], &ctx.counters->shadowCount[seg], sh, a, s);
        uint32_t slot = seg * next.segmentCapacity + a;
        if (alive) storePath(next, slot, path);
        if (wantsShadow) {
            storeShadow(ctx.shadow, seg * ctx.shadow.segmentCapacity + s, outcome.shadow,
                alive ? slot : kShadowTerminated, path);
        } else if (valid && !alive) {
            finishPath(ctx, path.pixel, outcome.isDebug ? outcome.debugColor : path.radiance);
        }
    }
}

__device__ __forceinline__ Vec3 eventContribution(const EventBuffer& b, uint32_t e, uint32_t slot) {
    return Vec3(__uint_as_float(*eventItemWord(b, e, kEiCx, slot)), __uint_as_float(*eventItemWord(b, e, kEiCy, slot)),
        __uint_as_float(*eventItemWord(b, e, kEiCz, slot)));
}

__device__ __forceinline__ Ray eventRay(const EventBuffer& b, uint32_t e, uint32_t slot) {
    return Ray(Vec3(__uint_as_float(*eventItemWord(b, e, kEiOx, slot)), __uint_as_float(*eventItemWord(b, e, kEiOy, slot)),
                   __uint_as_float(*eventItemWord(b, e, kEiOz, slot))),
        Vec3(__uint_as_float(*eventItemWord(b, e, kEiDx, slot)), __uint_as_float(*eventItemWord(b, e, kEiDy, slot)),
            __uint_as_float(*eventItemWord(b, e, kEiDz, slot))));
}

// Staged shading that traces no ray (scenes whose shading adds radiance
// after an occlusion test): each path's occlusion-dependent additions become
// ordered events at its slot (EventOcclusion), continuing paths are appended
// to the workgroup's segment of the next queue and the traced events to its
// segment of the event ray list; kShadowEvents then traces the list and
// kResolveEvents completes the radiance.  Paths without events finish or
// continue here as in kShade.
template <bool Full>
__global__ void __launch_bounds__(512) WPT_SHADE_OCCUPANCY kShadeEvents(LaunchContext ctx, uint32_t sel) {
    __shared__ AppendShared sh;
    SegmentView v = loadSegments(ctx.counters->count[sel]);
    uint32_t count = v.start[kSegments];
    const PathQueue q = sel ? ctx.queue[1] : ctx.queue[0];
    const PathQueue next = sel ? ctx.queue[0] : ctx.queue[1];
    const uint32_t seg = blockIdx.x % kSegments;
    const uint32_t maxEvents = ctx.events.maxEvents;
    for (uint32_t first = blockIdx.x * blockDim.x; first < count; first += gridDim.x * blockDim.x) {
        uint32_t g = first + threadIdx.x;
        bool valid = g < count;
        uint32_t i = 0, events = 0, traced = 0, gather = 0;
        PathState path;
        ShadeOutcome outcome;
        outcome.alive = false;
        outcome.isDebug = false;
        if (valid) {
            i = segmentSlot(v, q.segmentCapacity, g);
            path = loadPath(q, i);
            if (path.active) {
                StagedEventOcclusion occ{{ctx.events, i, 0u, 0u}, 0u};
                shadePath<Full>(ctx.scene, path, loadHit(ctx.hits, i), ctx.shade, occ, outcome);
                events = occ.count;
                traced = occ.traced;
                gather = occ.gather;
            } else {
                valid = false;  // ended by an event of the previous bounce
            }
            *eventWord(ctx.events, kEvCount, i) = events;
        }
        bool alive = valid && outcome.alive;
        uint32_t a, at, gi;
        blockAppendEvents(alive, traced, maxEvents, gather != 0u, &ctx.counters->count[sel ^ 1u][seg],
            &ctx.counters->shadowCount[seg], &ctx.counters->gatherCount[seg], sh, a, at, gi);
        uint32_t slot = seg * next.segmentCapacity + a;
        if (alive) storePath(next, slot, path);
        if (gather) {
            size_t k = static_cast<size_t>(seg) * ctx.events.gatherSegment + gi;
            ctx.events.gatherSlot[k] = i;
            ctx.events.gatherEvent[k] = static_cast<uint8_t>(gather - 1u);
        }
        if (events) {
            *eventWord(ctx.events, kEvTarget, i) = alive ? slot : kShadowTerminated;
            *eventWord(ctx.events, kEvPixel, i) = path.pixel;
            *eventWord(ctx.events, kEvRx, i) = __float_as_uint(path.radiance.x);
            *eventWord(ctx.events, kEvRy, i) = __float_as_uint(path.radiance.y);
            *eventWord(ctx.events, kEvRz, i) = __float_as_uint(path.radiance.z);
        } else if (valid && !alive) {
            finishPath(ctx, path.pixel, outcome.isDebug ? outcome.debugColor : path.radiance);
        }
        // The wave's event rays, grouped by event index: rays of one kind
        // (the area-light sample, the sky sample, ...) sit side by side.
        const size_t rayBase = static_cast<size_t>(seg) * ctx.events.raySegment;
        for (uint32_t e = 0; e < maxEvents; ++e) {
            bool has = (traced >> e) & 1u;
            unsigned long long m = __ballot(has);
            if (has) {
                uint32_t k = at + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
                ctx.events.raySlot[rayBase + k] = i;
                ctx.events.rayEvent[rayBase + k] = static_cast<uint8_t>(e);
            }
            at += static_cast<uint32_t>(__popcll(m));
        }
    }
}

// The bounce's deferred caustic gathers, one per thread: each fills in its
// event's contribution, throughput x (albedo x irradiance), clamped as
// shading clamps it (device_path_shading.cuh:99-102).
__global__ void __launch_bounds__(512) WPT_GATHER_OCCUPANCY kGatherEvents(LaunchContext ctx) {
    SegmentView v = loadSegments(ctx.counters->gatherCount);
    uint32_t count = v.start[kSegments];
    const EventBuffer& b = ctx.events;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
        uint32_t k = segmentSlot(v, b.gatherSegment, g);
        uint32_t i = b.gatherSlot[k], e = b.gatherEvent[k];
        Ray at = eventRay(b, e, i);  // query point and shading normal
        Vec3 albedo = eventContribution(b, e, i);
        float clamp = __uint_as_float(*eventItemWord(b, e, kEiTmax, i));
        Vec3 throughput(__uint_as_float(*eventWord(b, kEvGx, i)), __uint_as_float(*eventWord(b, kEvGy, i)),
            __uint_as_float(*eventWord(b, kEvGz, i)));
        Vec3 caustic = gatherCaustics(*ctx.shade.photons, at.origin, at.direction);
        Vec3 c = hadamard(throughput, hadamard(albedo, caustic));
        if (clamp > 0.0f) c = clampMagnitude(c, clamp);
        *eventItemWord(b, e, kEiCx, i) = __float_as_uint(c.x);
        *eventItemWord(b, e, kEiCy, i) = __float_as_uint(c.y);
        *eventItemWord(b, e, kEiCz, i) = __float_as_uint(c.z);
    }
}

// The bounce's event rays, one per thread: a blocked ray marks its event.
template <bool Full>
__global__ void __launch_bounds__(512) WPT_TRACE_OCCUPANCY kShadowEvents(LaunchContext ctx) {
    SegmentView v = loadSegments(ctx.counters->shadowCount);
    uint32_t count = v.start[kSegments];
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0 && count) atomicAdd(&ctx.counters->shadowQueries, static_cast<unsigned long long>(count));
    const uint32_t stride = gridDim.x * blockDim.x;
    const EventBuffer& b = ctx.events;
    LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
    auto entry = [&](uint32_t at, uint32_t& slot, uint32_t& event) {
        uint32_t k = segmentSlot(v, b.raySegment, at);
        slot = b.raySlot[k];
        event = b.rayEvent[k];
    };
    // The walks of a lane's rays interleave with the other lanes' as in
    // kTrace.  Prefetched: the next ray (its slot, event, ray and distance)
    // and the list entry after it.
    uint32_t i1 = 0, e1 = 0, i2 = 0, e2 = 0;
    Ray ray1;
    float tmax1 = 0.0f;
    if (gid < count) {
        entry(gid, i1, e1);
        ray1 = eventRay(b, e1, i1);
        tmax1 = __uint_as_float(*eventItemWord(b, e1, kEiTmax, i1));
    }
    if (gid + stride < count) entry(gid + stride, i2, e2);
    TlasWalk<Full, Query::kAny> w;
    Walk state = Walk::kDone;
    bool have = false;  // the lane holds a ray
    uint32_t g = gid, i = 0, e = 0;
    while (true) {
        bool walking = state == Walk::kTlas;
        unsigned long long waiting = __ballot(!walking), left = __ballot(true);
        if (walking || (__popcll(waiting) < kTraceRefill && waiting != left)) {
            if (walking) state = w.step(ctx.scene, stack);
            continue;
        }
        if (have) {  // the ray ended: a blocked ray marks its event
            if (state == Walk::kAfter) w.finish(ctx.scene, stack);
            if (w.best.kind != kHitNone) *eventItemWord(b, e, kEiFlags, i) |= kEventBlocked;
            g += stride;
        }
        if (g >= count) break;
        i = i1;
        e = e1;
        Ray ray = ray1;
        float tmax = tmax1;
        if (g + stride < count) {
            i1 = i2;
            e1 = e2;
            ray1 = eventRay(b, e1, i1);
            tmax1 = __uint_as_float(*eventItemWord(b, e1, kEiTmax, i1));
            if (g + 2u * stride < count) entry(g + 2u * stride, i2, e2);
        }
        state = w.start(ctx.scene, ray, tmax, true, stack);
        have = true;
    }
}

// The events of one shaded path, in order: the contributions whose ray was
// not blocked are added to the radiance the path had before its first
// event, and the result goes to the continuing path or, when the path ended
// (or an escaping event ends it), to its pixel.
__device__ __forceinline__ void resolveEvents(const LaunchContext& ctx, const SegmentView& v, uint32_t sel, uint32_t g) {
    const PathQueue& q = ctx.queue[sel];
    const PathQueue& next = ctx.queue[sel ^ 1u];
    uint32_t i = segmentSlot(v, q.segmentCapacity, g);
    uint32_t n = *eventWord(ctx.events, kEvCount, i);
    if (n == 0) return;
    if (n > ctx.events.maxEvents) n = ctx.events.maxEvents;
    Vec3 r(__uint_as_float(*eventWord(ctx.events, kEvRx, i)), __uint_as_float(*eventWord(ctx.events, kEvRy, i)),
        __uint_as_float(*eventWord(ctx.events, kEvRz, i)));
    bool ends = false;
    for (uint32_t e = 0; e < n; ++e) {
        uint32_t flags = *eventItemWord(ctx.events, e, kEiFlags, i);
        if (!(flags & kEventBlocked)) {
            r += eventContribution(ctx.events, e, i);
            if (flags & kEventTerminates) ends = true;
        }
    }
    uint32_t target = *eventWord(ctx.events, kEvTarget, i);
    if (target != kShadowTerminated && !ends) {
        sf(next, kRx, target, r.x);
        sf(next, kRy, target, r.y);
        sf(next, kRz, target, r.z);
    } else {
        finishPath(ctx, *eventWord(ctx.events, kEvPixel, i), r);
        if (target != kShadowTerminated) {
            su(next, kDepthFlags, target, qu(next, kDepthFlags, target) | kEndedBit);
            // The next bounce traces the ended path with the others; it is
            // not a query of the path tracer.
            atomicAdd(&ctx.counters->closestQueries, ~0ull);
        }
    }
}

__global__ void __launch_bounds__(512) kResolveEvents(LaunchContext ctx, uint32_t sel) {
    SegmentView v = loadSegments(ctx.counters->count[sel]);
    uint32_t count = v.start[kSegments];
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x)
        resolveEvents(ctx, v, sel, g);
}

// One shadow request: the contribution joins the continuing path's
// radiance, or the finished path's pixel, unless the ray is blocked.
template <bool Full>
__device__ __forceinline__ void resolveShadow(const LaunchContext& ctx, const SegmentView& v, uint32_t sel, uint32_t g) {
    LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
    const ShadowQueue& q = ctx.shadow;
    const PathQueue& next = ctx.queue[sel ^ 1u];
    uint32_t i = segmentSlot(v, q.segmentCapacity, g);
    Ray ray(Vec3(__uint_as_float(sfield(q, kSOx)[i]), __uint_as_float(sfield(q, kSOy)[i]),
                __uint_as_float(sfield(q, kSOz)[i])),
        Vec3(__uint_as_float(sfield(q, kSDx)[i]), __uint_as_float(sfield(q, kSDy)[i]),
            __uint_as_float(sfield(q, kSDz)[i])));
    bool occluded = traceOccluded<Full>(ctx.scene, ray, __uint_as_float(sfield(q, kSTmax)[i]), stack);
    Vec3 c(__uint_as_float(sfield(q, kSCx)[i]), __uint_as_float(sfield(q, kSCy)[i]),
        __uint_as_float(sfield(q, kSCz)[i]));
    uint32_t target = sfield(q, kSTarget)[i];
    if (target != kShadowTerminated) {
        if (!occluded) {
            Vec3 r(qf(next, kRx, target), qf(next, kRy, target), qf(next, kRz, target));
            r += c;
            sf(next, kRx, target, r.x);
            sf(next, kRy, target, r.y);
            sf(next, kRz, target, r.z);
        }
    } else {
        Vec3 r(__uint_as_float(sfield(q, kSRx)[i]), __uint_as_float(sfield(q, kSRy)[i]),
            __uint_as_float(sfield(q, kSRz)[i]));
        if (!occluded) r += c;
        finishPath(ctx, sfield(q, kSPixel)[i], r);
    }
}

// Resolves the bounce's shadow requests in grid stride: unoccluded
// contributions are added to the continuing path's radiance, and finished
// paths reach the pixel.
template <bool Full>
__global__ void __launch_bounds__(512) WPT_TRACE_OCCUPANCY kShadow(LaunchContext ctx, uint32_t sel) {
    SegmentView v = loadSegments(ctx.counters->shadowCount);
    uint32_t count = v.start[kSegments];
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid == 0) atomicAdd(&ctx.counters->shadowQueries, static_cast<unsigned long long>(count));
    for (uint32_t g = gid; g < count; g += gridDim.x * blockDim.x)
        resolveShadow<Full>(ctx, v, sel, g);
}

__global__ void kAdvance(LaunchContext ctx) {
    if (ctx.tile->advanceSample) ctx.counters->sampleIndex += ctx.tile->samples;
}

// Persistent kernel: one grid-wide barrier per bounce.  The resident grid
// walks the queue's items with a static stride (every workgroup runs the same
// number of rounds, so the workgroup-wide appends line up); each thread
// traces, shades and resolves its shadow ray inline, and continuations are
// appended to the workgroup's segment of the other queue.  Bounce b reads
// pcount[b % 3], appends to pcount[(b + 1) % 3] and clears pcount[(b + 2) % 3],
// which no workgroup touches again before the barrier that ends bounce b.
template <bool Full, bool Events>
__global__ void __launch_bounds__(512) WPT_PERSISTENT_OCCUPANCY(Full, Events) kPersistent(LaunchContext ctx) {
    __shared__ AppendShared sh;
    cg::grid_group grid = cg::this_grid();
    WavefrontCounters* counters = ctx.counters;
    RenderTile tile = *ctx.tile;
    uint32_t count0 = passPaths(ctx, tile);
    uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t stride = gridDim.x * blockDim.x;
    uint32_t sample = loadAgent(&counters->sampleIndex);
    for (uint32_t g = gid; g < count0; g += stride)
        storePath(ctx.queue[0], primarySlot(g, ctx.blockSize, ctx.queue[0].segmentCapacity),
            passPath(ctx, tile, sample, g));
    if (gid < kSegments) {
        storeAgent(&counters->pcount[0][gid], primarySegmentCount(count0, ctx.blockSize, gid));
        storeAgent(&counters->pcount[1][gid], 0u);
    }
    LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    grid.sync();
    uint32_t c = 0;
    const uint32_t seg = blockIdx.x % kSegments;
    for (uint32_t bounce = 0; bounce < ctx.shade.maxDepth; ++bounce) {
        uint32_t sel = bounce & 1u;
        uint32_t c1 = c == 2u ? 0u : c + 1u;
        uint32_t c2 = c1 == 2u ? 0u : c1 + 1u;
        SegmentView v;
        {
            uint32_t acc = 0;
            for (uint32_t s = 0; s < kSegments; ++s) {
                v.start[s] = acc;
                acc += loadAgent(&counters->pcount[c][s]);
            }
            v.start[kSegments] = acc;
        }
        uint32_t count = v.start[kSegments];
        if (count == 0) break;
        if (gid == 0) {
            atomicAdd(&counters->closestQueries, static_cast<unsigned long long>(count));
            if (ctx.profileQueues && bounce < 64) counters->queueCounts[bounce] = count;
        }
        if (gid < kSegments) storeAgent(&counters->pcount[c2][gid], 0u);
        const PathQueue& q = ctx.queue[sel];
        const PathQueue& next = ctx.queue[sel ^ 1u];
        uint32_t shadowQueries = 0;
        uint32_t rounds = (count + stride - 1) / stride;
        for (uint32_t round = 0; round < rounds; ++round) {
            uint32_t g = round * stride + gid;
            bool valid = g < count;
            PathState path;
            ShadeOutcome outcome;
            outcome.alive = false;
            if (valid) {
                path = loadPath(q, segmentSlot(v, q.segmentCapacity, g));
                if constexpr (Full && Events) {
                    // Shading stores its occlusion-dependent additions as
                    // ordered events in this thread's column; the same
                    // traversal call then resolves them one by one.
                    bool shadowQuery = false;
                    Ray ray = path.ray;
                    float tmax = kNoHit;
                    uint32_t eventCount = 0, eventNext = 0;
                    bool eventEnds = false;
                    while (true) {
                        HitRecord h = traceQuery<Full, Query::kEither>(ctx.scene, ray, tmax, shadowQuery, stack);
                        if (shadowQuery) {
                            ++shadowQueries;
                            if (h.kind == kHitNone) {
                                path.radiance += eventContribution(ctx.events, eventNext, gid);
                                if (*eventItemWord(ctx.events, eventNext, kEiFlags, gid) & kEventTerminates)
                                    eventEnds = true;
                            }
                            ++eventNext;
                        } else {
                            EventOcclusion occ{ctx.events, gid, 0u, 0u};
                            shadePath<Full>(ctx.scene, path, h, ctx.shade, occ, outcome);
                            eventCount = occ.count < ctx.events.maxEvents ? occ.count : ctx.events.maxEvents;
                        }
                        while (eventNext < eventCount &&
                            !(*eventItemWord(ctx.events, eventNext, kEiFlags, gid) & kEventTraced)) {
                            path.radiance += eventContribution(ctx.events, eventNext, gid);
                            ++eventNext;
                        }
                        if (eventNext >= eventCount) break;
                        shadowQuery = true;
                        ray = eventRay(ctx.events, eventNext, gid);
                        tmax = __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiTmax, gid));
                    }
                    if (eventEnds) outcome.alive = false;
                } else if constexpr (Full) {
                    HitRecord h = traceClosest<Full>(ctx.scene, path.ray, stack);
                    InlineOcclusion<LdsStack> occ{stack, 0u};
                    shadePath<Full>(ctx.scene, path, h, ctx.shade, occ, outcome);
                    shadowQueries += occ.queries;
                } else {
                    // Without the full feature set the area-light sample is
                    // the bounce's only occlusion test and its last radiance
                    // addition: shading defers it, and the same traversal
                    // call then resolves it (as in the regenerating kernel).
                    DeferOcclusion occ;
                    bool shadowQuery = false;
                    Ray ray = path.ray;
                    float tmax = kNoHit;
                    while (true) {
                        HitRecord h = traceQuery<Full, Query::kEither>(ctx.scene, ray, tmax, shadowQuery, stack);
                        if (shadowQuery) {
                            ++shadowQueries;
                            if (h.kind == kHitNone) path.radiance += outcome.shadow.contribution;
                            break;
                        }
                        shadePath<Full>(ctx.scene, path, h, ctx.shade, occ, outcome);
                        if (!outcome.shadow.valid) break;
                        shadowQuery = true;
                        ray = Ray(outcome.shadow.origin, outcome.shadow.direction);
                        tmax = outcome.shadow.tmax;
                    }
                }
            }
            bool alive = valid && outcome.alive;
            uint32_t a, unused;
            blockAppend(alive, false, &counters->pcount[c1][seg], &counters->pcount[c1][seg], sh, a, unused);
            if (alive)
                storePath(next, seg * next.segmentCapacity + a, path);
            else if (valid)
                finishPath(ctx, path.pixel, outcome.isDebug ? outcome.debugColor : path.radiance);
        }
        if (shadowQueries) atomicAdd(&counters->shadowQueries, static_cast<unsigned long long>(shadowQueries));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
        grid.sync();
        c = c1;
    }
}

__global__ void kTraceClosestAbi(DevScene scene, const float* rays, uint32_t n, wpt_hit* out, uint2* spill) {
    LdsStack stack = makeStack(spill, gridDim.x * blockDim.x);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Ray ray(Vec3(rays[i], rays[n + i], rays[2 * n + i]), Vec3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]));
        HitRecord h = traceClosest(scene, ray, stack);
        wpt_hit r;
        r.t = h.t;
        r.u = h.u;
        r.v = h.v;
        r.kind = h.kind;
        r.instance = h.instance;
        r.primitive = h.prim;
        r.material = h.material;
        r.light_index = kInvalid;
        if (h.kind != kHitNone) r.light_index = surfaceAttributes(scene, ray, h).lightIndex;
        out[i] = r;
    }
}

__global__ void kTraceOccludedAbi(DevScene scene, const float* rays, const float* tmax, uint32_t n, uint8_t* out,
    uint2* spill) {
    LdsStack stack = makeStack(spill, gridDim.x * blockDim.x);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        Ray ray(Vec3(rays[i], rays[n + i], rays[2 * n + i]), Vec3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]));
        out[i] = traceOccluded(scene, ray, tmax[i], stack) ? 1 : 0;
    }
}

__global__ void kCompactReset(uint32_t* scratch, uint32_t words, uint32_t* count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = 0;
    for (uint32_t j = i; j < words; j += gridDim.x * blockDim.x) scratch[j] = 0;
}

__global__ void kCompactAbi(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity,
    uint32_t* count, uint32_t* scratch) {
    __shared__ CompactShared sh;
    if (threadIdx.x == 0) sh.chunk = atomicAdd(&scratch[0], 1u);
    __syncthreads();
    uint32_t chunk = sh.chunk;
    if (chunk * blockDim.x >= n) return;
    uint32_t i = chunk * blockDim.x + threadIdx.x;
    bool alive = i < n && flags[i] != 0;
    uint32_t item = i < n ? items[i] : 0u;
    uint32_t slot;
    uint32_t inclusive = orderedSlots(alive, chunk, scratch + 1, sh, slot);
    if (alive && slot < capacity) out[slot] = item;
    uint32_t chunks = (n + blockDim.x - 1) / blockDim.x;
    if (threadIdx.x == 0 && chunk + 1 == chunks) *count = inclusive;
}

}  // namespace

inline uint32_t ceilDiv(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

namespace {

// One staged sample; `timer` (optional) brackets every launch with events.
template <class Timer>
hipError_t stagedSample(const LaunchContext& ctx, hipStream_t stream, Timer&& timer) {
    size_t lds = traversalLdsBytes(ctx.blockSize);
    dim3 grid(ctx.gridBlocks), block(ctx.blockSize);
    // The caustic gather adds radiance after the area light's, so a
    // deferred shadow ray would reorder the sum.
    const bool inlineShadows = shadesWithInlineShadows(ctx.scene) || ctx.shade.photons != nullptr;
    const bool full = ctx.scene.fullFeatures != 0;
    // Such scenes shade without tracing when the renderer holds an event
    // buffer (the usual case), and trace inline otherwise.
    const bool eventShading = inlineShadows && full && ctx.events.words != nullptr;
    hipError_t e = timer(0, [&] { hipLaunchKernelGGL(kPrimary, grid, block, 0, stream, ctx); });
    for (uint32_t bounce = 0; e == hipSuccess && bounce < ctx.shade.maxDepth; ++bounce) {
        uint32_t sel = bounce & 1u;
        e = timer(1, [&] { hipLaunchKernelGGL(kReset, dim3(1), dim3(64), 0, stream, ctx, sel ^ 1u); });
        if (e == hipSuccess)
            e = timer(2, [&] {
                if (full)
                    hipLaunchKernelGGL(kTrace<true>, grid, block, lds, stream, ctx, bounce, sel);
                else
                    hipLaunchKernelGGL(kTrace<false>, grid, block, lds, stream, ctx, bounce, sel);
            });
        if (e == hipSuccess)
            e = timer(3, [&] {
                if (eventShading)
                    hipLaunchKernelGGL(kShadeEvents<true>, grid, block, 0, stream, ctx, sel);
                else if (inlineShadows)
                    hipLaunchKernelGGL((kShade<true, true>), grid, block, lds, stream, ctx, sel);
                else if (full)
                    hipLaunchKernelGGL((kShade<false, true>), grid, block, 0, stream, ctx, sel);
                else
                    hipLaunchKernelGGL((kShade<false, false>), grid, block, 0, stream, ctx, sel);
            });
        if (e == hipSuccess && eventShading && ctx.shade.photons)
            e = timer(3, [&] { hipLaunchKernelGGL(kGatherEvents, grid, block, 0, stream, ctx); });
        if (e == hipSuccess && eventShading)
            e = timer(5, [&] { hipLaunchKernelGGL(kShadowEvents<true>, grid, block, lds, stream, ctx); });
        if (e == hipSuccess && eventShading)
            e = timer(5, [&] { hipLaunchKernelGGL(kResolveEvents, grid, block, 0, stream, ctx, sel); });
        if (e == hipSuccess && !inlineShadows && ctx.scene.lightCount != 0 && !ctx.scene.ignoreDirectLightOcclusion)
            e = timer(5, [&] {
                if (full)
                    hipLaunchKernelGGL(kShadow<true>, grid, block, lds, stream, ctx, sel);
                else
                    hipLaunchKernelGGL(kShadow<false>, grid, block, lds, stream, ctx, sel);
            });
    }
    if (e != hipSuccess) return e;
    if (ctx.pieces) {
        uint32_t blocks = (ctx.passPixels + ctx.blockSize - 1) / ctx.blockSize;
        e = timer(6, [&] { hipLaunchKernelGGL(kResolvePieces, dim3(blocks > 0 ? blocks : 1u), block, 0, stream, ctx); });
        return e != hipSuccess ? e : hipGetLastError();
    }
    e = timer(6, [&] { hipLaunchKernelGGL(kResolve, dim3(ctx.resolveBlocks), block, 0, stream, ctx); });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kAdvance, dim3(1), dim3(1), 0, stream, ctx);
    return hipGetLastError();
}

hipError_t launchPersistent(const LaunchContext& ctx, hipStream_t stream) {
    LaunchContext local = ctx;
    void* args[] = {&local};
    const void* kernel = !ctx.scene.fullFeatures ? reinterpret_cast<const void*>(kPersistent<false, false>)
        : ctx.events.words                       ? reinterpret_cast<const void*>(kPersistent<true, true>)
                                                 : reinterpret_cast<const void*>(kPersistent<true, false>);
    return hipLaunchCooperativeKernel(kernel, dim3(ctx.persistentBlocks),
        dim3(ctx.blockSize), args, stackLdsBytes(ctx.blockSize), stream);
}

}  // namespace

hipError_t enqueueStagedBatchPass(const LaunchContext& ctx, hipStream_t stream) {
    return stagedSample(ctx, stream, [](int, auto&& launch) {
        launch();
        return hipGetLastError();
    });
}

hipError_t enqueueStagedSample(const LaunchContext& ctx, hipStream_t stream) {
    return stagedSample(ctx, stream, [](int, auto&& launch) {
        launch();
        return hipGetLastError();
    });
}

hipError_t kernelOccupancy(uint32_t blockSize, bool full, Occupancy* out, bool events) {
    size_t lds = stackLdsBytes(blockSize), traversal = traversalLdsBytes(blockSize);
    int b = static_cast<int>(blockSize);
    hipError_t e = full ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->trace, kTrace<true>, b, traversal)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->trace, kTrace<false>, b, traversal);
    if (e == hipSuccess)
        e = full ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->shade, kShade<true, true>, b, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->shade, kShade<false, false>, b, 0);
    if (e == hipSuccess)
        e = full ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->shadow, kShadow<true>, b, traversal)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->shadow, kShadow<false>, b, traversal);
    // The cooperative grid must fit the variant that runs.
    if (e == hipSuccess)
        e = !full  ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->persistent, kPersistent<false, false>, b, lds)
            : events ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->persistent, kPersistent<true, true>, b, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&out->persistent, kPersistent<true, false>, b, lds);
    return e;
}

hipError_t kernelAttributes(bool persistent, bool full, hipFuncAttributes* out, bool events) {
    const void* k = persistent ? (full ? (events ? reinterpret_cast<const void*>(kPersistent<true, true>)
                                                 : reinterpret_cast<const void*>(kPersistent<true, false>))
                                       : reinterpret_cast<const void*>(kPersistent<false, false>))
                               : (full ? reinterpret_cast<const void*>(kTrace<true>)
                                       : reinterpret_cast<const void*>(kTrace<false>));
    return hipFuncGetAttributes(out, k);
}

hipError_t enqueuePersistentSample(const LaunchContext& ctx, hipStream_t stream) {
    hipError_t e = launchPersistent(ctx, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kResolve, dim3(ctx.resolveBlocks), dim3(ctx.blockSize), 0, stream, ctx);
    hipLaunchKernelGGL(kAdvance, dim3(1), dim3(1), 0, stream, ctx);
    return hipGetLastError();
}

hipError_t profileSample(const LaunchContext& ctx, hipStream_t stream, bool persistent, double ms[kProfileSlots],
    uint32_t launches[kProfileSlots]) {
    struct Span {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Span> spans;
    auto timer = [&](int kind, auto&& launch) -> hipError_t {
        Span sp{kind, nullptr, nullptr};
        hipError_t e = hipEventCreate(&sp.a);
        if (e == hipSuccess) e = hipEventCreate(&sp.b);
        if (e != hipSuccess) return e;
        spans.push_back(sp);
        e = hipEventRecord(sp.a, stream);
        if (e != hipSuccess) return e;
        launch();
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        return hipEventRecord(sp.b, stream);
    };
    hipError_t e;
    if (persistent) {
        hipError_t launchError = hipSuccess;
        e = timer(4, [&] { launchError = launchPersistent(ctx, stream); });
        if (e == hipSuccess) e = launchError;
        if (e == hipSuccess)
            e = timer(6, [&] { hipLaunchKernelGGL(kResolve, dim3(ctx.resolveBlocks), dim3(ctx.blockSize), 0, stream, ctx); });
        if (e == hipSuccess) {
            hipLaunchKernelGGL(kAdvance, dim3(1), dim3(1), 0, stream, ctx);
            e = hipGetLastError();
        }
    } else {
        e = stagedSample(ctx, stream, timer);
    }
    hipError_t s = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = s;
    for (Span& sp : spans) {
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, sp.a, sp.b);
        ms[sp.kind] += t;
        launches[sp.kind] += 1;
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    return e;
}

uint32_t abiTraceThreads() { return kAbiBlock * kAbiBlocks; }

hipError_t launchTraceClosest(const DevScene& scene, const float* rays6, uint32_t n, void* hitsOut, uint2* spill,
    hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kTraceClosestAbi, dim3(kAbiBlocks), dim3(kAbiBlock), stackLdsBytes(kAbiBlock), stream, scene,
        rays6, n, static_cast<wpt_hit*>(hitsOut), spill);
    return hipGetLastError();
}

hipError_t launchTraceOccluded(const DevScene& scene, const float* rays6, const float* tmax, uint32_t n,
    uint8_t* out, uint2* spill, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kTraceOccludedAbi, dim3(kAbiBlocks), dim3(kAbiBlock), stackLdsBytes(kAbiBlock), stream, scene,
        rays6, tmax, n, out, spill);
    return hipGetLastError();
}

uint32_t compactScratchWords(uint32_t n) { return 1u + ceilDiv(n > 0 ? n : 1u, 256u); }

hipError_t launchCompact(const uint32_t* items, const uint8_t* flags, uint32_t n, uint32_t* out, uint32_t capacity,
    uint32_t* count, uint32_t* scratch, uint32_t scratchWords, hipStream_t stream) {
    hipLaunchKernelGGL(kCompactReset, dim3(ceilDiv(scratchWords, 256u)), dim3(256), 0, stream, scratch, scratchWords,
        count);
    if (n == 0) return hipGetLastError();
    hipLaunchKernelGGL(kCompactAbi, dim3(ceilDiv(n, 256u)), dim3(256), 0, stream, items, flags, n, out, capacity,
        count, scratch);
    return hipGetLastError();
}

}  // namespace wpt
