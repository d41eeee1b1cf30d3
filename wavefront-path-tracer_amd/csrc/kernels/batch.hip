// Batched persistent consumer: many tasks (tiles x sample ranges) of several
// resident scenes rendered by one launch.  The reference renders tasks one
// at a time (RenderTaskRenderer::render, src/renderer/render_task.cu:26-114);
// this kernel has no reference counterpart.
//
// Each wave works on one task at a time, so the task's scene, camera and
// path parameters are wave-uniform and read through scalar loads.  Idle
// lanes claim tile pixels from the task's cursor with one atomic per wave.
// A lane renders every sample of its pixel in order, one bounce per loop
// iteration (closest hit, then shading with its shadow rays traced inline), and starts the pixel's next
// sample in the same iteration its path ends, so lanes do not wait for the
// longest path of the wave.  The pixel's sum is accumulated in sample order
// from zero (or from the pixel's current sum when the task accumulates into
// a framebuffer), the same additions the staged renderer makes into its
// framebuffer one sample at a time, which makes the outputs bit-identical.
// When the task has no unclaimed pixels and the wave's lanes have finished
// theirs, the wave moves on to the next task with unclaimed pixels; the
// kernel ends when none is left.  With no barrier between bounces, waves
// that finish early keep taking new pixels while a few long traversals run.
#include <hip/hip_runtime.h>

#include "device_queue.h"
#include "wavefront.h"

namespace wpt {

namespace {

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ unsigned long long waveSum(unsigned long long v) {
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    return v;
}

__global__ void __launch_bounds__(512) kBatchReset(BatchContext ctx) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctx.counters->startClock = ~0ull;
        ctx.counters->closestQueries = 0;
        ctx.counters->shadowQueries = 0;
    }
    for (uint32_t t = i; t < ctx.taskCount; t += gridDim.x * blockDim.x) {
        ctx.state[t].cursor = 0;
        ctx.state[t].done = 0;
        ctx.state[t].finishClock = 0;
    }
}

// Every lane runs one traversal per loop iteration, through one call site:
// the closest hit of its path, or a shadow ray its last shading step left.
// - Scenes without environment maps, punctual lights and texture bindings
//   (the kernel without `Full`) test occlusion only for the area-light
//   sample, and that sample's contribution is the last radiance a bounce
//   adds: the shading step returns it as one deferred request.
// - Full scenes with `Events` shade through ordered radiance events
//   (EventOcclusion) stored in the lane's column of the event buffer; the
//   lane resolves them in order in the following iterations, one traced ray
//   each, adding the visible contributions to the path's radiance and ending
//   the path when an escaping event says so.
// Both add the same contributions in the same order as tracing inline, which
// Full scenes without an event buffer do.
template <bool Full, bool Events>
__global__ void __launch_bounds__(512) WPT_BATCH_OCCUPANCY(Full) kBatch(BatchContext ctx) {
    constexpr bool kDefer = !Full;
    LdsStack stack = makeStack(ctx.spill, ctx.spillThreads);
    const uint32_t n = ctx.taskCount;
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t waveId = lane >> 6;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    if (__lane_id() == 0) atomicMin(&ctx.counters->startClock, static_cast<unsigned long long>(wall_clock64()));
    // Waves start spread evenly over the task list.
    uint32_t task = uniform(static_cast<uint32_t>((static_cast<unsigned long long>(waveId) * n) / waves));
    bool taskOpen = true;

    bool busy = false;
    uint32_t x = 0, y = 0, samplesDone = 0, local = 0;
    Vec3 sum(0.0f, 0.0f, 0.0f);
    PathState path;
    ShadowRequest pending;   // deferred shadow ray of the last shading step
    pending.valid = 0;
    bool pathEnded = false;  // the path ended at its last shading step
    uint32_t eventCount = 0, eventNext = 0;  // events of the last shading step, and the next to resolve
    bool eventEnds = false;                  // an escaping event ended the path
    unsigned long long closest = 0, shadow = 0;

    // Adds the untraced events from eventNext on, up to the next traced one;
    // when none is left, returns whether the path has ended.
    auto drainEvents = [&]() -> bool {
        while (eventNext < eventCount) {
            if (*eventItemWord(ctx.events, eventNext, kEiFlags, lane) & kEventTraced) return false;
            path.radiance += Vec3(__uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCx, lane)),
                __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCy, lane)),
                __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCz, lane)));
            ++eventNext;
        }
        eventCount = 0;
        eventNext = 0;
        return pathEnded || eventEnds;
    };

    while (true) {
        const BatchTask& T = ctx.tasks[task];
        if (taskOpen) {
            bool need = !busy;
            unsigned long long mask = __ballot(need);
            if (mask) {
                uint32_t want = static_cast<uint32_t>(__popcll(mask));
                uint32_t base = 0;
                if (__lane_id() == 0) base = atomicAdd(&ctx.state[task].cursor, want);
                base = uniform(base);
                if (base + want >= T.pixelCount) taskOpen = false;
                uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                    __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
                uint32_t u = base + rank;
                if (need && u < T.pixelCount) {
                    busy = true;
                    tilePixel(T.tile, u, x, y);
                    local = (y - T.tile.y) * T.outStride + (x - T.tile.x);
                    samplesDone = 0;
                    sum = Vec3(0.0f, 0.0f, 0.0f);
                    if (T.accumulate) {
                        const float* o = ctx.out + T.out + 3ull * local;
                        sum = Vec3(o[0], o[1], o[2]);
                    }
                    path = primaryPath(T.camera, x, y, T.width, T.height, T.sampleStart, T.shade.rngStrategy);
                }
            }
        }
        if (__ballot(busy) == 0) {
            // Only reached with the task closed: find one with unclaimed pixels.
            bool found = false;
            for (uint32_t k = 1; k <= n; ++k) {
                uint32_t t = task + k;
                if (t >= n) t -= n;
                uint32_t claimed = uniform(loadAgent(&ctx.state[t].cursor));
                if (claimed < ctx.tasks[t].pixelCount) {
                    task = t;
                    found = true;
                    break;
                }
            }
            if (!found) break;
            taskOpen = true;
            continue;
        }
        if (busy) {
            const DevScene& scene = ctx.scenes[T.scene];
            bool shadowQuery = false;
            Ray queryRay = path.ray;
            float queryTmax = 0.0f;
            if constexpr (kDefer) {
                if (pending.valid) {
                    shadowQuery = true;
                    queryRay = Ray(pending.origin, pending.direction);
                    queryTmax = pending.tmax;
                }
            } else if constexpr (Events) {
                if (eventNext < eventCount) {  // a traced event (untraced ones are drained already)
                    shadowQuery = true;
                    queryRay = Ray(Vec3(__uint_as_float(*eventItemWord(ctx.events, eventNext, kEiOx, lane)),
                                       __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiOy, lane)),
                                       __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiOz, lane))),
                        Vec3(__uint_as_float(*eventItemWord(ctx.events, eventNext, kEiDx, lane)),
                            __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiDy, lane)),
                            __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiDz, lane))));
                    queryTmax = __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiTmax, lane));
                }
            }
            HitRecord h = traceQuery<Full, (kDefer || Events) ? Query::kEither : Query::kClosest>(scene, queryRay,
                queryTmax, shadowQuery, stack);
            bool ended;
            if (shadowQuery) {
                ++shadow;
                if constexpr (kDefer) {
                    if (h.kind == kHitNone) path.radiance += pending.contribution;
                    pending.valid = 0;
                    ended = pathEnded;
                } else {
                    if (h.kind == kHitNone) {
                        path.radiance += Vec3(__uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCx, lane)),
                            __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCy, lane)),
                            __uint_as_float(*eventItemWord(ctx.events, eventNext, kEiCz, lane)));
                        if (*eventItemWord(ctx.events, eventNext, kEiFlags, lane) & kEventTerminates) eventEnds = true;
                    }
                    ++eventNext;
                    ended = drainEvents();
                    // Events still pending: the path waits for the next iteration.
                    if (eventCount != 0) ended = false;
                }
            } else {
                ++closest;
                ShadeOutcome outcome;
                if constexpr (kDefer) {
                    DeferOcclusion occ;
                    shadePath<Full>(scene, path, h, T.shade, occ, outcome);
                    pending = outcome.shadow;
                } else if constexpr (Events) {
                    EventOcclusion occ{ctx.events, lane, 0u, 0u};
                    shadePath<Full>(scene, path, h, T.shade, occ, outcome);
                    eventCount = occ.count < ctx.events.maxEvents ? occ.count : ctx.events.maxEvents;
                    eventNext = 0;
                    eventEnds = false;
                } else {
                    InlineOcclusion<LdsStack> occ{stack, 0u};
                    shadePath<Full>(scene, path, h, T.shade, occ, outcome);
                    shadow += occ.queries;
                }
                if (outcome.isDebug) path.radiance = outcome.debugColor;
                pathEnded = !outcome.alive;
                if constexpr (Events) {
                    ended = drainEvents();
                    if (eventCount != 0) ended = false;
                } else {
                    ended = pathEnded && !(kDefer && pending.valid);
                }
            }
            if (ended) {
                sum += path.radiance;
                ++samplesDone;
                if (samplesDone < T.sampleCount) {
                    path = primaryPath(T.camera, x, y, T.width, T.height, T.sampleStart + samplesDone,
                        T.shade.rngStrategy);
                } else {
                    float* o = ctx.out + T.out + 3ull * local;
                    o[0] = sum.x;
                    o[1] = sum.y;
                    o[2] = sum.z;
                    busy = false;
                    if (atomicAdd(&ctx.state[task].done, 1u) + 1u == T.pixelCount)
                        ctx.state[task].finishClock = wall_clock64();
                }
            }
        }
    }
    closest = waveSum(closest);
    shadow = waveSum(shadow);
    if (__lane_id() == 0) {
        atomicAdd(&ctx.counters->closestQueries, closest);
        atomicAdd(&ctx.counters->shadowQueries, shadow);
    }
}

}  // namespace

hipError_t enqueueBatch(const BatchContext& ctx, uint32_t blocks, uint32_t blockSize, hipStream_t stream) {
    uint32_t resetBlocks = (ctx.taskCount + 255u) / 256u;
    hipLaunchKernelGGL(kBatchReset, dim3(resetBlocks > 0 ? resetBlocks : 1u), dim3(256), 0, stream, ctx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || ctx.taskCount == 0) return e;
    if (ctx.full && ctx.events.words)
        hipLaunchKernelGGL((kBatch<true, true>), dim3(blocks), dim3(blockSize), stackLdsBytes(blockSize), stream, ctx);
    else if (ctx.full)
        hipLaunchKernelGGL((kBatch<true, false>), dim3(blocks), dim3(blockSize), stackLdsBytes(blockSize), stream, ctx);
    else
        hipLaunchKernelGGL((kBatch<false, false>), dim3(blocks), dim3(blockSize), stackLdsBytes(blockSize), stream,
            ctx);
    return hipGetLastError();
}

hipError_t batchOccupancy(uint32_t blockSize, bool full, int* blocksPerCu) {
    return full ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCu, kBatch<true, true>,
                      static_cast<int>(blockSize), stackLdsBytes(blockSize))
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocksPerCu, kBatch<false, false>,
                      static_cast<int>(blockSize), stackLdsBytes(blockSize));
}

hipError_t batchAttributes(bool full, hipFuncAttributes* out) {
    return full ? hipFuncGetAttributes(out, reinterpret_cast<const void*>(kBatch<true, true>))
                : hipFuncGetAttributes(out, reinterpret_cast<const void*>(kBatch<false, false>));
}

}  // namespace wpt
