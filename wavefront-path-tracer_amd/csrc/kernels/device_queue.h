// Device-side building blocks shared by the wavefront kernels and the batch
// consumer: agent-scope flag access, the LDS traversal stack with global
// spill, wave-level work claims and appends, and the tile pixel order.
#pragma once

#include <hip/hip_runtime.h>

#include "wavefront.h"

namespace wpt {

constexpr uint32_t kSpillEntries = kMaxStackDepth - kLdsStackEntries;

// Occupancy target of the traversal kernels (waves per SIMD), 0: compiler's
// choice.  The register budget it implies must not force scratch spills.
// The batch kernel has its own targets with (Full) and without the full
// feature set; a target of 1 leaves the choice to the compiler.
#ifndef WPT_BATCH_WAVES_PER_EU
#define WPT_BATCH_WAVES_PER_EU 0
#endif
#ifndef WPT_BATCH_LEAN_WAVES_PER_EU
#define WPT_BATCH_LEAN_WAVES_PER_EU 0
#endif
constexpr int batchWavesPerEu(bool full) {
    return (full ? WPT_BATCH_WAVES_PER_EU : WPT_BATCH_LEAN_WAVES_PER_EU) > 0
        ? (full ? WPT_BATCH_WAVES_PER_EU : WPT_BATCH_LEAN_WAVES_PER_EU)
        : 1;
}
#define WPT_BATCH_OCCUPANCY(full) __attribute__((amdgpu_waves_per_eu(batchWavesPerEu(full))))
#if defined(WPT_TRACE_WAVES_PER_EU) && WPT_TRACE_WAVES_PER_EU > 0
#define WPT_TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WPT_TRACE_WAVES_PER_EU)))
#else
#define WPT_TRACE_OCCUPANCY
#endif
// Occupancy targets of the cooperative kernel with the full feature set and
// radiance events, and without the full feature set; 0: the compiler's
// choice (249 and 142 VGPRs, 2 and 3 waves per SIMD).
#ifndef WPT_PERSISTENT_WAVES_PER_EU
#define WPT_PERSISTENT_WAVES_PER_EU 0
#endif
#ifndef WPT_PERSISTENT_LEAN_WAVES_PER_EU
#define WPT_PERSISTENT_LEAN_WAVES_PER_EU 0
#endif
constexpr int persistentWavesPerEu(bool full, bool events) {
    return full ? (events && WPT_PERSISTENT_WAVES_PER_EU > 0 ? WPT_PERSISTENT_WAVES_PER_EU : 1)
                : (WPT_PERSISTENT_LEAN_WAVES_PER_EU > 0 ? WPT_PERSISTENT_LEAN_WAVES_PER_EU : 1);
}
#define WPT_PERSISTENT_OCCUPANCY(full, events) \
    __attribute__((amdgpu_waves_per_eu(persistentWavesPerEu(full, events))))
// Occupancy target of the full-feature event shading kernel, 0: compiler's
// choice (206 VGPRs, 2 waves per SIMD).
#if defined(WPT_SHADE_WAVES_PER_EU) && WPT_SHADE_WAVES_PER_EU > 0
#define WPT_SHADE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WPT_SHADE_WAVES_PER_EU)))
#else
#define WPT_SHADE_OCCUPANCY
#endif
// Occupancy target of the deferred caustic gather kernel, 0: compiler's
// choice.
#if defined(WPT_GATHER_WAVES_PER_EU) && WPT_GATHER_WAVES_PER_EU > 0
#define WPT_GATHER_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WPT_GATHER_WAVES_PER_EU)))
#else
#define WPT_GATHER_OCCUPANCY
#endif

__device__ __forceinline__ uint32_t loadAgent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeAgent(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

extern __shared__ uint2 gLdsStack[];

// Ring of the newest kLdsStackEntries entries in LDS, older ones in global
// memory.  Entry e lives in ring slot e % K while e >= size - K.
struct LdsStack {
    uint2* ring;      // this lane's column: ring[slot * stride]
    uint32_t stride;  // workgroup size
    uint2* spill;     // this thread's column: spill[e * spillStride]
    uint32_t spillStride;
    uint32_t size;

    __device__ void push(uint32_t n, float t) {
        uint32_t slot = size & (kLdsStackEntries - 1u);
        if (size >= kLdsStackEntries) spill[(size - kLdsStackEntries) * spillStride] = ring[slot * stride];
        ring[slot * stride] = make_uint2(n, __float_as_uint(t));
        ++size;
    }
    __device__ void pop(uint32_t& n, float& t) {
        --size;
        uint32_t slot = size & (kLdsStackEntries - 1u);
        uint2 e = ring[slot * stride];
        if (size >= kLdsStackEntries) ring[slot * stride] = spill[(size - kLdsStackEntries) * spillStride];
        n = e.x;
        t = __uint_as_float(e.y);
    }
    // Entry k below the top, for k < 8 (always in the LDS ring).
    __device__ uint2 peek(uint32_t k) const { return ring[((size - 1u - k) & (kLdsStackEntries - 1u)) * stride]; }
    __device__ void poke(uint32_t k, uint2 e) { ring[((size - 1u - k) & (kLdsStackEntries - 1u)) * stride] = e; }
    __device__ void truncate(uint32_t base) {
        while (size > base) {
            uint32_t n;
            float t;
            pop(n, t);
        }
    }
};
static_assert((kLdsStackEntries & (kLdsStackEntries - 1u)) == 0, "LDS stack depth must be a power of two");
static_assert(kLdsStackEntries >= 8, "wide nodes reorder their children within the LDS part of the stack");

__device__ __forceinline__ LdsStack makeStack(uint2* spill, uint32_t spillThreads) {
    LdsStack s;
    s.ring = gLdsStack + threadIdx.x;
    s.stride = blockDim.x;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    s.spill = spill + g;
    s.spillStride = spillThreads;
    s.size = 0;
    return s;
}

// ---- segmented queues
// Flattened view of a queue's segments: segment s holds the items
// [start[s], start[s + 1]) of the enumeration.
struct SegmentView {
    uint32_t start[kSegments + 1];
};

__device__ __forceinline__ SegmentView loadSegments(const uint32_t* counts) {
    SegmentView v;
    uint32_t acc = 0;
#pragma unroll
    for (uint32_t s = 0; s < kSegments; ++s) {
        v.start[s] = acc;
        acc += __builtin_amdgcn_readfirstlane(counts[s]);
    }
    v.start[kSegments] = acc;
    return v;
}

// Queue slot of enumerated item g (< total).
// Constant indices only (a select chain), so the view stays in scalar
// registers.
__device__ __forceinline__ uint32_t segmentSlot(const SegmentView& v, uint32_t segmentCapacity, uint32_t g) {
    uint32_t s = 0, first = 0;
#pragma unroll
    for (uint32_t k = 1; k < kSegments; ++k) {
        bool beyond = g >= v.start[k];
        s = beyond ? k : s;
        first = beyond ? v.start[k] : first;
    }
    return s * segmentCapacity + (g - first);
}

// Layout of a freshly generated queue: item g belongs to virtual workgroup
// g / B of size B and goes to that workgroup's segment, packed in order.
WPT_HD uint32_t primarySlot(uint32_t g, uint32_t B, uint32_t segmentCapacity) {
    uint32_t b = g / B;
    return (b % kSegments) * segmentCapacity + (b / kSegments) * B + (g - b * B);
}

WPT_HD uint32_t primarySegmentCount(uint32_t total, uint32_t B, uint32_t s) {
    uint32_t nb = (total + B - 1) / B;
    if (nb <= s) return 0;
    uint32_t blocks = (nb - 1 - s) / kSegments + 1;
    uint32_t items = blocks * B;
    if ((nb - 1) % kSegments == s) items -= nb * B - total;
    return items;
}

// Workgroup-aggregated append of up to two kinds of items: one atomic per
// kind per workgroup on this workgroup's segment counters.  Every thread of
// the workgroup must call it.  Returns each active lane's index within its
// segment.
constexpr uint32_t kMaxWaves = 16;  // 1024-lane workgroups
struct AppendShared {
    uint32_t wave[kMaxWaves][3];
    uint32_t base[3];
};

__device__ __forceinline__ void blockAppend(bool a0, bool a1, uint32_t* counter0, uint32_t* counter1, AppendShared& sh,
    uint32_t& index0, uint32_t& index1) {
    unsigned long long m0 = __ballot(a0), m1 = __ballot(a1);
    uint32_t r0 = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m0 >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m0), 0u));
    uint32_t r1 = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m1 >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m1), 0u));
    uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        sh.wave[w][0] = static_cast<uint32_t>(__popcll(m0));
        sh.wave[w][1] = static_cast<uint32_t>(__popcll(m1));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t waves = blockDim.x >> 6;
        uint32_t t0 = 0, t1 = 0;
        for (uint32_t k = 0; k < waves; ++k) {
            uint32_t c0 = sh.wave[k][0], c1 = sh.wave[k][1];
            sh.wave[k][0] = t0;
            sh.wave[k][1] = t1;
            t0 += c0;
            t1 += c1;
        }
        sh.base[0] = t0 ? atomicAdd(counter0, t0) : 0u;
        sh.base[1] = t1 ? atomicAdd(counter1, t1) : 0u;
    }
    __syncthreads();
    index0 = sh.base[0] + sh.wave[w][0] + r0;
    index1 = sh.base[1] + sh.wave[w][1] + r1;
    __syncthreads();
}

// Workgroup-aggregated append of continuing paths (kind 0, one item per
// `alive` lane), traced events (kind 1, bit e of `mask` for event e <
// `events`) and deferred gathers (kind 2, one item per `gather` lane): one
// atomic per kind per workgroup.  Every thread of the workgroup must call
// it.  index0 and index2 are a lane's indices within its segments; the
// wave's event entries start at waveBase1, grouped by event index and in
// lane order within a group (what the lanes' own ballots over e < events
// reproduce).
__device__ __forceinline__ void blockAppendEvents(bool alive, uint32_t mask, uint32_t events, bool gather,
    uint32_t* counter0, uint32_t* counter1, uint32_t* counter2, AppendShared& sh, uint32_t& index0,
    uint32_t& waveBase1, uint32_t& index2) {
    unsigned long long m0 = __ballot(alive), m2 = __ballot(gather);
    uint32_t r0 = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m0 >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m0), 0u));
    uint32_t r2 = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m2 >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m2), 0u));
    uint32_t total1 = 0;
    for (uint32_t e = 0; e < events; ++e) total1 += static_cast<uint32_t>(__popcll(__ballot((mask >> e) & 1u)));
    uint32_t w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        sh.wave[w][0] = static_cast<uint32_t>(__popcll(m0));
        sh.wave[w][1] = total1;
        sh.wave[w][2] = static_cast<uint32_t>(__popcll(m2));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t waves = blockDim.x >> 6;
        uint32_t t[3] = {0, 0, 0};
        for (uint32_t k = 0; k < waves; ++k)
            for (uint32_t c = 0; c < 3; ++c) {
                uint32_t n = sh.wave[k][c];
                sh.wave[k][c] = t[c];
                t[c] += n;
            }
        sh.base[0] = t[0] ? atomicAdd(counter0, t[0]) : 0u;
        sh.base[1] = t[1] ? atomicAdd(counter1, t[1]) : 0u;
        sh.base[2] = t[2] ? atomicAdd(counter2, t[2]) : 0u;
    }
    __syncthreads();
    index0 = sh.base[0] + sh.wave[w][0] + r0;
    waveBase1 = sh.base[1] + sh.wave[w][1];
    index2 = sh.base[2] + sh.wave[w][2] + r2;
    __syncthreads();
}

// Tile-local index -> pixel.  Rows are walked in strips of 8 and each strip
// column-major, so a wave's 64 consecutive paths cover an 8x8 pixel block.
__device__ __forceinline__ void tilePixel(const RenderTile& tile, uint32_t i, uint32_t& x, uint32_t& y) {
    uint32_t stripPixels = 8u * tile.width;
    uint32_t strip = i / stripPixels;
    uint32_t j = i - strip * stripPixels;
    uint32_t rows = tile.height - strip * 8u;
    rows = rows < 8u ? rows : 8u;
    x = tile.x + j / rows;
    y = tile.y + strip * 8u + j % rows;
}

__device__ __forceinline__ void accumulate(float* fb, uint32_t pixel, const Vec3& v) {
    float* p = fb + 3u * static_cast<size_t>(pixel);
    p[0] += v.x;
    p[1] += v.y;
    p[2] += v.z;
}

}  // namespace wpt
