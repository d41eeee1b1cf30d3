// Device-side building blocks shared by the wavefront kernels and the batch
// consumer: agent-scope flag access, the LDS traversal stack with global
// spill, wave-level work claims and appends, and the tile pixel order.
#pragma once

#include <hip/hip_runtime.h>

#include "wavefront.h"

namespace wpt {

constexpr uint32_t kSpillEntries = kMaxStackDepth - kLdsStackEntries;

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
    __device__ void truncate(uint32_t base) {
        while (size > base) {
            uint32_t n;
            float t;
            pop(n, t);
        }
    }
};
static_assert((kLdsStackEntries & (kLdsStackEntries - 1u)) == 0, "LDS stack depth must be a power of two");

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

// Wave-aggregated append: one atomic per wave, returns this lane's index.
__device__ __forceinline__ uint32_t waveAppend(bool active, uint32_t* counter) {
    unsigned long long mask = __ballot(active);
    uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
    uint32_t base = 0;
    if (__lane_id() == 0 && mask) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(mask)));
    base = __builtin_amdgcn_readfirstlane(base);
    return base + rank;
}

// Wave-level work claim: 64 consecutive items per call.
__device__ __forceinline__ uint32_t claimWaveBatch(uint32_t* cursor) {
    uint32_t base = 0;
    if (__lane_id() == 0) base = atomicAdd(cursor, 64u);
    return __builtin_amdgcn_readfirstlane(base);
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
