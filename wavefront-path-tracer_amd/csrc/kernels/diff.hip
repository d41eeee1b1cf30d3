// Differentiable rendering kernels: forward paths, per-pixel resolve, the
// path-replay adjoint and an Adam step.  No reference counterpart.
//
// One thread renders one (pixel, sample) path at a time, walking the paths
// with a static grid stride (no shared cursor: device-scope atomics on one
// address serialize across XCDs); consecutive paths cover 8x8 pixel blocks
// of one sample.  The forward pass keeps every path's radiance (3 floats),
// which the adjoint needs as L_path; the resolve kernel averages a pixel's
// samples in sample order, so the image does not depend on scheduling.  The adjoint re-traces
// each path with the same random numbers; texture gradients go to global
// memory with hardware float atomics, roughness and emission gradients are
// summed per workgroup in LDS first (a few dozen slots shared by every path).
#include <hip/hip_runtime.h>

#include "device_queue.h"
#include "diff.h"
#include "diff_kernels.h"

namespace wpt {

namespace {

__device__ __forceinline__ void pathPixel(const DiffContext& c, uint32_t i, uint32_t& sample, uint32_t& x,
    uint32_t& y) {
    uint32_t pixels = c.width * c.height;
    sample = i / pixels;
    RenderTile frame{0u, 0u, c.width, c.height, 0u};
    tilePixel(frame, i - sample * pixels, x, y);
}

__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) kDiffForward(DiffContext c) {
    LdsStack stack = makeStack(c.spill, c.spillThreads);
    const uint32_t pixels = c.width * c.height;
    const uint32_t count = pixels * c.spp;
    NoGradients none;
    const Vec3 zero(0.0f, 0.0f, 0.0f);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint32_t sample, x, y;
        pathPixel(c, i, sample, x, y);
        Vec3 L = diffInlinePath<false>(c.scene, c.layout, c.params, c.camera, x, y, c.width, c.height,
            c.sampleStart + sample, c.rngStrategy, c.maxDepth, stack, zero, zero, none);
        size_t slot = static_cast<size_t>(sample) * pixels + y * c.width + x;
        c.pathRadiance[slot] = L.x;
        c.pathRadiance[static_cast<size_t>(count) + slot] = L.y;
        c.pathRadiance[2 * static_cast<size_t>(count) + slot] = L.z;
    }
}

__global__ void kDiffResolve(DiffContext c) {
    const uint32_t pixels = c.width * c.height;
    const size_t count = static_cast<size_t>(pixels) * c.spp;
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    Vec3 sum(0.0f, 0.0f, 0.0f);
    for (uint32_t s = 0; s < c.spp; ++s) {
        size_t slot = static_cast<size_t>(s) * pixels + p;
        sum += Vec3(c.pathRadiance[slot], c.pathRadiance[count + slot], c.pathRadiance[2 * count + slot]);
    }
    float inv = 1.0f / static_cast<float>(c.spp);
    c.image[3u * p] = sum.x * inv;
    c.image[3u * p + 1] = sum.y * inv;
    c.image[3u * p + 2] = sum.z * inv;
}

// Gradient sink of the adjoint kernel.
struct AdjointSink {
    const DiffContext* c;
    float* ldsRoughness;
    float* ldsEmission;
    __device__ void texture(const TexFootprint& f, const Vec3& g) {
        for (int k = 0; k < 4; ++k) {
            float w = f.weight[k];
            if (w == 0.0f) continue;
            float* t = c->grads.texture + 3u * f.index[k];
            unsafeAtomicAdd(t, g.x * w);
            unsafeAtomicAdd(t + 1, g.y * w);
            unsafeAtomicAdd(t + 2, g.z * w);
        }
    }
    __device__ void roughness(uint32_t slot, float g) {
        if (slot < kDiffLdsSlots)
            atomicAdd(&ldsRoughness[slot], g);
        else
            unsafeAtomicAdd(&c->grads.roughness[slot], g);
    }
    __device__ void emission(uint32_t slot, float g) {
        if (slot < kDiffLdsSlots)
            atomicAdd(&ldsEmission[slot], g);
        else
            unsafeAtomicAdd(&c->grads.emission[slot], g);
    }
};

__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3))) kDiffAdjoint(DiffContext c) {
    __shared__ float ldsRoughness[kDiffLdsSlots];
    __shared__ float ldsEmission[kDiffLdsSlots];
    for (uint32_t k = threadIdx.x; k < kDiffLdsSlots; k += blockDim.x) {
        ldsRoughness[k] = 0.0f;
        ldsEmission[k] = 0.0f;
    }
    __syncthreads();
    LdsStack stack = makeStack(c.spill, c.spillThreads);
    AdjointSink sink{&c, ldsRoughness, ldsEmission};
    const uint32_t pixels = c.width * c.height;
    const uint32_t count = pixels * c.spp;
    const float invSpp = 1.0f / static_cast<float>(c.spp);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        uint32_t sample, x, y;
        pathPixel(c, i, sample, x, y);
        uint32_t p = y * c.width + x;
        size_t slot = static_cast<size_t>(sample) * pixels + p;
        Vec3 total(c.pathRadiance[slot], c.pathRadiance[static_cast<size_t>(count) + slot],
            c.pathRadiance[2 * static_cast<size_t>(count) + slot]);
        Vec3 delta(c.imageGrad[3u * p] * invSpp, c.imageGrad[3u * p + 1] * invSpp, c.imageGrad[3u * p + 2] * invSpp);
        if (delta.x == 0.0f && delta.y == 0.0f && delta.z == 0.0f) continue;
        diffInlinePath<true>(c.scene, c.layout, c.params, c.camera, x, y, c.width, c.height, c.sampleStart + sample,
            c.rngStrategy, c.maxDepth, stack, delta, total, sink);
    }
    __syncthreads();
    uint32_t rc = c.layout.roughnessCount < kDiffLdsSlots ? c.layout.roughnessCount : kDiffLdsSlots;
    uint32_t ec = c.layout.emissionCount < kDiffLdsSlots ? c.layout.emissionCount : kDiffLdsSlots;
    for (uint32_t k = threadIdx.x; k < rc; k += blockDim.x)
        if (ldsRoughness[k] != 0.0f) unsafeAtomicAdd(&c.grads.roughness[k], ldsRoughness[k]);
    for (uint32_t k = threadIdx.x; k < ec; k += blockDim.x)
        if (ldsEmission[k] != 0.0f) unsafeAtomicAdd(&c.grads.emission[k], ldsEmission[k]);
}

__global__ void kDiffClear(float* p, size_t n, uint32_t* cursor) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0 && cursor) *cursor = 0;
    for (; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

__global__ void kAdam(AdamStep a) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.count) return;
    float g = a.grad[i];
    float m = a.beta1 * a.m[i] + (1.0f - a.beta1) * g;
    float v = a.beta2 * a.v[i] + (1.0f - a.beta2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    float mHat = m / a.bias1;
    float vHat = v / a.bias2;
    float p = a.params[i] - a.lr * mHat / (sqrtf(vHat) + a.epsilon);
    float lo, hi;
    if (i < a.textureFloats) {
        lo = a.textureRange[0];
        hi = a.textureRange[1];
    } else if (i < a.textureFloats + a.roughnessCount) {
        lo = a.roughnessRange[0];
        hi = a.roughnessRange[1];
    } else {
        lo = a.emissionRange[0];
        hi = a.emissionRange[1];
    }
    a.params[i] = fminf(hi, fmaxf(lo, p));
}

inline uint32_t blocksFor(size_t n, uint32_t block) {
    size_t b = (n + block - 1) / block;
    return static_cast<uint32_t>(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

hipError_t enqueueDiffForward(const DiffContext& c, uint32_t blocks, uint32_t blockSize, hipStream_t stream) {
    hipLaunchKernelGGL(kDiffClear, dim3(1), dim3(64), 0, stream, static_cast<float*>(nullptr), size_t{0}, c.cursor);
    hipLaunchKernelGGL(kDiffForward, dim3(blocks), dim3(blockSize), stackLdsBytes(blockSize), stream, c);
    uint32_t pixels = c.width * c.height;
    hipLaunchKernelGGL(kDiffResolve, dim3((pixels + 255) / 256), dim3(256), 0, stream, c);
    return hipGetLastError();
}

hipError_t enqueueDiffAdjoint(const DiffContext& c, uint32_t blocks, uint32_t blockSize, size_t gradFloats,
    hipStream_t stream) {
    hipLaunchKernelGGL(kDiffClear, dim3(blocksFor(gradFloats, 256)), dim3(256), 0, stream, c.grads.texture, gradFloats,
        c.cursor);
    hipLaunchKernelGGL(kDiffAdjoint, dim3(blocks), dim3(blockSize), stackLdsBytes(blockSize), stream, c);
    return hipGetLastError();
}

hipError_t enqueueAdam(const AdamStep& a, hipStream_t stream) {
    if (a.count == 0) return hipSuccess;
    hipLaunchKernelGGL(kAdam, dim3(static_cast<uint32_t>((a.count + 255) / 256)), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t diffOccupancy(uint32_t blockSize, int* forwardPerCu, int* adjointPerCu) {
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(forwardPerCu, kDiffForward,
        static_cast<int>(blockSize), stackLdsBytes(blockSize));
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(adjointPerCu, kDiffAdjoint, static_cast<int>(blockSize),
        stackLdsBytes(blockSize));
}

}  // namespace wpt
