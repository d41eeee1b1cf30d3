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
#include "photon_trace.h"
#include "wavefront.h"

namespace wpt {

namespace {

__global__ void __launch_bounds__(256) kPhotonTrace(DevScene scene, PhotonSettings ps, Photon* photons, uint2* spill,
    uint32_t spillThreads) {
    LdsStack stack = makeStack(spill, spillThreads);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ps.count; i += gridDim.x * blockDim.x)
        photons[i] = tracePhoton(scene, ps, i, stack);
}

// Sort keys: (cell << 32) | photon, with photons outside the grid or not
// stored keyed past the last cell; the stored count and the bounds of the
// occupied cells, reduced per block before one atomic each.
__global__ void __launch_bounds__(256) kPhotonKeys(PhotonGridView g, PhotonSettings ps, const Photon* photons,
    uint64_t* keys, uint32_t* stored, uint32_t* occupied) {
    __shared__ uint32_t box[6][256];
    __shared__ uint32_t count[256];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t b[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    uint32_t kept = 0;
    if (i < ps.count) {
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
                kept = 1;
                b[0] = b[3] = static_cast<uint32_t>(x);
                b[1] = b[4] = static_cast<uint32_t>(y);
                b[2] = b[5] = static_cast<uint32_t>(z);
            }
        }
        keys[i] = (static_cast<uint64_t>(cell) << 32) | i;
    }
    for (int k = 0; k < 6; ++k) box[k][threadIdx.x] = b[k];
    count[threadIdx.x] = kept;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                box[k][threadIdx.x] = min(box[k][threadIdx.x], box[k][threadIdx.x + s]);
                box[k + 3][threadIdx.x] = max(box[k + 3][threadIdx.x], box[k + 3][threadIdx.x + s]);
            }
            count[threadIdx.x] += count[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && count[0] != 0) {
        atomicAdd(stored, count[0]);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&occupied[k], box[k][0]);
            atomicMax(&occupied[k + 3], box[k + 3][0]);
        }
    }
}

__global__ void kFillWords(uint32_t* p, uint32_t n, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

// The photons in sorted order.
__global__ void kPhotonSorted(PhotonSettings ps, const uint64_t* sortedKeys, const Photon* photons, Photon* sorted) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < ps.count) sorted[p] = photons[static_cast<uint32_t>(sortedKeys[p])];
}

// cellStart[c]: the sorted keys of cells below c (a lower bound of c << 32;
// c == cells counts every stored photon).
__global__ void kPhotonCellStarts(PhotonSettings ps, uint32_t cells, const uint64_t* sortedKeys, uint32_t* cellStart) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > cells) return;
    const uint64_t key = static_cast<uint64_t>(c) << 32;
    uint32_t lo = 0, hi = ps.count;  // first index whose key >= key
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (sortedKeys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    cellStart[c] = lo;
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
    hipLaunchKernelGGL(kFillWords, dim3(1), dim3(64), 0, stream, b.occupied, 3u, 0xffffffffu);
    hipLaunchKernelGGL(kFillWords, dim3(1), dim3(64), 0, stream, b.occupied + 3, 3u, 0u);
    hipLaunchKernelGGL(kPhotonKeys, dim3(blocks), dim3(256), 0, stream, view, ps, b.photons, b.keys, b.stored,
        b.occupied);
    size_t bytes = b.tempBytes;
    hipError_t e = hipcub::DeviceRadixSort::SortKeys(b.temp, bytes, b.keys, b.sortedKeys, static_cast<int>(ps.count),
        0, static_cast<int>(keyBits(cells)), stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kPhotonSorted, dim3(blocks), dim3(256), 0, stream, ps, b.sortedKeys, b.photons, b.sorted);
    hipLaunchKernelGGL(kPhotonCellStarts, dim3(cells / 256 + 1), dim3(256), 0, stream, ps, cells, b.sortedKeys,
        b.cellStart);
    return hipGetLastError();
}

}  // namespace wpt
