// Device BVH2 builder (PLOC over Morton-sorted triangles).
//
// Passes, all over arrays in HBM:
//   kPrimBounds      triangle AABBs and per-block centroid bounds
//   kReduceBounds    the centroid bounds of the mesh (one block)
//   kMorton          63-bit Morton keys; hipCUB radix sort (stable) by key
//   kInitClusters    one cluster per slot; slot-order geometry and order
//   per pass, while more than one cluster is left:
//     kNearest       each cluster's nearest neighbour by merged surface
//                    area among the `radius` clusters on either side,
//                    staged through LDS (ties: the lower index)
//     kFlags         mutual pairs merge; the higher member leaves
//     2 exclusive scans: merge ranks (node numbers) and survivor positions
//     kMerge         new interior nodes and the compacted cluster list
//   kLeafDepth       tree depth from parent links
//   kPreorder, kPermute  interior nodes renumbered depth first
// Interior nodes are numbered downward from n-2 pass by pass, so the last
// merge, the root, is node 0; the final depth-first renumbering keeps the
// root at 0 and places every subtree's nodes in one contiguous range.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "bvh_build.h"

namespace wpt {

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kMaxRadius = 64;

struct Box {
    float mn[3];
    float mx[3];
};

__device__ inline Box unite(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = fminf(a.mn[k], b.mn[k]);
        r.mx[k] = fmaxf(a.mx[k], b.mx[k]);
    }
    return r;
}

// Half the surface area of the union of two boxes.
__device__ inline float mergedArea(const float* a, const float* b) {
    float dx = fmaxf(a[3], b[3]) - fminf(a[0], b[0]);
    float dy = fmaxf(a[4], b[4]) - fminf(a[1], b[1]);
    float dz = fmaxf(a[5], b[5]) - fminf(a[2], b[2]);
    return dx * dy + dy * dz + dz * dx;
}

__device__ inline float4 pack(const Vec3& v, uint32_t word) { return make_float4(v.x, v.y, v.z, __uint_as_float(word)); }

// Triangle bounds (as the host builder's triangleBounds) and per-block
// centroid bounds.
__global__ void __launch_bounds__(kBlock) kPrimBounds(const TriShade* tris, uint32_t n, Box* boxes, Box* partial) {
    __shared__ float lo[3][kBlock], hi[3][kBlock];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    float c[3] = {INFINITY, INFINITY, INFINITY}, d[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        const TriShade& t = tris[i];
        Box b;
        b.mn[0] = fminf(t.v0.x, fminf(t.v1.x, t.v2.x));
        b.mn[1] = fminf(t.v0.y, fminf(t.v1.y, t.v2.y));
        b.mn[2] = fminf(t.v0.z, fminf(t.v1.z, t.v2.z));
        b.mx[0] = fmaxf(t.v0.x, fmaxf(t.v1.x, t.v2.x));
        b.mx[1] = fmaxf(t.v0.y, fmaxf(t.v1.y, t.v2.y));
        b.mx[2] = fmaxf(t.v0.z, fmaxf(t.v1.z, t.v2.z));
        boxes[i] = b;
        for (int k = 0; k < 3; ++k) c[k] = d[k] = 0.5f * (b.mn[k] + b.mx[k]);
    }
    for (int k = 0; k < 3; ++k) {
        lo[k][threadIdx.x] = c[k];
        hi[k][threadIdx.x] = d[k];
    }
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                lo[k][threadIdx.x] = fminf(lo[k][threadIdx.x], lo[k][threadIdx.x + s]);
                hi[k][threadIdx.x] = fmaxf(hi[k][threadIdx.x], hi[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box r;
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = lo[k][0];
            r.mx[k] = hi[k][0];
        }
        partial[blockIdx.x] = r;
    }
}

__global__ void __launch_bounds__(kBlock) kReduceBounds(const Box* partial, uint32_t count, Box* out) {
    __shared__ float lo[3][kBlock], hi[3][kBlock];
    float c[3] = {INFINITY, INFINITY, INFINITY}, d[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = threadIdx.x; i < count; i += kBlock)
        for (int k = 0; k < 3; ++k) {
            c[k] = fminf(c[k], partial[i].mn[k]);
            d[k] = fmaxf(d[k], partial[i].mx[k]);
        }
    for (int k = 0; k < 3; ++k) {
        lo[k][threadIdx.x] = c[k];
        hi[k][threadIdx.x] = d[k];
    }
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int k = 0; k < 3; ++k) {
                lo[k][threadIdx.x] = fminf(lo[k][threadIdx.x], lo[k][threadIdx.x + s]);
                hi[k][threadIdx.x] = fmaxf(hi[k][threadIdx.x], hi[k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        Box r;
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = lo[k][0];
            r.mx[k] = hi[k][0];
        }
        *out = r;
    }
}

// 21 bits spread to every third bit of a 63-bit word.
__device__ inline uint64_t spread21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void kMorton(const Box* boxes, uint32_t n, const Box* bounds, uint64_t* keys, uint32_t* index) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Box& b = boxes[i];
    const Box& s = *bounds;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float extent = s.mx[k] - s.mn[k];
        float c = 0.5f * (b.mn[k] + b.mx[k]);
        float u = extent > 0.0f ? (c - s.mn[k]) / extent : 0.0f;
        float scaled = u * 2097152.0f;
        q[k] = scaled <= 0.0f ? 0u : (scaled >= 2097151.0f ? 2097151u : static_cast<uint32_t>(scaled));
    }
    keys[i] = spread21(q[0]) << 2 | spread21(q[1]) << 1 | spread21(q[2]);
    index[i] = i;
}

__global__ void kInitClusters(const TriShade* tris, const Box* boxes, const uint32_t* sorted, uint32_t n,
    uint32_t* clusterId, Box* clusterBox, uint32_t* order, float4* geometry, uint32_t* leafParent) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = sorted[i];
    order[i] = p;
    clusterId[i] = kLeafBit | i;
    clusterBox[i] = boxes[p];
    leafParent[i] = kInvalid;
    const TriShade& t = tris[p];
    geometry[3 * i + 0] = pack(t.v0, p);
    geometry[3 * i + 1] = pack(t.v1, t.material);
    geometry[3 * i + 2] = pack(t.v2, t.flags);
}

__global__ void kFillInvalid(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = kInvalid;
}

// Nearest neighbour within [i - radius, i + radius] by merged area; the
// block's clusters and a halo of `radius` on each side are staged in LDS.
__global__ void __launch_bounds__(kBlock) kNearest(const Box* boxes, uint32_t count, uint32_t radius, uint32_t* nn) {
    extern __shared__ float window[];  // (kBlock + 2 radius) boxes of 6 floats
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kBlock - radius;
    const uint32_t span = kBlock + 2 * radius;
    for (uint32_t k = threadIdx.x; k < span; k += kBlock) {
        int64_t g = base + k;
        if (g >= 0 && g < count) {
            const Box& b = boxes[g];
            for (int c = 0; c < 3; ++c) {
                window[6 * k + c] = b.mn[c];
                window[6 * k + 3 + c] = b.mx[c];
            }
        }
    }
    __syncthreads();
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    uint32_t first = i > radius ? i - radius : 0u;
    uint32_t last = i + radius < count - 1 ? i + radius : count - 1;
    const float* self = &window[6 * (i - base)];
    float best = INFINITY;
    uint32_t bestJ = kInvalid;
    for (uint32_t j = first; j <= last; ++j) {
        if (j == i) continue;
        float a = mergedArea(self, &window[6 * (j - base)]);
        if (a < best) {
            best = a;
            bestJ = j;
        }
    }
    nn[i] = bestJ;
}

__global__ void kFlags(const uint32_t* nn, uint32_t count, uint32_t* mergeFlag, uint32_t* keepFlag) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t j = nn[i];
    bool mutual = j < count && nn[j] == i;
    mergeFlag[i] = mutual && i < j ? 1u : 0u;
    keepFlag[i] = mutual && i > j ? 0u : 1u;
}

__global__ void kTotal(const uint32_t* rank, const uint32_t* flag, uint32_t count, uint32_t* total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *total = rank[count - 1] + flag[count - 1];
}

__global__ void kMerge(uint32_t count, const uint32_t* nn, const uint32_t* mergeFlag, const uint32_t* mergeRank,
    const uint32_t* keepFlag, const uint32_t* newPos, uint32_t nodeBase, const uint32_t* clusterId,
    const Box* clusterBox, uint32_t* outId, Box* outBox, BvhNode* nodes, uint32_t* leafParent,
    uint32_t* nodeParent, uint32_t* leafCount) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count || !keepFlag[i]) return;
    uint32_t dst = newPos[i];
    if (!mergeFlag[i]) {
        outId[dst] = clusterId[i];
        outBox[dst] = clusterBox[i];
        return;
    }
    uint32_t j = nn[i];
    uint32_t node = nodeBase + mergeRank[i];
    uint32_t a = clusterId[i], b = clusterId[j];
    const Box A = clusterBox[i], B = clusterBox[j];
    BvhNode r;
    r.lminx = A.mn[0];
    r.lminy = A.mn[1];
    r.lminz = A.mn[2];
    r.lmaxx = A.mx[0];
    r.lmaxy = A.mx[1];
    r.lmaxz = A.mx[2];
    r.rminx = B.mn[0];
    r.rminy = B.mn[1];
    r.rminz = B.mn[2];
    r.rmaxx = B.mx[0];
    r.rmaxy = B.mx[1];
    r.rmaxz = B.mx[2];
    r.left = a;
    r.right = b;
    r.pad0 = r.pad1 = 0;
    nodes[node] = r;
    if (a & kLeafBit) leafParent[a & ~kLeafBit] = node; else nodeParent[a] = node;
    if (b & kLeafBit) leafParent[b & ~kLeafBit] = node; else nodeParent[b] = node;
    // Children were made in earlier passes, so their counts are final.
    leafCount[node] = ((a & kLeafBit) ? 1u : leafCount[a]) + ((b & kLeafBit) ? 1u : leafCount[b]);
    outId[dst] = node;
    outBox[dst] = unite(A, B);
}

// Depth-first (preorder) number of each interior node: the root is 0, a
// left child follows its parent, a right child follows the parent's whole
// left subtree (leaves - 1 interior nodes).  Found by walking to the root.
__global__ void kPreorder(const BvhNode* nodes, const uint32_t* nodeParent, const uint32_t* leafCount, uint32_t count,
    uint32_t* preorder) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= count) return;
    uint32_t index = 0, c = v, p = nodeParent[v];
    for (uint32_t step = 0; p != kInvalid && step < count; ++step) {
        uint32_t left = nodes[p].left;
        index += c == left ? 1u : 1u + ((left & kLeafBit) ? 0u : leafCount[left] - 1u);
        c = p;
        p = nodeParent[p];
    }
    preorder[v] = index;
}

// Nodes moved to their preorder slots, interior child links renumbered.
__global__ void kPermute(const BvhNode* nodes, const uint32_t* preorder, uint32_t count, BvhNode* out) {
    uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= count) return;
    BvhNode r = nodes[v];
    if (!(r.left & kLeafBit)) r.left = preorder[r.left];
    if (!(r.right & kLeafBit)) r.right = preorder[r.right];
    out[preorder[v]] = r;
}

// Edges from each leaf to the root; the block maximum goes to *depth.
__global__ void __launch_bounds__(kBlock) kLeafDepth(const uint32_t* leafParent, const uint32_t* nodeParent,
    uint32_t n, uint32_t* depth) {
    __shared__ uint32_t best[kBlock];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t d = 0;
    if (i < n) {
        uint32_t p = leafParent[i];
        while (p != kInvalid && d < n) {
            ++d;
            p = nodeParent[p];
        }
    }
    best[threadIdx.x] = d;
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) best[threadIdx.x] = max(best[threadIdx.x], best[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(depth, best[0]);
}

// Device allocations of one build, released on every exit.
struct Scratch {
    std::vector<void*> blocks;
    ~Scratch() {
        for (void* p : blocks) (void)hipFree(p);
    }
    template <class T>
    hipError_t get(size_t count, T** out) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) blocks.push_back(p);
        *out = static_cast<T*>(p);
        return e;
    }
};

inline uint32_t blocksFor(uint32_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t buildTriangleBvh(const TriShade* tris, uint32_t n, uint32_t radius, BvhNode* nodes, float4* geometry,
    uint32_t* order, hipStream_t stream, DeviceBvhBuild* result) {
#define B_HIP(x)                          \
    do {                                  \
        hipError_t e_ = (x);              \
        if (e_ != hipSuccess) return e_;  \
    } while (0)
    *result = DeviceBvhBuild{};
    if (n == 0) return hipSuccess;
    if (radius < 1 || radius > kMaxRadius || n > 0x7fffffffu) return hipErrorInvalidValue;
    hipEvent_t start, end;
    B_HIP(hipEventCreate(&start));
    B_HIP(hipEventCreate(&end));
    struct Events {
        hipEvent_t a, b;
        ~Events() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } events{start, end};
    B_HIP(hipEventRecord(start, stream));
    Scratch s;
    uint32_t blocks = blocksFor(n);
    Box *boxes, *partial, *bounds, *boxA, *boxB;
    uint64_t *keys, *sortedKeys;
    uint32_t *index, *sorted, *idA, *idB, *nn, *mergeFlag, *mergeRank, *keepFlag, *newPos, *leafParent, *nodeParent,
        *leafCount, *preorder, *scalars;
    BvhNode* staged;
    B_HIP(s.get(n, &boxes));
    B_HIP(s.get(blocks, &partial));
    B_HIP(s.get(1, &bounds));
    B_HIP(s.get(n, &keys));
    B_HIP(s.get(n, &sortedKeys));
    B_HIP(s.get(n, &index));
    B_HIP(s.get(n, &sorted));
    B_HIP(s.get(n, &boxA));
    B_HIP(s.get(n, &boxB));
    B_HIP(s.get(n, &idA));
    B_HIP(s.get(n, &idB));
    B_HIP(s.get(n, &nn));
    B_HIP(s.get(n, &mergeFlag));
    B_HIP(s.get(n, &mergeRank));
    B_HIP(s.get(n, &keepFlag));
    B_HIP(s.get(n, &newPos));
    B_HIP(s.get(n, &leafParent));
    B_HIP(s.get(n, &nodeParent));
    B_HIP(s.get(n, &leafCount));
    B_HIP(s.get(n, &preorder));
    B_HIP(s.get(n, &staged));
    B_HIP(s.get(2, &scalars));

    hipLaunchKernelGGL(kPrimBounds, dim3(blocks), dim3(kBlock), 0, stream, tris, n, boxes, partial);
    hipLaunchKernelGGL(kReduceBounds, dim3(1), dim3(kBlock), 0, stream, partial, blocks, bounds);
    hipLaunchKernelGGL(kMorton, dim3(blocks), dim3(kBlock), 0, stream, boxes, n, bounds, keys, index);
    B_HIP(hipGetLastError());
    size_t sortBytes = 0, scanBytes = 0;
    B_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, keys, sortedKeys, index, sorted,
        static_cast<int>(n), 0, 63, stream));
    B_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, mergeFlag, mergeRank, static_cast<int>(n), stream));
    void* temp = nullptr;
    B_HIP(s.get(std::max(sortBytes, scanBytes), reinterpret_cast<uint8_t**>(&temp)));
    size_t bytes = sortBytes;
    B_HIP(hipcub::DeviceRadixSort::SortPairs(temp, bytes, keys, sortedKeys, index, sorted, static_cast<int>(n), 0,
        63, stream));
    hipLaunchKernelGGL(kInitClusters, dim3(blocks), dim3(kBlock), 0, stream, tris, boxes, sorted, n, idA, boxA, order,
        geometry, leafParent);
    if (n > 1) hipLaunchKernelGGL(kFillInvalid, dim3(blocksFor(n - 1)), dim3(kBlock), 0, stream, nodeParent, n - 1);
    B_HIP(hipGetLastError());

    uint32_t count = n, remaining = n - 1, iterations = 0;
    size_t lds = sizeof(float) * 6 * (kBlock + 2 * radius);
    while (count > 1) {
        uint32_t b = blocksFor(count);
        hipLaunchKernelGGL(kNearest, dim3(b), dim3(kBlock), lds, stream, boxA, count, radius, nn);
        hipLaunchKernelGGL(kFlags, dim3(b), dim3(kBlock), 0, stream, nn, count, mergeFlag, keepFlag);
        B_HIP(hipGetLastError());
        bytes = scanBytes;
        B_HIP(hipcub::DeviceScan::ExclusiveSum(temp, bytes, mergeFlag, mergeRank, static_cast<int>(count), stream));
        bytes = scanBytes;
        B_HIP(hipcub::DeviceScan::ExclusiveSum(temp, bytes, keepFlag, newPos, static_cast<int>(count), stream));
        hipLaunchKernelGGL(kTotal, dim3(1), dim3(1), 0, stream, mergeRank, mergeFlag, count, scalars);
        B_HIP(hipGetLastError());
        uint32_t merges = 0;
        B_HIP(hipMemcpyAsync(&merges, scalars, sizeof merges, hipMemcpyDeviceToHost, stream));
        B_HIP(hipStreamSynchronize(stream));
        if (merges == 0 || merges > remaining) return hipErrorUnknown;  // no mutual pair: cannot happen
        hipLaunchKernelGGL(kMerge, dim3(b), dim3(kBlock), 0, stream, count, nn, mergeFlag, mergeRank, keepFlag, newPos,
            remaining - merges, idA, boxA, idB, boxB, nodes, leafParent, nodeParent, leafCount);
        B_HIP(hipGetLastError());
        remaining -= merges;
        count -= merges;
        std::swap(idA, idB);
        std::swap(boxA, boxB);
        ++iterations;
    }
    B_HIP(hipMemsetAsync(scalars + 1, 0, sizeof(uint32_t), stream));
    hipLaunchKernelGGL(kLeafDepth, dim3(blocks), dim3(kBlock), 0, stream, leafParent, nodeParent, n, scalars + 1);
    B_HIP(hipGetLastError());
    if (n > 2) {
        // Clustering numbers nodes by pass; preorder puts each subtree's
        // nodes together, as the host builder does, for cache locality.
        uint32_t inner = n - 1;
        hipLaunchKernelGGL(kPreorder, dim3(blocksFor(inner)), dim3(kBlock), 0, stream, nodes, nodeParent, leafCount,
            inner, preorder);
        hipLaunchKernelGGL(kPermute, dim3(blocksFor(inner)), dim3(kBlock), 0, stream, nodes, preorder, inner, staged);
        B_HIP(hipGetLastError());
        B_HIP(hipMemcpyAsync(nodes, staged, sizeof(BvhNode) * inner, hipMemcpyDeviceToDevice, stream));
    }
    Box root;
    uint32_t depth = 0;
    B_HIP(hipMemcpyAsync(&root, boxA, sizeof root, hipMemcpyDeviceToHost, stream));
    B_HIP(hipMemcpyAsync(&depth, scalars + 1, sizeof depth, hipMemcpyDeviceToHost, stream));
    B_HIP(hipEventRecord(end, stream));
    B_HIP(hipEventSynchronize(end));
    float ms = 0.0f;
    B_HIP(hipEventElapsedTime(&ms, start, end));
    result->depth = depth;
    result->iterations = iterations;
    for (int k = 0; k < 3; ++k) {
        result->rootMin[k] = root.mn[k];
        result->rootMax[k] = root.mx[k];
    }
    result->milliseconds = ms;
    return hipSuccess;
#undef B_HIP
}

}  // namespace wpt
