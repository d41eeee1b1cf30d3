// BVH2 construction on the device for triangle meshes: parallel locally
// ordered clustering (PLOC; Meister and Bittner, "Parallel Locally-Ordered
// Clustering for Bounding Volume Hierarchy Construction", IEEE TVCG 2018)
// over 63-bit Morton codes of the triangle centroids.
//
// The reference builds its BLASes on the GPU with NexusBVH
// (src/scene/scene.cu:123-239); that library is not available, so this is a
// builder of its own.  Its output has the layout of the host's binned-SAH
// builder: interior nodes with both children's bounds, root at node 0,
// single-triangle leaves kLeafBit|slot, and the triangles' positions in
// slot order.  Every step is a deterministic function of the input (stable
// radix sort, ties broken toward the lower index, node numbers from prefix
// sums), so repeated builds give the same tree.
#pragma once

#include <hip/hip_runtime.h>

#include "../scene/device_scene.h"

namespace wpt {

struct DeviceBvhBuild {
    uint32_t depth;       // longest root-to-leaf edge count
    uint32_t iterations;  // clustering passes
    float rootMin[3];
    float rootMax[3];
    float milliseconds;   // device time of the whole build
};

// Builds over n triangles given in original order on the device.  Outputs
// (device memory, caller-allocated): nodes[n-1], geometry[3n] (v0 | original
// index, v1 | material, v2 | flags per slot), order[n] (slot -> original
// index).  radius is PLOC's neighbour search radius (1..64).
hipError_t buildTriangleBvh(const TriShade* triangles, uint32_t n, uint32_t radius, BvhNode* nodes,
    float4* geometry, uint32_t* order, hipStream_t stream, DeviceBvhBuild* result);

}  // namespace wpt
