// Host-side scene build: BVH2 construction, AccelerationMode partition and the
// triangle-light table, producing the arrays that are uploaded to HBM.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/wpt.h"
#include "device_scene.h"

namespace wpt {

struct Aabb {
    float mn[3];
    float mx[3];
};

// A BVH2 with single-primitive leaves in the 64 B child-bounds layout.
// order[slot] is the primitive stored at leaf slot `slot`.
struct HostBvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> order;
    std::vector<Aabb> primBounds;  // original primitive bounds (for export)
    std::vector<Aabb> slotBounds;  // per-slot bounds when they differ from primBounds (spatial splits)
    uint32_t root = 0;
    Aabb rootBounds{};
    uint32_t depth = 0;  // longest root-to-leaf edge count
    uint32_t primCount() const { return static_cast<uint32_t>(order.size()); }
};

// Binned-SAH build over primitive bounds.  The node numbering is fixed by the
// primitive partition alone (a subtree over n primitives owns n-1 consecutive
// interior nodes), so the result is identical for any thread count.
// `median` builds the median-split tree instead (object median on the
// longest centroid axis).
HostBvh buildBvh2(const std::vector<Aabb>& primBounds, unsigned threads = 0, bool median = false);

// Spatial-split BVH over triangles: triangles crossing a chosen plane are
// referenced on both sides, bounded by their clipped parts, so a triangle
// can occupy several leaf slots (order[slot] names it; slotBounds holds the
// clipped bounds).
HostBvh buildSpatialSplitBvh2(const std::vector<TriShade>& tris);

// Eight-wide nodes collapsed from a binary tree: each wide node takes the
// children of a binary node and keeps opening its largest interior child
// until it has eight.  Leaves keep the binary tree's slots.  *stackNeed
// receives the deepest traversal stack the wide walk can need.
std::vector<WideNode> collapseWide(const HostBvh& bvh, uint32_t* stackNeed);

// Export in the reference node convention (NXB::BVH2: own bounds per node,
// leaf iff left == 0xffffffff, root = count - 1).
std::vector<wpt_bvh2_node> exportBvh2(const HostBvh& bvh);

Aabb triangleBounds(const Vec3& a, const Vec3& b, const Vec3& c);

struct HostMesh {
    std::vector<TriShade> triangles;  // mesh-local, original order
    HostBvh bvh;                      // empty unless an instance uses AccelerationMode::Bvh
    std::vector<float4> geometry;     // slot order
    bool buildOnDevice = false;       // BVH left to the device builder
};

// Everything uploadScene (reference src/scene/scene.cu:336-389) resolves on
// the host, in device layout.
struct HostSceneData {
    std::vector<float4> spheres;
    std::vector<uint32_t> sphereMaterial;
    std::vector<uint32_t> directSpheres;
    std::vector<uint32_t> directInstances;
    std::vector<uint32_t> tlasLeavesOriginal;  // reference SceneLeaf order
    std::vector<uint32_t> tlasLeavesSlot;      // BVH slot order
    HostBvh tlas;
    std::vector<Instance> instances;
    std::vector<HostMesh> meshes;
    std::vector<TriShade> staticTriangles;
    std::vector<TriShade> flatTriangles;
    HostBvh flat;
    std::vector<float4> flatGeometry;
    uint32_t flatAccelerator = 0;    // WPT_FLAT_*
    std::vector<WideNode> flatWide;  // WPT_FLAT_WIDE8: nodes over flat's slots
    uint32_t flatWideStack = 0;
    std::vector<Material> materials;
    std::vector<uint8_t> materialCutout;
    std::vector<TriangleLight> lights;
    std::vector<LightAlias> lightAlias;
    std::vector<float> lightWeights;
    std::vector<uint32_t> meshLightIndex;  // concatenated mesh triangles
    uint32_t blackBackground = 0;
    uint32_t ignoreDirectLightOcclusion = 0;
    // Environment map (empty texels: none).
    std::vector<float4> envTexels;
    std::vector<float> envProbabilities;
    std::vector<EnvAlias> envAlias;
    uint32_t envWidth = 0, envHeight = 0;
    float envIntensity = 1.0f, envRotation = 0.0f;
    std::vector<PunctualLight> punctual;  // directional lights first
    uint32_t directionalCount = 0;
    // Textures (RGBA8 packed per texel), texture bindings of the textured
    // materials, and the second-set / tangent data of textured triangles.
    std::vector<std::vector<uint32_t>> textureTexels;
    std::vector<uint32_t> textureWidth, textureHeight;
    std::vector<MaterialExt> materialExt;
    std::vector<TriExtra> triExtra;
    uint32_t useNormalMaps = 0;
    float normalMapMinimumCosine = 0.0f;
    // An environment map, punctual lights or texture bindings: the scene
    // needs the full-feature kernels.
    bool fullFeatures() const {
        return !envTexels.empty() || !punctual.empty() || !materialExt.empty() || !textureTexels.empty();
    }
    uint32_t maxStackDepth = 0;  // worst-case traversal stack use
    uint32_t maxBlasDepth = 0;   // of the BLASes built so far
    uint32_t bvhBuilder = 0;     // WPT_BVH_BUILDER_*
    uint32_t bvhSearchRadius = 16;
    double blasBuildMilliseconds = 0.0;
};

// Returns an empty string on success, else the validation error.
std::string buildHostScene(const wpt_scene_desc& desc, HostSceneData& out, unsigned threads = 0);

// Worst-case traversal stack use of the TLAS over the deepest BLAS, or of
// the flat BVH.
uint32_t stackDepth(const HostSceneData& s);

}  // namespace wpt
