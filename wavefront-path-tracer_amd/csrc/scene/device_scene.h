// Device-resident scene layout.
//
// Layout choices for HBM (reference scene model: include/scene/scene.cuh:117-161,
// geometry.cuh:6-82):
//  * BVH2 interior nodes are 64 B and carry both children's bounds, so one
//    aligned 64 B load tests both children (the reference's 32 B nodes with
//    own bounds cost three node loads per interior visit).  Children are
//    either interior node indices or kLeafBit|slot.  Single-primitive leaves,
//    like NexusBVH's BVH2 (device_traversal.cuh:254-255).
//  * Triangle positions used by traversal are separate from shading data:
//    three float4 (v0|original index, v1|material, v2|flags) = 48 B per
//    triangle in BVH leaf order, instead of the reference's 228 B AoS record.
//  * Shading attributes (TriShade) stay in the caller's original order and are
//    read once per winning hit.
#pragma once

#include <cstdint>

#include "../common/vec.h"

#if !defined(__HIPCC__)
struct alignas(16) float4 {
    float x, y, z, w;
};
#endif

namespace wpt {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kInvalid = 0xffffffffu;

struct alignas(16) BvhNode {
    float lminx, lminy, lminz, lmaxx;
    float lmaxy, lmaxz, rminx, rminy;
    float rminz, rmaxx, rmaxy, rmaxz;
    uint32_t left, right, pad0, pad1;
};
static_assert(sizeof(BvhNode) == 64, "BvhNode must be one 64 B chunk");

// Eight-wide node of the flat-triangle accelerator WPT_FLAT_WIDE8: the
// bounds of up to eight children as coordinate arrays, then the children
// (interior node index, kLeafBit|slot, or kInvalid for an unused child).
constexpr uint32_t kWideChildren = 8;
struct alignas(16) WideNode {
    float minX[kWideChildren], minY[kWideChildren], minZ[kWideChildren];
    float maxX[kWideChildren], maxY[kWideChildren], maxZ[kWideChildren];
    uint32_t child[kWideChildren];
    uint32_t pad[kWideChildren];
};
static_assert(sizeof(WideNode) == 256, "WideNode must be 256 B");

struct Bvh {
    const BvhNode* nodes;
    uint32_t nodeCount;  // interior nodes
    uint32_t root;       // 0 when nodeCount > 0, else kLeafBit|0
    uint32_t primCount;  // 0: no BVH
    uint32_t pad;
    float rootMin[3];
    float rootMax[3];
};

// Shading-side triangle record, original order (reference Triangle fields used
// by the untextured shading path).
struct TriShade {
    Vec3 v0, v1, v2;
    Vec3 n0, n1, n2;
    Vec2 uv0, uv1, uv2;
    Vec3 tangent, bitangent;
    uint32_t material;
    uint32_t lightIndex;
    uint32_t flags;  // kTriHas*
    uint32_t extra;  // index into DevScene.triExtra, kInvalid: none
};

constexpr uint32_t kTriHasNormals = 1u;
constexpr uint32_t kTriHasTexcoords = 2u;
constexpr uint32_t kTriHasTexcoords1 = 4u;
constexpr uint32_t kTriHasVertexTangents = 8u;

// Second texture-coordinate set and tangent data of a triangle (reference
// Triangle uv1_*, vertexTangent*, vertexTangentSign*, tangent1, bitangent1),
// kept out of TriShade because only textured triangles carry it.
struct TriExtra {
    Vec2 uv1_0, uv1_1, uv1_2;
    Vec3 vertexTangent0, vertexTangent1, vertexTangent2;
    float sign0, sign1, sign2;
    Vec3 tangent1, bitangent1;
};

struct InstanceTransform {
    Vec3 localToWorldX, localToWorldY, localToWorldZ;
    Vec3 worldToLocalX, worldToLocalY, worldToLocalZ;
    Vec3 translation;
};

// transformVector / inverseTransformVector / transformNormal, reference
// include/scene/scene.cuh:38-55.
WPT_HD Vec3 transformVector(const InstanceTransform& t, const Vec3& v) {
    return t.localToWorldX * v.x + t.localToWorldY * v.y + t.localToWorldZ * v.z;
}
WPT_HD Vec3 inverseTransformVector(const InstanceTransform& t, const Vec3& v) {
    return t.worldToLocalX * v.x + t.worldToLocalY * v.y + t.worldToLocalZ * v.z;
}
WPT_HD Vec3 transformNormal(const InstanceTransform& t, const Vec3& n) {
    return normalize(Vec3(dot(t.worldToLocalX, n), dot(t.worldToLocalY, n), dot(t.worldToLocalZ, n)));
}

struct Instance {
    InstanceTransform transform;
    uint32_t mesh;
    uint32_t lightOffset;
    uint32_t materialOverride;
    uint32_t pad;
};

struct Blas {
    Bvh bvh;
    const float4* geometry;     // BVH slot order, 3 float4 per triangle
    const TriShade* triangles;  // original order
    uint32_t triangleCount;
    uint32_t pad[3];
};

struct Material {
    uint32_t type;
    Vec3 albedo;
    Vec3 emission;
    float roughness;
    float ior;
    float metallic;
    float opacity;
    float alphaCutoff;
    uint32_t alphaMasked;
    Vec3 attenuationColor;
    float attenuationDistance;
    float volumeDensity;
    Vec3 specular;
    float glossiness;
    uint32_t ext;      // index into DevScene.materialExt, kInvalid: untextured
    float dispersion;  // IOR step per spectral channel of caustic photons
};

// Reference TextureBinding (material.cuh:14-21).
struct TexBinding {
    uint32_t texture;  // kInvalid: none
    uint32_t texCoord;
    Vec2 offset;
    Vec2 scale;
    float rotation;
    uint32_t hasTransform;
};

// Texture bindings and glTF workflow fields (the rest of reference Material).
struct MaterialExt {
    TexBinding albedo, metallicRoughness, emissive, normal, thickness, specularGlossiness, specularWeight,
        specularColor, clearcoat, clearcoatRoughness, clearcoatNormal;
    float normalScale;
    uint32_t specularGlossinessWorkflow;
    uint32_t gltfPbrWorkflow;
    float specularWeightFactor;
    Vec3 specularColorFactor;
    float clearcoatFactor;
    float clearcoatRoughnessFactor;
    float clearcoatNormalScale;
    float dispersion;
};

// RGBA8 texels packed r | g << 8 | b << 16 | a << 24, rows top to bottom.
struct DevTexture {
    const uint32_t* texels;
    uint32_t width, height;
};

enum MaterialType : uint32_t {
    kDiffuse = 0,
    kMetal = 1,
    kDielectric = 2,
    kEmissive = 3,
    kSpecularGlossiness = 4,
};

struct TriangleLight {
    Vec3 v0, v1, v2;
    Vec3 normal;
    float area;
    uint32_t material;
    float selectionPdf;
    uint32_t pad;
};

struct LightAlias {
    float probability;
    uint32_t alias;
};

// Alpha-mask policy of a material in traversal (reference
// intersectTriangleHit, device_traversal.cuh:148-151): opaque, never hit
// (masked, untextured, opacity below the cutoff), or tested per hit against
// opacity x albedo-texture alpha.
constexpr uint8_t kCutNever = 0;
constexpr uint8_t kCutAlways = 1;
constexpr uint8_t kCutTexture = 2;

// Reference PunctualLightType and PunctualLight (punctual_light.cuh:6-16);
// the values are the C ABI's WPT_LIGHT_*.
enum PunctualLightType : uint32_t {
    kLightPoint = 0,
    kLightSpot = 1,
    kLightDirectional = 2,
};

struct PunctualLight {
    uint32_t type;  // PunctualLightType
    Vec3 position;
    Vec3 direction;  // of emission
    Vec3 intensity;
    float range;     // 0: unlimited
    float innerConeCos;
    float outerConeCos;
};

struct EnvAlias {
    float probability;
    uint32_t alias;
};

// Equirectangular environment (texels == nullptr: none).
struct EnvironmentMap {
    const float4* texels;         // width * height RGB(+1), row 0 at +Y
    const float* probabilities;   // importance per texel, sums to 1
    const EnvAlias* alias;
    uint32_t width, height;
    float intensity;
    float rotation;               // radians about +Y
};

// Passed by value to every kernel (reference Scene, scene.cuh:117-161).
struct DevScene {
    const float4* spheres;  // center xyz, radius
    const uint32_t* sphereMaterial;
    uint32_t sphereCount;
    uint32_t directSphereCount;
    const uint32_t* directSpheres;
    const uint32_t* directInstances;
    uint32_t directInstanceCount;
    uint32_t staticCount;
    const TriShade* staticTriangles;
    Bvh tlas;
    const uint32_t* tlasLeaves;  // slot -> (1<<31 | sphere) or instance
    const Instance* instances;
    const Blas* blases;
    uint32_t instanceCount;
    uint32_t blasCount;
    Bvh flat;
    const WideNode* flatWide;  // non-null: the flat triangles use eight-wide nodes
    const float4* flatGeometry;
    const TriShade* flatTriangles;
    uint32_t flatCount;
    uint32_t materialCount;
    const Material* materials;
    const uint8_t* materialCutout;  // kCut*: alpha-mask policy per material
    const TriangleLight* lights;
    const LightAlias* lightAlias;
    uint32_t lightCount;
    uint32_t blackBackground;
    uint32_t ignoreDirectLightOcclusion;
    uint32_t pad;
    EnvironmentMap environment;
    const PunctualLight* punctual;  // directional lights first
    uint32_t punctualCount;
    uint32_t directionalCount;
    const MaterialExt* materialExt;
    const DevTexture* textures;
    const TriExtra* triExtra;
    uint32_t textureCount;
    uint32_t useNormalMaps;
    float normalMapMinimumCosine;
    uint32_t fullFeatures;  // environment map, punctual lights or texture bindings present
};

// Shading traces its shadow rays where it makes them (rather than deferring
// the area-light ray to a later pass) when radiance is added after the
// area-light NEE within a bounce: environment and punctual NEE, and the
// dielectric environment splits.
WPT_HD bool shadesWithInlineShadows(const DevScene& s) { return s.environment.texels || s.punctualCount; }

constexpr uint32_t kTlasSphereBit = 0x80000000u;

}  // namespace wpt
