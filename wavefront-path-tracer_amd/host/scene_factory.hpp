// Scene assembly for presets and procedural scenes (reference
// src/scene/scene_host.cpp, src/scene/scene_factory.cpp, src/scene/camera.cu).
#pragma once

#include <string>
#include <vector>

#include "../../include/wpt.h"
#include "config.hpp"

namespace wpt {

// Host-owned scene arrays behind a wpt_scene_desc (reference HostScene,
// include/scene/scene_host.hpp:19-33).
struct HostScene {
    std::vector<wpt_sphere> spheres;
    std::vector<wpt_triangle> staticTriangles;
    std::vector<wpt_triangle> flatTriangles;
    std::vector<wpt_triangle> meshTriangles;
    std::vector<wpt_mesh> meshes;
    std::vector<wpt_instance> instances;
    std::vector<wpt_material> materials;
    std::vector<wpt_punctual_light> punctualLights;
    bool blackBackground = false;
    bool ignoreDirectLightOcclusion = false;
    // Textured scenes: RGBA8 texels of each texture (textures[i].rgba points
    // into textureData[i]), per-material bindings (empty: none), the
    // equirectangular environment (empty: none) and the normal-map switch.
    std::vector<std::vector<uint8_t>> textureData;
    std::vector<wpt_texture> textures;
    std::vector<wpt_material_ext> materialExt;
    std::vector<float> environment;
    uint32_t environmentWidth = 0, environmentHeight = 0;
    float environmentIntensity = 1.0f, environmentRotationDegrees = 0.0f;
    bool useNormalMaps = true;
    float normalMapMinimumCosine = 0.0f;
    uint32_t flatAccelerator = WPT_FLAT_SAH_BVH2;  // tree over flatTriangles

    wpt_scene_desc desc() const;
};

wpt_material makeMaterial(uint32_t type, float r, float g, float b, float er = 0.0f, float eg = 0.0f, float eb = 0.0f);
// No textures, unit scales and weights, no dispersion (material.cuh defaults).
wpt_material_ext plainMaterialExt();
wpt_triangle makeTriangle(const float v0[3], const float v1[3], const float v2[3], uint32_t material);
wpt_transform makeScaledTransform(float tx, float ty, float tz, float scale);

// Displaced geodesic icosphere: 20 * frequency^2 triangles, smooth normals.
std::vector<wpt_triangle> displacedIcosphere(uint32_t frequency, uint32_t seed, float amplitude);

// Shading-heavy terrain with caustics: a noise-displaced terrain of
// 2 r^2 triangles, four displaced rock meshes (icosphereFrequency / 2),
// alpha-masked foliage cards and dielectric crystals with Beer attenuation
// and dispersion as instanced props, 256 lamp quads (512 emissive triangles,
// log-normal radiance), 16 square RGBA8 noise textures (albedo, normal,
// alpha and metallic-roughness), and a log-normal sky with a sun lobe as an
// HDR environment.  Seeded by c.seed.
void terrainScene(HostScene& s, const SceneConfig& c);

// Returns false with *error set for presets that need external assets.
bool createSceneFromConfig(const SceneConfig& config, HostScene& out, std::string* error);
wpt_camera createSceneCamera(const SceneConfig& config, uint32_t width, uint32_t height);

}  // namespace wpt
