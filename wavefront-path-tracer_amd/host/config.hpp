// Render and scene configuration, command line and scene manifests.
//
// Mirrors the reference interface: RenderConfig / SceneConfig
// (include/config.hpp:8-63), parseCommandLine / parseCommandLineText /
// scenePreset / parseScenePreset (src/config.cpp:91-319) and
// loadSceneManifest (src/scene/scene_manifest.cpp:79-224).  Distributed
// (coordinator/worker) options parse as in the reference but this renderer
// runs them as locally leased tasks on one GPU.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wpt.h"

namespace wpt {

enum class ScenePreset {
    Sponza, Cornell, Hurricane, Prism, Crystal, Deer, Demo, RtWeek,
    // Procedural scenes of the measured configurations (no reference preset).
    CornellBox,   // unit-scale Cornell box, 2 spheres, emissive quad
    SphereField,  // 484-sphere field + 3 instances of a displaced icosphere
    DiffCornell,  // Cornell box tessellated for the differentiable renderer
    Terrain,      // textured terrain, instanced props, 512 lamps, HDR sky, crystals with caustics
};

enum class DistributedRole { Local, Coordinator, Worker };
enum class CornellObjectSource { Obj, Gltf, ProceduralPrism };
enum class CornellLightProfile { Standard, Prism, Crystal };
enum class ObjAccelerationPolicy { TlasBlasBvh2, NexusBvh2, NexusBvh8, MedianBvh, SpatialSplitBvh };

struct Vec3f {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};

struct RenderConfig {
    uint32_t width = 1920, height = 1080, tileWidth = 960, tileHeight = 540;
    uint32_t maxDepth = 32, russianRouletteStartDepth = 12, samplesPerPixel = 512, blockSize = 256;
    bool tiledRendering = false, intersectionDebug = false, shadingNormalDebug = false, resumeFromCheckpoint = false;
    bool persistentWavefront = false, profileQueues = false;
    bool regenerate = false;  // extension: persistent kernel with path regeneration
    std::string checkpointFilename = "render.checkpoint";
    uint32_t checkpointProgressPercent = 25, checkpointBufferCount = 3;
    std::string outputFilename = "render.ppm";
    bool enableCaustics = false, enableCausticGather = true;
    uint32_t causticPhotonCount = 1048576, causticMaxDepth = 8;
    float causticGatherRadius = 0.06f;
    float indirectRadianceClamp = 10.0f;
    DistributedRole distributedRole = DistributedRole::Local;
    std::string listenHost = "0.0.0.0";
    uint16_t listenPort = 9000;
    std::string coordinatorHost = "127.0.0.1";
    uint16_t coordinatorPort = 9000;
    std::string workerId;
    std::string distributedCheckpointFilename = "render.distributed.checkpoint";
    uint32_t distributedCheckpointIntervalSeconds = 30, distributedSamplesPerTask = 16;
    uint64_t distributedLeaseDurationMs = 30000;
    bool pushAssets = false;
    std::string assetCacheDirectory = ".pathtracer-assets";
    bool pushConfigurationToWorkers = false;
    uint32_t rngStrategy = WPT_RNG_XORSHIFT32;
    uint32_t persistentBlocks = 0;
};

struct SceneConfig {
    ScenePreset preset = ScenePreset::Sponza;
    std::string name = "sponza";
    std::string objectFilename = "../assets/deer-obj.obj";
    std::string gltfFilename = "../assets/sponza/Sponza.gltf";
    float gltfScale = 1.0f, objectScale = 0.1f;
    Vec3f objectTranslation{0.181f, -0.906f, -0.2f};
    CornellObjectSource objectSource = CornellObjectSource::Obj;
    CornellLightProfile lightProfile = CornellLightProfile::Standard;
    ObjAccelerationPolicy acceleration = ObjAccelerationPolicy::TlasBlasBvh2;
    bool neutralRoom = false, removeBackdrop = false, convertObjectMaterialsToDielectric = false;
    int32_t objectMaterialOverride = -1;
    bool addSponzaTopLight = true, ignoreSponzaLightOcclusion = false, useNormalMaps = true;
    float normalMapMinimumCosine = 0.0f;
    std::string environmentFilename;
    float environmentIntensity = 1.0f, environmentRotationDegrees = 0.0f;
    // Procedural generators.
    uint32_t seed = 1;
    uint32_t icosphereFrequency = 224;  // 20 * f^2 triangles
    uint32_t bvhBuilder = WPT_BVH_BUILDER_HOST_SAH;  // mesh BLAS builder (--bvh-builder)
    // Terrain preset: grid cells per side (2 r^2 triangles; rocks use
    // icosphereFrequency / 2) and the edge of its 16 square RGBA8 textures.
    uint32_t terrainResolution = 708;
    uint32_t textureSize = 4096;
};

const char* presetName(ScenePreset preset);
bool parseScenePreset(const char* value, ScenePreset& preset);
// Preset defaults, overlaid by assets/scenes/<name>.json when one is found
// next to the working directory or the library.
SceneConfig scenePreset(ScenePreset preset);
bool loadSceneManifest(const char* filename, SceneConfig& scene, std::string* error);
// Returns false when the program should stop (help, errors).
bool parseCommandLine(int argc, char** argv, RenderConfig& render, SceneConfig& scene);
bool parseCommandLineText(const std::string& command, RenderConfig& render, SceneConfig& scene,
    std::vector<std::string>& argumentStorage);

// The C ABI render configuration for a RenderConfig.  Caustic photons are
// traced only for the caustic presets (Cornell, Hurricane, Prism, Crystal),
// with one spectral channel per photon for Prism and Crystal
// (render_driver.cu:37-40, 180-196).
wpt_render_config toRenderConfig(const RenderConfig& c, ScenePreset preset);

// 64-bit FNV-1a over the scene (with its name and the bytes of its asset
// files) and the output-affecting render settings, in the field order of
// src/renderer/render_fingerprint.cpp:82-163: the key of checkpoints and
// tasks.
uint64_t computeRenderFingerprint(const RenderConfig& render, const SceneConfig& scene);
// The scene part alone, without the name (computeSceneFingerprint,
// render_fingerprint.cpp:157-161).
uint64_t computeSceneFingerprint(const SceneConfig& scene);

}  // namespace wpt
