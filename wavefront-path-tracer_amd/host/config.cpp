#include "config.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include "json.hpp"

namespace wpt {

namespace {

void printHelp() {
    std::cout << "Usage: pathtracer [options]\n\n"
              << "Options:\n"
              << "  --help, -h              Show this help text\n"
              << "  --scene NAME             Select a scene preset\n"
              << "                           NAME: sponza, cornell, hurricane, prism, crystal, deer, rtweek, demo,\n"
              << "                                 cornell-box, sphere-field, diff-cornell\n"
              << "  --scene-file FILE        Load a scene manifest JSON file\n"
              << "  --list-scenes            List available scene presets\n"
              << "  --caustics               Enable photon-mapped caustics (Cornell/Hurricane/Prism/Crystal)\n"
              << "  --caustic-photons N      Set the photon count (default: 1048576)\n"
              << "  --caustic-radius R       Set the gather radius (default: 0.06)\n"
              << "  --no-caustics            Disable caustic photon tracing\n"
              << "  --no-caustic-gather      Trace photons but omit camera gathering\n"
              << "  --persistent-wavefront  Use the persistent cooperative wavefront kernel\n"
              << "  --regenerate             Use the persistent kernel with path regeneration (no bounce barriers)\n"
              << "  --bvh-builder host|device  Build mesh BLASes with host binned SAH (default) or on the GPU\n"
              << "  --profile-queues        Print queue occupancy by bounce\n"
              << "  --resume                 Resume from the selected checkpoint\n"
              << "  --coordinator            Run as a distributed coordinator (local task leases)\n"
              << "  --listen-host HOST       Coordinator bind host (default: 0.0.0.0)\n"
              << "  --listen-port PORT       Coordinator bind port (default: 9000)\n"
              << "  --worker                 Run as a distributed worker (local task leases)\n"
              << "  --worker-id ID           Unique worker identifier\n"
              << "  --coordinator-host HOST  Worker coordinator host (default: 127.0.0.1)\n"
              << "  --coordinator-port PORT  Worker coordinator port (default: 9000)\n"
              << "  --distributed-checkpoint FILE\n"
              << "                           Distributed checkpoint path\n"
              << "  --distributed-checkpoint-seconds N\n"
              << "                           Coordinator checkpoint interval (default: 30)\n"
              << "  --samples-per-task N     Distributed sample batch size (default: 16)\n"
              << "  --distributed-lease-seconds N\n"
              << "                           Worker lease duration (default: 30)\n"
              << "  --push-assets            Transfer scene assets to distributed workers\n"
              << "  --asset-cache DIR        Worker asset cache directory\n"
              << "  --rng STRATEGY           RNG strategy: xorshift32 or pcg32\n"
              << "  --clamp-indirect VALUE   Limit indirect sample RGB mean (default: 10; 0 disables)\n"
              << "Extensions:\n"
              << "  --resolution WxH         Image size (default: 1920x1080)\n"
              << "  --spp N                  Samples per pixel (default: 512)\n"
              << "  --max-depth N            Maximum path depth (default: 32)\n"
              << "  --rr-depth N             Russian roulette start depth (default: 12)\n"
              << "  --tile WxH               Render in tiles of this size\n"
              << "  --block-size N           Workgroup size, a multiple of 64 (default: 256)\n"
              << "  --persistent-blocks N    Persistent grid size (default: occupancy)\n"
              << "  --output FILE            Output PPM (default: render.ppm)\n"
              << "  --seed N                 Seed of procedural scenes (default: 1)\n"
              << "  --intersection-debug     Shade by surface normal\n";
}

bool parseUint(const char* value, uint32_t& out, bool positive) {
    if (value == nullptr || value[0] == '\0') return false;
    char* end = nullptr;
    unsigned long long v = std::strtoull(value, &end, 10);
    if (*end != '\0' || v > 0xffffffffull || (positive && v == 0)) return false;
    out = static_cast<uint32_t>(v);
    return true;
}

bool parsePort(const char* value, uint16_t& port) {
    uint32_t v = 0;
    if (!parseUint(value, v, true) || v > 65535) return false;
    port = static_cast<uint16_t>(v);
    return true;
}

bool parseSize(const char* value, uint32_t& w, uint32_t& h) {
    if (!value) return false;
    std::string s(value);
    size_t x = s.find('x');
    if (x == std::string::npos) return false;
    return parseUint(s.substr(0, x).c_str(), w, true) && parseUint(s.substr(x + 1).c_str(), h, true);
}

std::string resolveAssetPath(const std::filesystem::path& manifest, const std::string& file) {
    if (file.empty() || file.rfind("__", 0) == 0) return file;
    std::filesystem::path p(file);
    if (p.is_absolute()) return p.lexically_normal().string();
    return (manifest.parent_path() / p).lexically_normal().string();
}

bool readVec3(const Json& v, Vec3f& out, const char* field, std::string* error) {
    if (!v.isArray() || v.items().size() != 3 || !v.items()[0].isNumber() || !v.items()[1].isNumber() ||
        !v.items()[2].isNumber()) {
        if (error) *error = std::string("scene manifest field '") + field + "' must be an array of three numbers";
        return false;
    }
    out.x = static_cast<float>(v.items()[0].number());
    out.y = static_cast<float>(v.items()[1].number());
    out.z = static_cast<float>(v.items()[2].number());
    return true;
}

bool parseObjectSource(const std::string& v, CornellObjectSource& s) {
    if (v == "obj") s = CornellObjectSource::Obj;
    else if (v == "gltf" || v == "glb") s = CornellObjectSource::Gltf;
    else if (v == "procedural_prism") s = CornellObjectSource::ProceduralPrism;
    else return false;
    return true;
}

bool parseLightProfile(const std::string& v, CornellLightProfile& p) {
    if (v == "standard") p = CornellLightProfile::Standard;
    else if (v == "prism") p = CornellLightProfile::Prism;
    else if (v == "crystal") p = CornellLightProfile::Crystal;
    else return false;
    return true;
}

bool parseAcceleration(const std::string& v, ObjAccelerationPolicy& a) {
    if (v == "tlas_blas_bvh2") a = ObjAccelerationPolicy::TlasBlasBvh2;
    else if (v == "nexus_bvh2") a = ObjAccelerationPolicy::NexusBvh2;
    else if (v == "nexus_bvh8") a = ObjAccelerationPolicy::NexusBvh8;
    else if (v == "median_bvh") a = ObjAccelerationPolicy::MedianBvh;
    else if (v == "spatial_split_bvh") a = ObjAccelerationPolicy::SpatialSplitBvh;
    else return false;
    return true;
}

// Built-in defaults of each preset (the reference keeps them in manifests).
SceneConfig presetDefaults(ScenePreset preset) {
    SceneConfig c;
    c.preset = preset;
    c.name = presetName(preset);
    switch (preset) {
    case ScenePreset::Prism:
        c.objectSource = CornellObjectSource::ProceduralPrism;
        c.objectFilename = "__caustic_prism__";
        c.objectScale = 1.0f;
        c.objectTranslation = Vec3f{0.0f, 0.0f, 0.0f};
        c.convertObjectMaterialsToDielectric = true;
        c.lightProfile = CornellLightProfile::Prism;
        break;
    case ScenePreset::RtWeek:
        c.objectFilename.clear();
        c.useNormalMaps = false;
        break;
    case ScenePreset::Demo:
    case ScenePreset::CornellBox:
    case ScenePreset::SphereField:
    case ScenePreset::DiffCornell:
    case ScenePreset::Terrain:
        c.objectFilename.clear();
        break;
    default:
        break;
    }
    return c;
}

}  // namespace

const char* presetName(ScenePreset preset) {
    switch (preset) {
    case ScenePreset::Sponza: return "sponza";
    case ScenePreset::Cornell: return "cornell";
    case ScenePreset::Hurricane: return "hurricane";
    case ScenePreset::Prism: return "prism";
    case ScenePreset::Crystal: return "crystal";
    case ScenePreset::Deer: return "deer";
    case ScenePreset::Demo: return "demo";
    case ScenePreset::RtWeek: return "rtweek";
    case ScenePreset::CornellBox: return "cornell-box";
    case ScenePreset::SphereField: return "sphere-field";
    case ScenePreset::DiffCornell: return "diff-cornell";
    case ScenePreset::Terrain: return "terrain";
    }
    return "unknown";
}

bool parseScenePreset(const char* v, ScenePreset& p) {
    static const ScenePreset all[] = {ScenePreset::Sponza, ScenePreset::Cornell, ScenePreset::Hurricane,
        ScenePreset::Prism, ScenePreset::Crystal, ScenePreset::Deer, ScenePreset::Demo, ScenePreset::RtWeek,
        ScenePreset::CornellBox, ScenePreset::SphereField, ScenePreset::DiffCornell, ScenePreset::Terrain};
    for (ScenePreset candidate : all) {
        if (!std::strcmp(v, presetName(candidate))) {
            p = candidate;
            return true;
        }
    }
    return false;
}

SceneConfig scenePreset(ScenePreset preset) {
    SceneConfig result = presetDefaults(preset);
    const std::string manifest = std::string(presetName(preset)) + ".json";
    const std::string candidates[] = {"assets/scenes/" + manifest, "../assets/scenes/" + manifest,
        "wavefront-path-tracer_amd/assets/scenes/" + manifest};
    for (const std::string& candidate : candidates) {
        std::ifstream in(candidate);
        if (!in) continue;
        std::string error;
        if (!loadSceneManifest(candidate.c_str(), result, &error))
            std::cerr << "Could not load scene manifest '" << candidate << "': " << error << '\n';
        break;
    }
    return result;
}

bool loadSceneManifest(const char* filename, SceneConfig& scene, std::string* error) {
    auto setError = [&](const std::string& m) {
        if (error) *error = m;
        return false;
    };
    if (filename == nullptr || filename[0] == '\0') return setError("scene manifest filename is empty");
    std::ifstream input(filename);
    if (!input) return setError(std::string("could not open scene manifest: ") + filename);
    std::stringstream buffer;
    buffer << input.rdbuf();
    try {
        Json root = Json::parse(buffer.str());
        if (!root.isObject()) return setError("scene manifest root must be an object");
        if (!root.contains("preset") || !root.at("preset").isString())
            return setError("scene manifest requires a string 'preset' field");
        ScenePreset preset;
        const std::string presetValue = root.at("preset").str();
        if (!parseScenePreset(presetValue.c_str(), preset))
            return setError("scene manifest contains an unknown preset: " + presetValue);
        SceneConfig loaded = presetDefaults(preset);
        loaded.name = root.value("name", std::string(presetName(preset)));
        std::filesystem::path manifestPath(filename);
        if (root.contains("object")) {
            const Json& object = root.at("object");
            if (!object.isObject()) return setError("scene manifest field 'object' must be an object");
            const std::string source = object.value("source", std::string("obj"));
            if (!parseObjectSource(source, loaded.objectSource))
                return setError("scene manifest contains an unknown object source: " + source);
            if (object.contains("file")) {
                if (!object.at("file").isString()) return setError("scene manifest object.file must be a string");
                std::string path = resolveAssetPath(manifestPath, object.at("file").str());
                loaded.objectFilename = path;
                loaded.gltfFilename = path;
            }
            loaded.objectScale = static_cast<float>(object.value("scale", static_cast<double>(loaded.objectScale)));
            if (object.contains("translation") &&
                !readVec3(object.at("translation"), loaded.objectTranslation, "object.translation", error))
                return false;
            loaded.convertObjectMaterialsToDielectric =
                object.value("dielectric", loaded.convertObjectMaterialsToDielectric);
            loaded.objectMaterialOverride = object.value("material_override", static_cast<int>(loaded.objectMaterialOverride));
        }
        loaded.gltfScale = static_cast<float>(root.value("gltf_scale", static_cast<double>(loaded.gltfScale)));
        if (root.contains("environment")) {
            const Json& env = root.at("environment");
            if (!env.isObject() || !env.contains("file") || !env.at("file").isString() || env.at("file").str().empty())
                return setError("scene environment requires a nonempty file path");
            for (const auto& kv : env.members())
                if (kv.first != "file" && kv.first != "intensity" && kv.first != "rotation_degrees")
                    return setError("unknown scene environment field: " + kv.first);
            loaded.environmentFilename = resolveAssetPath(manifestPath, env.at("file").str());
            for (const char* key : {"intensity", "rotation_degrees"})
                if (env.contains(key) && !env.at(key).isNumber())
                    return setError(std::string("environment.") + key + " must be numeric");
            loaded.environmentIntensity = static_cast<float>(env.value("intensity", 1.0));
            loaded.environmentRotationDegrees = static_cast<float>(env.value("rotation_degrees", 0.0));
            if (!std::isfinite(loaded.environmentIntensity) || loaded.environmentIntensity < 0 ||
                !std::isfinite(loaded.environmentRotationDegrees))
                return setError("environment intensity must be finite/nonnegative and rotation finite");
        }
        if (root.contains("room")) {
            const Json& room = root.at("room");
            if (!room.isObject()) return setError("scene manifest field 'room' must be an object");
            loaded.neutralRoom = room.value("neutral", loaded.neutralRoom);
            loaded.removeBackdrop = room.value("remove_backdrop", loaded.removeBackdrop);
        }
        if (root.contains("lighting")) {
            const Json& lighting = root.at("lighting");
            if (!lighting.isObject()) return setError("scene manifest field 'lighting' must be an object");
            const std::string profile = lighting.value("profile", std::string("standard"));
            if (!parseLightProfile(profile, loaded.lightProfile))
                return setError("scene manifest contains an unknown light profile: " + profile);
            loaded.addSponzaTopLight = lighting.value("add_sponza_top_light", loaded.addSponzaTopLight);
            loaded.ignoreSponzaLightOcclusion =
                lighting.value("ignore_sponza_light_occlusion", loaded.ignoreSponzaLightOcclusion);
        }
        if (root.contains("normal_maps")) {
            const Json& nm = root.at("normal_maps");
            if (!nm.isObject()) return setError("scene manifest field 'normal_maps' must be an object");
            loaded.useNormalMaps = nm.value("enabled", loaded.useNormalMaps);
            loaded.normalMapMinimumCosine =
                static_cast<float>(nm.value("minimum_cosine", static_cast<double>(loaded.normalMapMinimumCosine)));
        }
        if (root.contains("acceleration")) {
            if (!root.at("acceleration").isString() || !parseAcceleration(root.at("acceleration").str(), loaded.acceleration))
                return setError("scene manifest contains an unknown acceleration policy");
        }
        if (root.contains("procedural")) {
            const Json& p = root.at("procedural");
            if (!p.isObject()) return setError("scene manifest field 'procedural' must be an object");
            loaded.seed = static_cast<uint32_t>(p.value("seed", static_cast<double>(loaded.seed)));
            loaded.icosphereFrequency =
                static_cast<uint32_t>(p.value("icosphere_frequency", static_cast<double>(loaded.icosphereFrequency)));
            loaded.terrainResolution =
                static_cast<uint32_t>(p.value("terrain_resolution", static_cast<double>(loaded.terrainResolution)));
            loaded.textureSize = static_cast<uint32_t>(p.value("texture_size", static_cast<double>(loaded.textureSize)));
            if (loaded.icosphereFrequency < 2 || loaded.icosphereFrequency > 1024 || loaded.terrainResolution < 2 ||
                loaded.terrainResolution > 4096 || loaded.textureSize < 4 || loaded.textureSize > 8192)
                return setError("scene manifest procedural sizes are out of range");
        }
        scene = std::move(loaded);
        return true;
    } catch (const std::exception& e) {
        return setError(std::string("could not parse scene manifest: ") + e.what());
    }
}

bool parseCommandLine(int argc, char** argv, RenderConfig& render, SceneConfig& scene) {
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        if (!std::strcmp(a, "--clamp-indirect")) {
            const char* v = next();
            if (!v) {
                std::cerr << "--clamp-indirect requires a finite nonnegative number\n";
                return false;
            }
            char* end = nullptr;
            float limit = std::strtof(v, &end);
            if (end == v || *end != '\0' || !std::isfinite(limit) || limit < 0.0f) {
                std::cerr << "Indirect clamp must be finite and nonnegative\n";
                return false;
            }
            render.indirectRadianceClamp = limit;
            continue;
        }
        if (!std::strcmp(a, "--list-scenes")) {
            std::cout << "Scene presets: sponza, cornell, hurricane, prism, crystal, deer, rtweek, demo, cornell-box, sphere-field, diff-cornell, terrain\n";
            return false;
        }
        if (!std::strcmp(a, "--help") || !std::strcmp(a, "-h")) {
            printHelp();
            return false;
        }
        if (!std::strcmp(a, "--rng") && i + 1 < argc) {
            const char* v = next();
            if (!std::strcmp(v, "xorshift32")) render.rngStrategy = WPT_RNG_XORSHIFT32;
            else if (!std::strcmp(v, "pcg32")) render.rngStrategy = WPT_RNG_PCG32;
            else {
                std::cerr << "RNG strategy must be xorshift32 or pcg32\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--no-caustics")) { render.enableCaustics = false; continue; }
        if (!std::strcmp(a, "--caustics")) { render.enableCaustics = true; continue; }
        if (!std::strcmp(a, "--caustic-photons") && i + 1 < argc) {
            render.causticPhotonCount = static_cast<uint32_t>(std::strtoul(next(), nullptr, 10));
            if (render.causticPhotonCount == 0) {
                std::cerr << "Photon count must be greater than zero\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--caustic-radius") && i + 1 < argc) {
            render.causticGatherRadius = std::strtof(next(), nullptr);
            if (render.causticGatherRadius <= 0.0f) {
                std::cerr << "Gather radius must be greater than zero\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--no-caustic-gather")) { render.enableCausticGather = false; continue; }
        if (!std::strcmp(a, "--persistent-wavefront")) { render.persistentWavefront = true; continue; }
        if (!std::strcmp(a, "--regenerate")) { render.regenerate = true; continue; }
        if (!std::strcmp(a, "--bvh-builder") && i + 1 < argc) {
            const char* v = next();
            if (!std::strcmp(v, "host")) scene.bvhBuilder = WPT_BVH_BUILDER_HOST_SAH;
            else if (!std::strcmp(v, "device")) scene.bvhBuilder = WPT_BVH_BUILDER_DEVICE_PLOC;
            else {
                std::cerr << "Unknown BVH builder: " << v << " (expected host or device)\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--profile-queues")) { render.profileQueues = true; continue; }
        if (!std::strcmp(a, "--resume")) { render.resumeFromCheckpoint = true; continue; }
        if (!std::strcmp(a, "--intersection-debug")) { render.intersectionDebug = true; continue; }
        if (!std::strcmp(a, "--coordinator")) {
            if (render.distributedRole == DistributedRole::Worker) {
                std::cerr << "Cannot combine --coordinator and --worker\n";
                return false;
            }
            render.distributedRole = DistributedRole::Coordinator;
            continue;
        }
        if (!std::strcmp(a, "--worker")) {
            if (render.distributedRole == DistributedRole::Coordinator) {
                std::cerr << "Cannot combine --coordinator and --worker\n";
                return false;
            }
            render.distributedRole = DistributedRole::Worker;
            continue;
        }
        if (!std::strcmp(a, "--listen-host") && i + 1 < argc) { render.listenHost = next(); continue; }
        if (!std::strcmp(a, "--listen-port") && i + 1 < argc) {
            if (!parsePort(next(), render.listenPort)) { std::cerr << "Invalid listen port\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--coordinator-host") && i + 1 < argc) { render.coordinatorHost = next(); continue; }
        if (!std::strcmp(a, "--coordinator-port") && i + 1 < argc) {
            if (!parsePort(next(), render.coordinatorPort)) { std::cerr << "Invalid coordinator port\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--worker-id") && i + 1 < argc) { render.workerId = next(); continue; }
        if (!std::strcmp(a, "--distributed-checkpoint") && i + 1 < argc) {
            render.distributedCheckpointFilename = next();
            continue;
        }
        if (!std::strcmp(a, "--distributed-checkpoint-seconds") && i + 1 < argc) {
            if (!parseUint(next(), render.distributedCheckpointIntervalSeconds, true)) {
                std::cerr << "Distributed checkpoint interval must be greater than zero\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--samples-per-task") && i + 1 < argc) {
            if (!parseUint(next(), render.distributedSamplesPerTask, true)) {
                std::cerr << "Samples per task must be greater than zero\n";
                return false;
            }
            continue;
        }
        if (!std::strcmp(a, "--distributed-lease-seconds") && i + 1 < argc) {
            uint32_t s = 0;
            if (!parseUint(next(), s, true)) {
                std::cerr << "Distributed lease duration must be greater than zero\n";
                return false;
            }
            render.distributedLeaseDurationMs = static_cast<uint64_t>(s) * 1000ull;
            continue;
        }
        if (!std::strcmp(a, "--push-assets")) { render.pushAssets = true; continue; }
        if (!std::strcmp(a, "--asset-cache") && i + 1 < argc) { render.assetCacheDirectory = next(); continue; }
        // Extensions: resolution, sampling and launch shape.
        if (!std::strcmp(a, "--resolution") && i + 1 < argc) {
            if (!parseSize(next(), render.width, render.height)) { std::cerr << "Resolution must be WxH\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--tile") && i + 1 < argc) {
            if (!parseSize(next(), render.tileWidth, render.tileHeight)) { std::cerr << "Tile must be WxH\n"; return false; }
            render.tiledRendering = true;
            continue;
        }
        if (!std::strcmp(a, "--spp") && i + 1 < argc) {
            if (!parseUint(next(), render.samplesPerPixel, true)) { std::cerr << "Samples per pixel must be positive\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--max-depth") && i + 1 < argc) {
            if (!parseUint(next(), render.maxDepth, true)) { std::cerr << "Maximum depth must be positive\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--rr-depth") && i + 1 < argc) {
            if (!parseUint(next(), render.russianRouletteStartDepth, false)) { std::cerr << "Invalid roulette depth\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--block-size") && i + 1 < argc) {
            if (!parseUint(next(), render.blockSize, true)) { std::cerr << "Invalid block size\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--persistent-blocks") && i + 1 < argc) {
            if (!parseUint(next(), render.persistentBlocks, true)) { std::cerr << "Invalid persistent grid size\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--output") && i + 1 < argc) { render.outputFilename = next(); continue; }
        if (!std::strcmp(a, "--seed") && i + 1 < argc) {
            if (!parseUint(next(), scene.seed, false)) { std::cerr << "Invalid seed\n"; return false; }
            continue;
        }
        if (!std::strcmp(a, "--scene") && i + 1 < argc) {
            ScenePreset preset;
            const char* v = next();
            if (!parseScenePreset(v, preset)) {
                std::cerr << "Unknown scene preset: " << v << '\n';
                return false;
            }
            uint32_t seed = scene.seed, builder = scene.bvhBuilder;
            scene = scenePreset(preset);
            scene.seed = seed;
            scene.bvhBuilder = builder;
            continue;
        }
        if (!std::strcmp(a, "--scene-file") && i + 1 < argc) {
            const char* f = next();
            std::string error;
            uint32_t builder = scene.bvhBuilder;
            if (!loadSceneManifest(f, scene, &error)) {
                std::cerr << "Could not load scene manifest '" << f << "': " << error << '\n';
                return false;
            }
            scene.bvhBuilder = builder;
            continue;
        }
        std::cerr << "Unknown option: " << a << "\n";
        printHelp();
        return false;
    }
    if (render.distributedRole != DistributedRole::Local) {
        render.tiledRendering = true;
        if (render.distributedRole == DistributedRole::Worker && render.workerId.empty()) {
            std::cerr << "--worker requires --worker-id\n";
            return false;
        }
    }
    if (render.intersectionDebug || render.shadingNormalDebug) render.samplesPerPixel = 1;
    return true;
}

bool parseCommandLineText(const std::string& command, RenderConfig& render, SceneConfig& scene,
    std::vector<std::string>& storage) {
    storage.clear();
    storage.emplace_back("pathtracer");
    std::string token;
    bool single = false, dbl = false, escaped = false, started = false;
    auto flush = [&] {
        if (!started) return;
        storage.push_back(std::move(token));
        token.clear();
        started = false;
    };
    for (char c : command) {
        if (escaped) {
            token.push_back(c);
            started = true;
            escaped = false;
        } else if (c == '\\' && !single) {
            escaped = true;
            started = true;
        } else if (c == '\'' && !dbl) {
            single = !single;
            started = true;
        } else if (c == '"' && !single) {
            dbl = !dbl;
            started = true;
        } else if (std::isspace(static_cast<unsigned char>(c)) && !single && !dbl) {
            flush();
        } else {
            token.push_back(c);
            started = true;
        }
    }
    if (escaped || single || dbl) {
        std::cerr << "Unterminated escape or quote in coordinator command\n";
        return false;
    }
    flush();
    if (storage.size() == 1) {
        std::cerr << "Coordinator command cannot be empty\n";
        return false;
    }
    std::vector<char*> argv;
    for (std::string& s : storage) argv.push_back(s.data());
    return parseCommandLine(static_cast<int>(argv.size()), argv.data(), render, scene);
}

wpt_render_config toRenderConfig(const RenderConfig& c, ScenePreset preset) {
    wpt_render_config r{};
    r.width = c.width;
    r.height = c.height;
    r.tiled = c.tiledRendering ? 1u : 0u;
    r.tile_width = c.tileWidth;
    r.tile_height = c.tileHeight;
    r.max_depth = c.maxDepth;
    r.rr_start_depth = c.russianRouletteStartDepth;
    r.samples_per_pixel = c.samplesPerPixel;
    r.block_size = c.blockSize;
    r.rng_strategy = c.rngStrategy;
    r.indirect_clamp = c.indirectRadianceClamp;
    r.mode = c.regenerate ? WPT_MODE_REGENERATE : (c.persistentWavefront ? WPT_MODE_PERSISTENT : WPT_MODE_GRAPH);
    r.persistent_blocks = c.persistentBlocks;
    r.intersection_debug = c.intersectionDebug ? 1u : 0u;
    r.shading_normal_debug = c.shadingNormalDebug ? 1u : 0u;
    r.profile_queues = c.profileQueues ? 1u : 0u;
    bool causticPreset = preset == ScenePreset::Cornell || preset == ScenePreset::Hurricane ||
        preset == ScenePreset::Prism || preset == ScenePreset::Crystal || preset == ScenePreset::Terrain;
    if (c.enableCaustics && causticPreset) {
        r.caustic_photons = c.causticPhotonCount;
        r.caustic_max_depth = c.causticMaxDepth;
        r.caustic_gather_radius = c.causticGatherRadius;
        r.caustic_gather = c.enableCausticGather ? 1u : 0u;
        r.caustic_spectral =
            preset == ScenePreset::Prism || preset == ScenePreset::Crystal || preset == ScenePreset::Terrain ? 1u : 0u;
    }
    return r;
}

namespace {
struct Fnv {
    uint64_t v = 1469598103934665603ull;
    void byte(uint8_t b) {
        v ^= b;
        v *= 1099511628211ull;
    }
    void u32(uint32_t x) {
        for (int i = 0; i < 4; ++i) byte(static_cast<uint8_t>((x >> (8 * i)) & 0xffu));
    }
    void u64(uint64_t x) {
        for (int i = 0; i < 8; ++i) byte(static_cast<uint8_t>((x >> (8 * i)) & 0xffu));
    }
    void f32(float f) {
        uint32_t b;
        std::memcpy(&b, &f, 4);
        u32(b);
    }
    void boolean(bool b) { byte(b ? 1 : 0); }
    void str(const std::string& s) {
        u64(s.size());
        for (unsigned char c : s) byte(c);
    }
};
}  // namespace

namespace {
// Versions of the random streams and of the integrator that the reference
// folds into every fingerprint (render_fingerprint.cpp:9,
// render_fingerprint.hpp:6).  Images of this renderer match the
// reference's integrator at that version.
constexpr uint32_t kRngImplementationVersion = 1;
constexpr uint32_t kIntegratorVersion = 12;

// A file's contents: byte 0 if it cannot be opened, else byte 1 and every
// byte of the file.
void appendAsset(Fnv& f, const std::string& filename) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) {
        f.byte(0);
        return;
    }
    f.byte(1);
    std::vector<char> buffer(64 * 1024);
    while (in.read(buffer.data(), static_cast<std::streamsize>(buffer.size())) || in.gcount() > 0)
        for (std::streamsize i = 0; i < in.gcount(); ++i) f.byte(static_cast<uint8_t>(buffer[static_cast<size_t>(i)]));
}

// render_fingerprint.cpp:82-120.  The procedural presets that the
// reference lacks add their generator seed and icosphere frequency last.
void appendScene(Fnv& f, const SceneConfig& scene, bool includeName) {
    f.u32(static_cast<uint32_t>(scene.preset));
    if (includeName) f.str(scene.name);
    switch (scene.preset) {
    case ScenePreset::Sponza: appendAsset(f, scene.gltfFilename); break;
    case ScenePreset::Cornell:
    case ScenePreset::Hurricane:
    case ScenePreset::Crystal:
    case ScenePreset::Deer:
    case ScenePreset::RtWeek: appendAsset(f, scene.objectFilename); break;
    default: break;
    }
    f.f32(scene.gltfScale);
    f.f32(scene.objectScale);
    f.f32(scene.objectTranslation.x);
    f.f32(scene.objectTranslation.y);
    f.f32(scene.objectTranslation.z);
    f.u32(static_cast<uint32_t>(scene.objectSource));
    f.u32(static_cast<uint32_t>(scene.lightProfile));
    f.u32(static_cast<uint32_t>(scene.acceleration));
    f.boolean(scene.neutralRoom);
    f.boolean(scene.removeBackdrop);
    f.boolean(scene.convertObjectMaterialsToDielectric);
    f.u32(static_cast<uint32_t>(scene.objectMaterialOverride));
    f.boolean(scene.addSponzaTopLight);
    f.boolean(scene.ignoreSponzaLightOcclusion);
    f.boolean(scene.useNormalMaps);
    f.f32(scene.normalMapMinimumCosine);
    f.boolean(!scene.environmentFilename.empty());
    if (!scene.environmentFilename.empty()) {
        appendAsset(f, scene.environmentFilename);
        f.f32(scene.environmentIntensity);
        f.f32(scene.environmentRotationDegrees);
    }
    if (scene.preset == ScenePreset::CornellBox || scene.preset == ScenePreset::SphereField ||
        scene.preset == ScenePreset::DiffCornell || scene.preset == ScenePreset::Terrain) {
        f.u32(scene.seed);
        f.u32(scene.icosphereFrequency);
    }
    if (scene.preset == ScenePreset::Terrain) {
        f.u32(scene.terrainResolution);
        f.u32(scene.textureSize);
    }
}

// render_fingerprint.cpp:122-145.
void appendRender(Fnv& f, const RenderConfig& render) {
    f.u32(render.width);
    f.u32(render.height);
    f.u32(render.tileWidth);
    f.u32(render.tileHeight);
    f.u32(render.maxDepth);
    f.u32(render.russianRouletteStartDepth);
    f.u32(render.samplesPerPixel);
    f.u32(render.blockSize);
    f.boolean(render.tiledRendering);
    f.boolean(render.intersectionDebug);
    f.boolean(render.shadingNormalDebug);
    f.boolean(render.persistentWavefront);
    f.boolean(render.enableCaustics);
    f.boolean(render.enableCausticGather);
    f.u32(render.causticPhotonCount);
    f.u32(render.causticMaxDepth);
    f.f32(render.causticGatherRadius);
    f.f32(render.indirectRadianceClamp);
    f.u32(render.rngStrategy);
    f.u32(kRngImplementationVersion);
    f.u32(kIntegratorVersion);
}
}  // namespace

uint64_t computeRenderFingerprint(const RenderConfig& render, const SceneConfig& scene) {
    Fnv f;
    appendScene(f, scene, true);
    appendRender(f, render);
    return f.v;
}

uint64_t computeSceneFingerprint(const SceneConfig& scene) {
    Fnv f;
    appendScene(f, scene, false);
    return f.v;
}

}  // namespace wpt
