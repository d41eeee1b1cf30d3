// C ABI of the host layer (include/wpt_host.h).
#include <cstring>
#include <iostream>
#include <string>

#include "../../include/wpt_host.h"
#include "config.hpp"
#include "render_driver.hpp"
#include "scene_factory.hpp"

struct wpt_host_scene {
    wpt::SceneConfig config;
    wpt::HostScene scene;
};

extern "C" {

WPT_API int wpt_host_scene_create(const char* name, uint32_t seed, uint32_t frequency, wpt_host_scene** out) {
    if (!name || !out) return WPT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    wpt::SceneConfig config;
    std::string n(name);
    if (n.size() > 5 && n.compare(n.size() - 5, 5, ".json") == 0) {
        std::string error;
        if (!wpt::loadSceneManifest(name, config, &error)) {
            std::cerr << error << '\n';
            return WPT_ERROR_INVALID_ARGUMENT;
        }
    } else {
        wpt::ScenePreset preset;
        if (!wpt::parseScenePreset(name, preset)) return WPT_ERROR_INVALID_ARGUMENT;
        config = wpt::scenePreset(preset);
    }
    if (seed) config.seed = seed;
    if (frequency) config.icosphereFrequency = frequency;
    wpt_host_scene* s = new wpt_host_scene();
    s->config = config;
    std::string error;
    if (!wpt::createSceneFromConfig(config, s->scene, &error)) {
        std::cerr << error << '\n';
        delete s;
        return WPT_ERROR_UNSUPPORTED;
    }
    *out = s;
    return WPT_OK;
}

WPT_API int wpt_host_scene_get_desc(const wpt_host_scene* s, wpt_scene_desc* out) {
    if (!s || !out) return WPT_ERROR_INVALID_ARGUMENT;
    *out = s->scene.desc();
    return WPT_OK;
}

WPT_API int wpt_host_scene_get_camera(const wpt_host_scene* s, uint32_t w, uint32_t h, wpt_camera* out) {
    if (!s || !out || w == 0 || h == 0) return WPT_ERROR_INVALID_ARGUMENT;
    *out = wpt::createSceneCamera(s->config, w, h);
    return WPT_OK;
}

WPT_API int wpt_host_scene_destroy(wpt_host_scene* s) {
    delete s;
    return WPT_OK;
}

WPT_API int wpt_host_parse_command_line(int argc, const char* const* argv, wpt_render_config* render, char* sceneName,
    uint32_t capacity, uint64_t* fingerprint) {
    wpt::RenderConfig rc;
    wpt::SceneConfig sc = wpt::scenePreset(wpt::ScenePreset::Sponza);
    if (!wpt::parseCommandLine(argc, const_cast<char**>(argv), rc, sc)) return WPT_ERROR_INVALID_ARGUMENT;
    if (render) *render = wpt::toRenderConfig(rc);
    if (sceneName && capacity) {
        std::strncpy(sceneName, sc.name.c_str(), capacity - 1);
        sceneName[capacity - 1] = '\0';
    }
    if (fingerprint) *fingerprint = wpt::computeRenderFingerprint(rc, sc);
    return WPT_OK;
}

WPT_API uint32_t wpt_host_create_tasks(uint32_t width, uint32_t height, uint32_t tw, uint32_t th, uint32_t spp,
    uint32_t perTask, uint64_t fingerprint, wpt_task* out, uint32_t capacity) {
    std::vector<wpt::RenderTask> tasks = wpt::createRenderTasks(width, height, tw, th, spp, perTask, fingerprint);
    for (uint32_t i = 0; out && i < tasks.size() && i < capacity; ++i) {
        const wpt::RenderTask& t = tasks[i];
        out[i] = wpt_task{t.jobFingerprint, t.id.tileOrdinal, t.id.sampleStart, t.id.sampleCount, t.tile.x, t.tile.y,
            t.tile.width, t.tile.height};
    }
    return static_cast<uint32_t>(tasks.size());
}

WPT_API uint64_t wpt_host_task_key(uint64_t fingerprint, const wpt_task* task) {
    if (!task) return 0;
    return wpt::computeRenderTaskKey(fingerprint, wpt::RenderTaskId{task->tile_ordinal, task->sample_start, task->sample_count});
}

WPT_API void wpt_host_sha256(const void* data, size_t bytes, char out[65]) {
    std::string h = wpt::sha256Hex(data, bytes);
    std::memcpy(out, h.c_str(), 65);
}

WPT_API int wpt_host_main(int argc, char** argv) {
    wpt::RenderConfig config;
    wpt::SceneConfig sceneConfig = wpt::scenePreset(wpt::ScenePreset::Sponza);
    if (!wpt::parseCommandLine(argc, argv, config, sceneConfig)) return 0;
    if (config.distributedRole != wpt::DistributedRole::Local) {
        std::cerr << "Distributed roles run as local task leases on one GPU; networked coordination is not available\n";
        return 1;
    }
    std::cout << "Starting wavefront path tracer\n";
    wpt::HostScene host;
    std::string error;
    if (!wpt::createSceneFromConfig(sceneConfig, host, &error)) {
        std::cerr << error << '\n';
        return 1;
    }
    wpt_camera camera = wpt::createSceneCamera(sceneConfig, config.width, config.height);
    wpt_scene_desc desc = host.desc();
    wpt_scene* scene = nullptr;
    if (wpt_scene_create(&desc, &scene) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        return 1;
    }
    std::cout << "Scene preset: " << sceneConfig.name << '\n';
    int result;
    {
        wpt::RenderDriver driver(scene, camera, config, sceneConfig);
        result = driver.run();
    }
    wpt_scene_destroy(scene);
    return result;
}

}  // extern "C"
