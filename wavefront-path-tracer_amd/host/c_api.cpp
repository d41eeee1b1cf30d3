// C ABI of the host layer (include/wpt_host.h).
#include <cstring>
#include <optional>
#include <iostream>
#include <string>

#include "../../include/wpt_host.h"
#include "../csrc/api/last_error.h"
#include "checkpoint.hpp"
#include "config.hpp"
#include "ledger.hpp"
#include "image_decode.hpp"
#include "loaders.hpp"
#include "render_driver.hpp"
#include "scene_factory.hpp"

struct wpt_host_scene {
    wpt::SceneConfig config;
    wpt::HostScene scene;
};

struct wpt_ledger {
    wpt::LeaseLedger ledger;
};

struct wpt_ledger_writer {
    wpt::LedgerCheckpointWriter writer;
    explicit wpt_ledger_writer(const char* filename) : writer(filename) {}
};

namespace {
wpt::RenderTaskId taskId(const wpt_task& t) { return wpt::RenderTaskId{t.tile_ordinal, t.sample_start, t.sample_count}; }

int invalid(const std::string& message) { return wpt::setLastError(WPT_ERROR_INVALID_ARGUMENT, message); }
}  // namespace

extern "C" {

WPT_API int wpt_host_scene_create(const char* name, uint32_t seed, uint32_t frequency, wpt_host_scene** out) {
    if (!name || !out) return WPT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    wpt::SceneConfig config;
    std::string n(name);
    if (n.size() > 5 && n.compare(n.size() - 5, 5, ".json") == 0) {
        std::string error;
        if (!wpt::loadSceneManifest(name, config, &error)) return invalid(error);
    } else {
        wpt::ScenePreset preset;
        if (!wpt::parseScenePreset(name, preset)) return invalid(std::string("unknown scene preset: ") + name);
        config = wpt::scenePreset(preset);
    }
    if (seed) config.seed = seed;
    if (frequency) config.icosphereFrequency = frequency;
    wpt_host_scene* s = new wpt_host_scene();
    s->config = config;
    std::string error;
    if (!wpt::createSceneFromConfig(config, s->scene, &error)) {
        delete s;
        return wpt::setLastError(WPT_ERROR_UNSUPPORTED, error);
    }
    *out = s;
    return WPT_OK;
}

WPT_API int wpt_host_scene_fingerprint(const char* name, uint64_t* out) {
    if (!name || !out) return WPT_ERROR_INVALID_ARGUMENT;
    wpt::SceneConfig config;
    std::string n(name);
    if (n.size() > 5 && n.compare(n.size() - 5, 5, ".json") == 0) {
        std::string error;
        if (!wpt::loadSceneManifest(name, config, &error)) return invalid(error);
    } else {
        wpt::ScenePreset preset;
        if (!wpt::parseScenePreset(name, preset)) return WPT_ERROR_INVALID_ARGUMENT;
        config = wpt::scenePreset(preset);
    }
    *out = wpt::computeSceneFingerprint(config);
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
    if (render) *render = wpt::toRenderConfig(rc, sc.preset);
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

WPT_API int wpt_host_load_hdr(const char* filename, float* rgb, uint64_t capacity, uint32_t* width,
    uint32_t* height) {
    if (!filename || !width || !height) return invalid("null argument");
    std::vector<float> pixels;
    std::string error;
    if (!wpt::loadRadianceHdr(filename, pixels, *width, *height, &error)) return invalid(error);
    if (rgb) {
        if (capacity < pixels.size()) return invalid("HDR output buffer too small");
        std::memcpy(rgb, pixels.data(), sizeof(float) * pixels.size());
    }
    return WPT_OK;
}

WPT_API int wpt_host_decode_image(const uint8_t* data, uint64_t size, uint8_t* rgba, uint64_t capacity,
    uint32_t* width, uint32_t* height) {
    if (!data || !width || !height) return invalid("null argument");
    std::vector<uint8_t> pixels;
    std::string error;
    if (!wpt::decodeImage(data, static_cast<size_t>(size), pixels, *width, *height, &error)) return invalid(error);
    if (rgba) {
        if (capacity < pixels.size()) return invalid("image output buffer too small");
        std::memcpy(rgba, pixels.data(), pixels.size());
    }
    return WPT_OK;
}

WPT_API int wpt_host_render_checkpoint_write(const char* filename, const float* rgb, uint32_t width, uint32_t height,
    uint32_t sampleIndex, uint64_t fingerprint) {
    if (!filename || !rgb) return invalid("null argument");
    std::string error;
    if (!wpt::writeRenderCheckpoint(filename, rgb, width, height, sampleIndex, fingerprint, &error))
        return invalid(error);
    return WPT_OK;
}

WPT_API int wpt_host_render_checkpoint_read(const char* filename, uint32_t width, uint32_t height,
    uint64_t fingerprint, float* rgb, uint32_t* sampleIndex) {
    if (!filename || !rgb || !sampleIndex) return invalid("null argument");
    std::vector<float> pixels;
    std::string error;
    if (!wpt::readRenderCheckpoint(filename, width, height, fingerprint, pixels, *sampleIndex, &error))
        return invalid(error);
    std::memcpy(rgb, pixels.data(), sizeof(float) * pixels.size());
    return WPT_OK;
}

WPT_API int wpt_ledger_create(const wpt_ledger_job* job, const wpt_task* tasks, uint32_t count, wpt_ledger** out) {
    if (!job || !out || (count && !tasks)) return invalid("null argument");
    *out = nullptr;
    std::vector<wpt::RenderTask> list;
    list.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
        const wpt_task& t = tasks[i];
        if (t.x >= job->width || t.y >= job->height || t.width > job->width - t.x || t.height > job->height - t.y)
            return invalid("ledger task tile lies outside the job's image");
        list.push_back(wpt::RenderTask{t.job_fingerprint, taskId(t), wpt::DistributedTile{t.x, t.y, t.width, t.height}});
    }
    wpt::LedgerJob j{job->job_fingerprint, job->width, job->height, job->tile_width, job->tile_height,
        job->samples_per_pixel, job->samples_per_task};
    *out = new wpt_ledger{wpt::LeaseLedger(j, list)};
    return WPT_OK;
}

WPT_API int wpt_ledger_destroy(wpt_ledger* l) {
    delete l;
    return WPT_OK;
}

WPT_API int wpt_ledger_lease_next(wpt_ledger* l, const char* worker, uint64_t nowMs, uint64_t leaseMs, wpt_task* task,
    uint64_t* expiresAtMs) {
    if (!l || !worker || !task) return -invalid("null argument");
    std::optional<wpt::LedgerLease> lease = l->ledger.leaseNext(worker, nowMs, leaseMs);
    if (!lease) return 0;
    const wpt::RenderTask& t = lease->task;
    *task = wpt_task{t.jobFingerprint, t.id.tileOrdinal, t.id.sampleStart, t.id.sampleCount, t.tile.x, t.tile.y,
        t.tile.width, t.tile.height};
    if (expiresAtMs) *expiresAtMs = lease->expiresAtMs;
    return 1;
}

WPT_API int wpt_ledger_commit(wpt_ledger* l, const wpt_task* task, uint32_t sampleCount, const float* radiance,
    uint64_t floats) {
    if (!l || !task || (floats && !radiance)) return invalid("null argument");
    wpt::LedgerResult r{task->job_fingerprint, taskId(*task), wpt::DistributedTile{task->x, task->y, task->width,
        task->height}, sampleCount, std::vector<float>(radiance, radiance + floats)};
    switch (l->ledger.commit(r)) {
    case wpt::LedgerCommit::Accepted: return WPT_LEDGER_ACCEPTED;
    case wpt::LedgerCommit::Duplicate: return WPT_LEDGER_DUPLICATE;
    default: return WPT_LEDGER_REJECTED;
    }
}

WPT_API uint32_t wpt_ledger_renew_leases(wpt_ledger* l, const char* worker, uint64_t nowMs, uint64_t leaseMs) {
    return l && worker ? l->ledger.renewWorkerLeases(worker, nowMs, leaseMs) : 0u;
}

WPT_API uint32_t wpt_ledger_release_leases(wpt_ledger* l, const char* worker) {
    return l && worker ? l->ledger.releaseWorkerLeases(worker) : 0u;
}

WPT_API uint32_t wpt_ledger_requeue_expired(wpt_ledger* l, uint64_t nowMs) {
    return l ? l->ledger.requeueExpired(nowMs) : 0u;
}

WPT_API int wpt_ledger_task_state(const wpt_ledger* l, const wpt_task* task) {
    if (!l || !task) return -invalid("null argument");
    switch (l->ledger.taskState(taskId(*task))) {
    case wpt::LedgerTaskState::Leased: return WPT_TASK_LEASED;
    case wpt::LedgerTaskState::Completed: return WPT_TASK_COMPLETED;
    default: return WPT_TASK_PENDING;
    }
}

WPT_API int wpt_ledger_counts(const wpt_ledger* l, uint64_t* tasks, uint64_t* pending, uint64_t* completed) {
    if (!l) return invalid("null ledger");
    if (tasks) *tasks = l->ledger.taskCount();
    if (pending) *pending = l->ledger.pendingTaskCount();
    if (completed) *completed = l->ledger.completedTaskCount();
    return WPT_OK;
}

WPT_API int wpt_ledger_read(const wpt_ledger* l, float* radiance, uint32_t* sampleCounts) {
    if (!l) return invalid("null ledger");
    wpt::LedgerCheckpoint c = l->ledger.snapshot(0);
    if (radiance) std::memcpy(radiance, c.radiance.data(), sizeof(float) * c.radiance.size());
    if (sampleCounts) std::memcpy(sampleCounts, c.sampleCounts.data(), sizeof(uint32_t) * c.sampleCounts.size());
    return WPT_OK;
}

WPT_API int wpt_ledger_write_checkpoint(const wpt_ledger* l, const char* filename, uint64_t sequence) {
    if (!l || !filename) return invalid("null argument");
    std::string error;
    if (!wpt::writeLedgerCheckpoint(filename, l->ledger.snapshot(sequence), &error)) return invalid(error);
    return WPT_OK;
}

WPT_API int wpt_ledger_restore_checkpoint(wpt_ledger* l, const char* filename, uint64_t* sequence) {
    if (!l || !filename) return invalid("null argument");
    wpt::LedgerCheckpoint c;
    std::string error;
    if (!wpt::readLedgerCheckpoint(filename, c, &error) || !l->ledger.restore(c, &error)) return invalid(error);
    if (sequence) *sequence = c.sequence;
    return WPT_OK;
}

WPT_API int wpt_ledger_writer_create(const char* filename, wpt_ledger_writer** out) {
    if (!filename || !out) return invalid("null argument");
    *out = new wpt_ledger_writer(filename);
    return WPT_OK;
}

WPT_API int wpt_ledger_writer_enqueue(wpt_ledger_writer* w, const wpt_ledger* l, uint64_t sequence) {
    if (!w || !l) return invalid("null argument");
    std::string error;
    if (!w->writer.enqueue(l->ledger.snapshot(sequence), &error)) return invalid(error);
    return WPT_OK;
}

WPT_API int wpt_ledger_writer_flush(wpt_ledger_writer* w) {
    if (!w) return invalid("null writer");
    std::string error;
    if (!w->writer.flush(&error)) return invalid(error);
    return WPT_OK;
}

WPT_API int wpt_ledger_writer_destroy(wpt_ledger_writer* w) {
    if (!w) return WPT_OK;
    w->writer.shutdown();
    std::string error;
    bool failed = w->writer.failed(&error);
    delete w;
    return failed ? invalid(error) : WPT_OK;
}

WPT_API int wpt_host_main(int argc, char** argv) {
    wpt::RenderConfig config;
    wpt::SceneConfig sceneConfig = wpt::scenePreset(wpt::ScenePreset::Sponza);
    if (!wpt::parseCommandLine(argc, argv, config, sceneConfig)) return 0;
    if (config.distributedRole == wpt::DistributedRole::Worker) {
        std::cerr << "--worker needs a networked coordinator, which this renderer does not provide; "
                     "--coordinator renders the task table through a local task ledger\n";
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
    desc.bvh_builder = sceneConfig.bvhBuilder;
    wpt_scene* scene = nullptr;
    if (wpt_scene_create(&desc, &scene) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        return 1;
    }
    std::cout << "Scene preset: " << sceneConfig.name << '\n';
    int result;
    {
        wpt::RenderDriver driver(scene, camera, config, sceneConfig);
        result = config.distributedRole == wpt::DistributedRole::Coordinator
            ? wpt::runLocalLeases(driver, config, sceneConfig)
            : driver.run();
    }
    wpt_scene_destroy(scene);
    return result;
}

}  // extern "C"
