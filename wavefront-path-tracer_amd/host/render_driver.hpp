// Render orchestration over the C ABI.
//
// Mirrors reference RenderDriver (include/renderer/render_driver.cuh:52-88,
// src/renderer/render_driver.cu:105-482), RenderTaskRenderer
// (include/renderer/render_task.cuh:11-29) and the task model
// (include/distributed/task.hpp:6-29).  Device work happens behind libwpt's
// C entry points; this layer owns configuration, tiling, output files and
// the task seam.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/wpt.h"
#include "config.hpp"

namespace wpt {

struct DistributedTile {
    uint32_t x, y, width, height;
};

struct RenderTaskId {
    uint32_t tileOrdinal, sampleStart, sampleCount;
    bool operator==(const RenderTaskId& o) const {
        return tileOrdinal == o.tileOrdinal && sampleStart == o.sampleStart && sampleCount == o.sampleCount;
    }
};

struct RenderTask {
    uint64_t jobFingerprint;
    RenderTaskId id;
    DistributedTile tile;
};

struct RenderTaskResult {
    bool success = false;
    RenderTaskId taskId{};
    DistributedTile tile{};
    uint32_t sampleCount = 0;
    std::vector<float> radiance;  // width*height*3 unnormalized sums
};

struct RenderTile {
    uint32_t x, y, width, height, advanceSample;
};

// Row-major tiles clamped at the image edge (src/renderer/tile_scheduler.cpp:5-21).
std::vector<RenderTile> createRenderTiles(uint32_t width, uint32_t height, uint32_t tileWidth, uint32_t tileHeight);

// Deterministic task table, tile-major then sample ranges (src/distributed/task.cpp:42-75).
std::vector<RenderTask> createRenderTasks(uint32_t width, uint32_t height, uint32_t tileWidth, uint32_t tileHeight,
    uint32_t samplesPerPixel, uint32_t samplesPerTask, uint64_t jobFingerprint);

// FNV-1a task key for duplicate rejection and completion bitmaps
// (src/distributed/task.cpp:33-40).
uint64_t computeRenderTaskKey(uint64_t jobFingerprint, const RenderTaskId& id);

std::string sha256Hex(const void* data, size_t bytes);
// Gamma-2 8-bit PPM (src/renderer/image_output.cpp:31-53).
bool writePpm(const std::string& filename, const std::vector<float>& rgb, uint32_t width, uint32_t height);

class RenderDriver {
public:
    RenderDriver(wpt_scene* scene, const wpt_camera& camera, const RenderConfig& config, const SceneConfig& sceneConfig);
    ~RenderDriver();
    RenderDriver(const RenderDriver&) = delete;
    RenderDriver& operator=(const RenderDriver&) = delete;

    int run();
    bool prepare();
    wpt_renderer* renderer() const { return renderer_; }
    uint64_t fingerprint() const { return fingerprint_; }
    const RenderConfig& config() const { return config_; }

private:
    wpt_scene* scene_;
    wpt_camera camera_;
    RenderConfig config_;
    SceneConfig sceneConfig_;
    uint64_t fingerprint_;
    wpt_renderer* renderer_ = nullptr;
};

class RenderTaskRenderer {
public:
    explicit RenderTaskRenderer(RenderDriver& driver) : driver_(driver) {}
    RenderTaskResult render(const RenderTask& task);

private:
    RenderDriver& driver_;
};

}  // namespace wpt
