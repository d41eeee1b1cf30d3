#include "render_driver.hpp"

#include "checkpoint.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

namespace wpt {

std::vector<RenderTile> createRenderTiles(uint32_t width, uint32_t height, uint32_t tw, uint32_t th) {
    std::vector<RenderTile> tiles;
    for (uint32_t y = 0; y < height; y += th)
        for (uint32_t x = 0; x < width; x += tw)
            tiles.push_back(RenderTile{x, y, std::min(tw, width - x), std::min(th, height - y), 0});
    return tiles;
}

std::vector<RenderTask> createRenderTasks(uint32_t width, uint32_t height, uint32_t tw, uint32_t th, uint32_t spp,
    uint32_t perTask, uint64_t fingerprint) {
    std::vector<RenderTask> tasks;
    if (width == 0 || height == 0 || tw == 0 || th == 0 || spp == 0 || perTask == 0) return tasks;
    uint32_t ordinal = 0;
    for (const RenderTile& tile : createRenderTiles(width, height, tw, th)) {
        for (uint32_t start = 0; start < spp; start += perTask) {
            uint32_t count = std::min(perTask, spp - start);
            tasks.push_back(RenderTask{fingerprint, RenderTaskId{ordinal, start, count},
                DistributedTile{tile.x, tile.y, tile.width, tile.height}});
        }
        ++ordinal;
    }
    return tasks;
}

uint64_t computeRenderTaskKey(uint64_t fingerprint, const RenderTaskId& id) {
    uint64_t v = 1469598103934665603ull;
    auto bytes = [&](uint64_t x, int n) {
        for (int i = 0; i < n; ++i) {
            v ^= static_cast<uint8_t>((x >> (8 * i)) & 0xffu);
            v *= 1099511628211ull;
        }
    };
    bytes(fingerprint, 8);
    bytes(id.tileOrdinal, 4);
    bytes(id.sampleStart, 4);
    bytes(id.sampleCount, 4);
    return v;
}

// ---- SHA-256 (FIPS 180-4)
namespace {
struct Sha256 {
    uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    uint8_t block[64];
    size_t used = 0;
    uint64_t total = 0;
    static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
    void compress(const uint8_t* p) {
        static const uint32_t k[64] = {0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u,
            0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
            0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau,
            0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u,
            0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu,
            0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u,
            0xf40e3585u, 0x106aa070u, 0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au,
            0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu,
            0xbef9a3f7u, 0xc67178f2u};
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t(p[4 * i]) << 24) | (uint32_t(p[4 * i + 1]) << 16) | (uint32_t(p[4 * i + 2]) << 8) | p[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
        for (int i = 0; i < 64; ++i) {
            uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
            uint32_t ch = (e & f) ^ (~e & g);
            uint32_t t1 = hh + S1 + ch + k[i] + w[i];
            uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
            uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
            uint32_t t2 = S0 + maj;
            hh = g;
            g = f;
            f = e;
            e = d + t1;
            d = c;
            c = b;
            b = a;
            a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
    }
    void update(const uint8_t* p, size_t n) {
        total += n;
        while (n) {
            size_t take = std::min(n, 64 - used);
            std::memcpy(block + used, p, take);
            used += take;
            p += take;
            n -= take;
            if (used == 64) {
                compress(block);
                used = 0;
            }
        }
    }
    std::string hex() {
        uint64_t bits = total * 8;
        uint8_t pad = 0x80;
        update(&pad, 1);
        uint8_t zero = 0;
        while (used != 56) update(&zero, 1);
        uint8_t len[8];
        for (int i = 0; i < 8; ++i) len[i] = static_cast<uint8_t>(bits >> (56 - 8 * i));
        update(len, 8);
        char out[65];
        for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
        return std::string(out, 64);
    }
};
}  // namespace

std::string sha256Hex(const void* data, size_t bytes) {
    Sha256 s;
    s.update(static_cast<const uint8_t*>(data), bytes);
    return s.hex();
}

bool writePpm(const std::string& filename, const std::vector<float>& rgb, uint32_t width, uint32_t height) {
    std::ofstream out(filename);
    if (!out) {
        std::cerr << "Could not open " << filename << " for writing\n";
        return false;
    }
    out << "P3\n" << width << ' ' << height << "\n255\n";
    auto byte = [](float v) {
        v = std::min(std::max(v, 0.0f), 1.0f);
        v = std::sqrt(v);
        return static_cast<uint32_t>(static_cast<uint8_t>(v * 255.999f));
    };
    for (size_t i = 0; i + 2 < rgb.size(); i += 3) out << byte(rgb[i]) << ' ' << byte(rgb[i + 1]) << ' ' << byte(rgb[i + 2]) << '\n';
    return true;
}

RenderDriver::RenderDriver(wpt_scene* scene, const wpt_camera& camera, const RenderConfig& config,
    const SceneConfig& sceneConfig)
    : scene_(scene), camera_(camera), config_(config), sceneConfig_(sceneConfig),
      fingerprint_(computeRenderFingerprint(config, sceneConfig)) {}

RenderDriver::~RenderDriver() {
    if (renderer_) wpt_renderer_destroy(renderer_);
}

bool RenderDriver::prepare() {
    if (renderer_) return true;
    wpt_render_config rc = toRenderConfig(config_, sceneConfig_.preset);
    if (wpt_renderer_create(scene_, &camera_, &rc, fingerprint_, &renderer_) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        renderer_ = nullptr;
        return false;
    }
    return true;
}

int RenderDriver::run() {
    if (!prepare()) return 1;
    size_t tiles = config_.tiledRendering ? createRenderTiles(config_.width, config_.height, config_.tileWidth,
        config_.tileHeight).size() : 1;
    std::cout << "Render mode: " << (config_.regenerate ? "persistent path regeneration" :
        (config_.persistentWavefront ? "persistent device wavefront" :
        (config_.tiledRendering ? "tiled" : "full frame"))) << ", " << tiles << " work tiles\n";
    const uint32_t spp = config_.samplesPerPixel;
    uint32_t completed = 0;
    if (config_.resumeFromCheckpoint) {
        std::vector<float> saved;
        std::string error;
        if (!readRenderCheckpoint(config_.checkpointFilename, config_.width, config_.height, fingerprint_, saved,
                completed, &error)) {
            std::cerr << error << '\n';
            return 1;
        }
        if (completed > spp) {
            std::cerr << "Checkpoint has more samples than the current render target\n";
            return 1;
        }
        if (wpt_renderer_load_framebuffer(renderer_, saved.data(), completed) != WPT_OK) {
            std::cerr << wpt_last_error() << '\n';
            return 1;
        }
        std::cout << "Resuming from " << completed << " completed samples\n";
    } else if (wpt_renderer_reset(renderer_) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        return 1;
    }
    wpt_stats stats{};
    wpt_renderer_last_stats(renderer_, &stats);
    if (stats.caustic_build_milliseconds > 0.0)
        std::cout << "Traced " << config_.causticPhotonCount << " photons; " << stats.caustic_photons_stored
                  << " stored in the caustic grid; photon map build: " << stats.caustic_build_milliseconds << " ms\n";
    // Samples run in steps of the checkpoint interval; a checkpoint is
    // queued after each step short of the target (render_driver.cu:359-383).
    const uint32_t step = std::max(1u, (spp * config_.checkpointProgressPercent + 99) / 100);
    RenderCheckpointWriter checkpoints(config_.width, config_.height, config_.checkpointFilename,
        config_.checkpointBufferCount, fingerprint_);
    double traceMs = 0.0;
    uint64_t queries = 0, samplesTraced = 0;
    while (completed < spp) {
        uint32_t count = std::min(step - completed % step, spp - completed);
        if (wpt_renderer_render_samples(renderer_, completed, count) != WPT_OK) {
            std::cerr << wpt_last_error() << '\n';
            return 1;
        }
        wpt_renderer_last_stats(renderer_, &stats);
        traceMs += stats.trace_milliseconds;
        queries += stats.closest_queries + stats.shadow_queries;
        samplesTraced += stats.samples;
        completed += count;
        std::cout << "Completed sample " << completed << " of " << spp << '\n';
        if (completed % step == 0 && completed < spp) {
            if (checkpoints.enqueue(renderer_, completed))
                std::cout << "Queued checkpoint at " << completed << " of " << spp << " samples\n";
            else
                std::cout << "Skipped checkpoint at " << completed << " of " << spp
                          << " samples because all checkpoint buffers are busy\n";
        }
    }
    std::string checkpointError;
    if (!checkpoints.finish(&checkpointError)) {
        std::cerr << checkpointError << '\n';
        return 1;
    }
    std::cout << "Trace time (" << (config_.tiledRendering ? "tiled" : "full frame") << "): " << traceMs << " ms\n";
    double seconds = traceMs / 1000.0;
    if (seconds > 0)
        std::cout << "Throughput: " << queries / seconds / 1e6 << " Mrays/s, " << samplesTraced / seconds / 1e6
                  << " Msamples/s\n";
    if (config_.profileQueues) {
        std::cout << "Queue occupancy (last launch):";
        for (uint32_t b = 0; b < stats.queue_count_depth; ++b) std::cout << ' ' << stats.queue_counts[b];
        std::cout << '\n';
    }
    std::vector<float> pixels(3ull * config_.width * config_.height);
    uint32_t samples = 0;
    if (wpt_renderer_read_framebuffer(renderer_, pixels.data(), &samples) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        return 1;
    }
    if (samples == 0) {
        std::cerr << "Render completed without accumulating any samples\n";
        return 1;
    }
    std::cout << "Raw framebuffer SHA-256: " << sha256Hex(pixels.data(), pixels.size() * sizeof(float)) << '\n';
    for (float& v : pixels) v = v / static_cast<float>(samples);
    if (!writePpm(config_.outputFilename, pixels, config_.width, config_.height)) return 1;
    std::cout << "Wrote " << config_.outputFilename << '\n';
    return 0;
}

RenderTaskResult RenderTaskRenderer::render(const RenderTask& task) {
    RenderTaskResult result;
    result.taskId = task.id;
    result.tile = task.tile;
    result.sampleCount = task.id.sampleCount;
    if (!driver_.prepare()) return result;
    wpt_task t{};
    t.job_fingerprint = task.jobFingerprint;
    t.tile_ordinal = task.id.tileOrdinal;
    t.sample_start = task.id.sampleStart;
    t.sample_count = task.id.sampleCount;
    t.x = task.tile.x;
    t.y = task.tile.y;
    t.width = task.tile.width;
    t.height = task.tile.height;
    result.radiance.assign(3ull * task.tile.width * task.tile.height, 0.0f);
    if (wpt_render_task(driver_.renderer(), &t, result.radiance.data()) != WPT_OK) {
        std::cerr << wpt_last_error() << '\n';
        result.radiance.clear();
        return result;
    }
    result.success = true;
    return result;
}

}  // namespace wpt
