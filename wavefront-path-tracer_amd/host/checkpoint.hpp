// Checkpoint files of the render driver and of the task ledger.
//
// Render checkpoints (src/renderer/render_checkpoint.cpp:17-195): a 32-byte
// header {magic "PTCHKPT1", version 3, width, height, completed samples,
// fingerprint} followed by the framebuffer's unnormalized RGB sums as
// width*height*3 floats.  Files are written to "<name>.tmp" and renamed, so
// a reader never sees a partial file.
//
// Ledger checkpoints (src/distributed/checkpoint.cpp): an 88-byte header
// {magic "PTDIST1\0", version 1, header size, job, sequence, pixel count,
// task count, payload size, FNV-1a payload checksum} followed by the
// accumulated radiance (3 floats per pixel), the per-pixel sample counts and
// one completion byte per task.
//
// The formats are the reference's byte for byte, so a checkpoint written by
// either renderer resumes in the other when the fingerprints agree.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wpt.h"

namespace wpt {

bool writeRenderCheckpoint(const std::string& filename, const float* rgb, uint32_t width, uint32_t height,
    uint32_t sampleIndex, uint64_t fingerprint, std::string* error);
// Reads a render checkpoint that must match the resolution and fingerprint
// (loadRenderSessionCheckpoint, render_checkpoint.cpp:163-195).
bool readRenderCheckpoint(const std::string& filename, uint32_t width, uint32_t height, uint64_t fingerprint,
    std::vector<float>& rgb, uint32_t& sampleIndex, std::string* error);

// Background writer of render checkpoints with a fixed number of host
// buffers (AsyncCheckpointWriter, render_checkpoint.cpp:66-161).  enqueue()
// copies the renderer's framebuffer into a free buffer and writes it on a
// thread of its own; it returns false when every buffer is still being
// written.  A file is only ever replaced by one with at least as many
// samples.
class RenderCheckpointWriter {
public:
    RenderCheckpointWriter(uint32_t width, uint32_t height, std::string filename, uint32_t bufferCount,
        uint64_t fingerprint);
    ~RenderCheckpointWriter();
    RenderCheckpointWriter(const RenderCheckpointWriter&) = delete;
    RenderCheckpointWriter& operator=(const RenderCheckpointWriter&) = delete;

    bool enqueue(wpt_renderer* renderer, uint32_t sampleIndex);
    // Waits for every pending write; false if any failed.
    bool finish(std::string* error = nullptr);

private:
    struct Slot {
        std::vector<float> pixels;
        std::thread writer;
        std::atomic<bool> busy{false};
    };
    uint32_t width_, height_;
    std::string filename_;
    uint64_t fingerprint_;
    std::mutex mutex_;  // file writes and the fields below
    uint32_t latestWritten_ = 0;
    bool failed_ = false;
    std::string error_;
    std::vector<std::unique_ptr<Slot>> slots_;
};

// A job of the task ledger (DistributedJobConfig, checkpoint.hpp:9-27).
struct LedgerJob {
    uint64_t jobFingerprint = 0;
    uint32_t width = 0, height = 0, tileWidth = 0, tileHeight = 0, samplesPerPixel = 0, samplesPerTask = 0;
    bool operator==(const LedgerJob& o) const {
        return jobFingerprint == o.jobFingerprint && width == o.width && height == o.height &&
            tileWidth == o.tileWidth && tileHeight == o.tileHeight && samplesPerPixel == o.samplesPerPixel &&
            samplesPerTask == o.samplesPerTask;
    }
};

// The ledger's full state (DistributedCheckpoint, checkpoint.hpp:29-38).
struct LedgerCheckpoint {
    LedgerJob job;
    uint64_t sequence = 0;
    std::vector<float> radiance;         // 3 floats per pixel
    std::vector<uint32_t> sampleCounts;  // per pixel
    std::vector<uint8_t> completedTasks; // 0 or 1 per task
};

bool writeLedgerCheckpoint(const std::string& filename, const LedgerCheckpoint& checkpoint, std::string* error);
bool readLedgerCheckpoint(const std::string& filename, LedgerCheckpoint& checkpoint, std::string* error);

// Writes ledger checkpoints on a background thread, keeping only the newest
// snapshot that waits (DistributedCheckpointWriter,
// checkpoint_writer.hpp:11-40): every snapshot is a full state, so an older
// unwritten one adds nothing.
class LedgerCheckpointWriter {
public:
    explicit LedgerCheckpointWriter(std::string filename);
    ~LedgerCheckpointWriter();
    LedgerCheckpointWriter(const LedgerCheckpointWriter&) = delete;
    LedgerCheckpointWriter& operator=(const LedgerCheckpointWriter&) = delete;

    bool enqueue(LedgerCheckpoint checkpoint, std::string* error = nullptr);
    bool flush(std::string* error = nullptr);
    bool failed(std::string* error = nullptr) const;
    // Writes what is still queued, then stops the thread.
    void shutdown();

private:
    void run();

    std::string filename_;
    mutable std::mutex mutex_;
    std::condition_variable workAvailable_, idle_;
    std::optional<LedgerCheckpoint> pending_;
    bool writing_ = false, stopping_ = false, failed_ = false;
    std::string error_;
    std::thread thread_;
};

}  // namespace wpt
