#include "checkpoint.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <utility>

namespace wpt {

namespace {

void setError(std::string* error, const std::string& message) {
    if (error) *error = message;
}

// ---- render checkpoints (render_checkpoint.cpp:17-64)

constexpr char kRenderMagic[8] = {'P', 'T', 'C', 'H', 'K', 'P', 'T', '1'};
constexpr uint32_t kRenderVersion = 3;

struct RenderHeader {
    char magic[8];
    uint32_t version;
    uint32_t width;
    uint32_t height;
    uint32_t sampleIndex;
    uint64_t fingerprint;
};
static_assert(sizeof(RenderHeader) == 32, "render checkpoint header is 32 bytes");

// Writes `parts` to filename + ".tmp" and renames it over filename.
bool replaceFile(const std::string& filename, const std::vector<std::pair<const void*, size_t>>& parts,
    std::string* error, const char* what) {
    std::string temporary = filename + ".tmp";
    {
        std::ofstream out(temporary, std::ios::binary | std::ios::trunc);
        if (!out) {
            setError(error, std::string("could not open ") + what + " for writing: " + temporary);
            return false;
        }
        for (const auto& p : parts) out.write(static_cast<const char*>(p.first), static_cast<std::streamsize>(p.second));
        out.flush();
        out.close();
        if (!out) {
            setError(error, std::string("could not write ") + what + ": " + temporary);
            return false;
        }
    }
    if (std::rename(temporary.c_str(), filename.c_str()) != 0) {
        setError(error, std::string("could not replace ") + what + ": " + filename);
        return false;
    }
    return true;
}

// ---- ledger checkpoints (src/distributed/checkpoint.cpp:9-84)

constexpr char kLedgerMagic[8] = {'P', 'T', 'D', 'I', 'S', 'T', '1', '\0'};
constexpr uint32_t kLedgerVersion = 1;

struct LedgerHeader {
    char magic[8];
    uint32_t version;
    uint32_t headerSize;
    uint64_t jobFingerprint;
    uint32_t width, height, tileWidth, tileHeight, samplesPerPixel, samplesPerTask;
    uint64_t sequence;
    uint64_t pixelCount;
    uint64_t taskCount;
    uint64_t payloadBytes;
    uint64_t payloadChecksum;
};
static_assert(sizeof(LedgerHeader) == 88, "ledger checkpoint header is 88 bytes");

uint64_t fnvBytes(uint64_t hash, const void* data, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(data);
    for (size_t i = 0; i < n; ++i) {
        hash ^= b[i];
        hash *= 1099511628211ull;
    }
    return hash;
}

// FNV-1a over radiance, sample counts and completion bytes in that order.
uint64_t payloadChecksum(const LedgerCheckpoint& c) {
    uint64_t h = 1469598103934665603ull;
    h = fnvBytes(h, c.radiance.data(), c.radiance.size() * sizeof(float));
    h = fnvBytes(h, c.sampleCounts.data(), c.sampleCounts.size() * sizeof(uint32_t));
    h = fnvBytes(h, c.completedTasks.data(), c.completedTasks.size());
    return h;
}

bool checkedProduct(uint64_t a, uint64_t b, uint64_t& out) {
    if (b != 0 && a > std::numeric_limits<uint64_t>::max() / b) return false;
    out = a * b;
    return true;
}

bool validShape(const LedgerCheckpoint& c, std::string* error) {
    uint64_t pixels = 0;
    if (!checkedProduct(c.job.width, c.job.height, pixels) || pixels > std::numeric_limits<uint64_t>::max() / 12) {
        setError(error, "ledger checkpoint dimensions overflow");
        return false;
    }
    if (c.radiance.size() != 3 * pixels || c.sampleCounts.size() != pixels) {
        setError(error, "ledger checkpoint pixel buffers have inconsistent sizes");
        return false;
    }
    for (uint8_t done : c.completedTasks)
        if (done > 1) {
            setError(error, "ledger checkpoint completion bitmap contains an invalid value");
            return false;
        }
    return true;
}

}  // namespace

bool writeRenderCheckpoint(const std::string& filename, const float* rgb, uint32_t width, uint32_t height,
    uint32_t sampleIndex, uint64_t fingerprint, std::string* error) {
    RenderHeader h{};
    std::memcpy(h.magic, kRenderMagic, sizeof h.magic);
    h.version = kRenderVersion;
    h.width = width;
    h.height = height;
    h.sampleIndex = sampleIndex;
    h.fingerprint = fingerprint;
    size_t bytes = sizeof(float) * 3 * static_cast<size_t>(width) * height;
    return replaceFile(filename, {{&h, sizeof h}, {rgb, bytes}}, error, "checkpoint");
}

bool readRenderCheckpoint(const std::string& filename, uint32_t width, uint32_t height, uint64_t fingerprint,
    std::vector<float>& rgb, uint32_t& sampleIndex, std::string* error) {
    std::ifstream in(filename, std::ios::binary);
    if (!in) {
        setError(error, "Could not open checkpoint: " + filename);
        return false;
    }
    RenderHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof h);
    if (!in) {
        setError(error, "Checkpoint is truncated or unreadable: " + filename);
        return false;
    }
    if (std::memcmp(h.magic, kRenderMagic, sizeof h.magic) != 0 || h.version != kRenderVersion) {
        setError(error, "Unsupported checkpoint format: " + filename);
        return false;
    }
    if (h.width != width || h.height != height) {
        setError(error, "Checkpoint resolution does not match the current render: " + filename);
        return false;
    }
    if (h.fingerprint != fingerprint) {
        setError(error, "Checkpoint scene/config fingerprint does not match the current render: " + filename);
        return false;
    }
    std::vector<float> pixels(3 * static_cast<size_t>(width) * height);
    in.read(reinterpret_cast<char*>(pixels.data()), static_cast<std::streamsize>(sizeof(float) * pixels.size()));
    if (!in) {
        setError(error, "Checkpoint is truncated or unreadable: " + filename);
        return false;
    }
    rgb = std::move(pixels);
    sampleIndex = h.sampleIndex;
    return true;
}

RenderCheckpointWriter::RenderCheckpointWriter(uint32_t width, uint32_t height, std::string filename,
    uint32_t bufferCount, uint64_t fingerprint)
    : width_(width), height_(height), filename_(std::move(filename)), fingerprint_(fingerprint) {
    for (uint32_t i = 0; i < bufferCount; ++i) {
        slots_.push_back(std::make_unique<Slot>());
        slots_.back()->pixels.resize(3 * static_cast<size_t>(width) * height);
    }
}

RenderCheckpointWriter::~RenderCheckpointWriter() { finish(); }

bool RenderCheckpointWriter::enqueue(wpt_renderer* renderer, uint32_t sampleIndex) {
    for (auto& owned : slots_) {
        Slot& slot = *owned;
        bool expected = false;
        if (!slot.busy.compare_exchange_strong(expected, true)) continue;
        if (slot.writer.joinable()) slot.writer.join();
        if (wpt_renderer_read_framebuffer(renderer, slot.pixels.data(), nullptr) != WPT_OK) {
            std::lock_guard<std::mutex> lock(mutex_);
            failed_ = true;
            error_ = wpt_last_error();
            slot.busy.store(false);
            return false;
        }
        slot.writer = std::thread([this, &slot, sampleIndex] {
            {
                std::lock_guard<std::mutex> lock(mutex_);
                if (sampleIndex >= latestWritten_) {
                    std::string error;
                    if (writeRenderCheckpoint(filename_, slot.pixels.data(), width_, height_, sampleIndex,
                            fingerprint_, &error))
                        latestWritten_ = sampleIndex;
                    else if (!failed_) {
                        failed_ = true;
                        error_ = error;
                    }
                }
            }
            slot.busy.store(false);
        });
        return true;
    }
    return false;
}

bool RenderCheckpointWriter::finish(std::string* error) {
    for (auto& slot : slots_)
        if (slot->writer.joinable()) slot->writer.join();
    std::lock_guard<std::mutex> lock(mutex_);
    if (failed_) setError(error, error_);
    return !failed_;
}

bool writeLedgerCheckpoint(const std::string& filename, const LedgerCheckpoint& c, std::string* error) {
    if (filename.empty()) {
        setError(error, "ledger checkpoint filename is empty");
        return false;
    }
    if (!validShape(c, error)) return false;
    uint64_t pixels = c.sampleCounts.size();
    uint64_t radianceBytes = pixels * 3 * sizeof(float);
    uint64_t sampleBytes = pixels * sizeof(uint32_t);
    uint64_t taskBytes = c.completedTasks.size();
    LedgerHeader h{};
    std::memcpy(h.magic, kLedgerMagic, sizeof h.magic);
    h.version = kLedgerVersion;
    h.headerSize = sizeof(LedgerHeader);
    h.jobFingerprint = c.job.jobFingerprint;
    h.width = c.job.width;
    h.height = c.job.height;
    h.tileWidth = c.job.tileWidth;
    h.tileHeight = c.job.tileHeight;
    h.samplesPerPixel = c.job.samplesPerPixel;
    h.samplesPerTask = c.job.samplesPerTask;
    h.sequence = c.sequence;
    h.pixelCount = pixels;
    h.taskCount = taskBytes;
    h.payloadBytes = radianceBytes + sampleBytes + taskBytes;
    h.payloadChecksum = payloadChecksum(c);
    return replaceFile(filename,
        {{&h, sizeof h}, {c.radiance.data(), radianceBytes}, {c.sampleCounts.data(), sampleBytes},
            {c.completedTasks.data(), taskBytes}},
        error, "ledger checkpoint");
}

bool readLedgerCheckpoint(const std::string& filename, LedgerCheckpoint& out, std::string* error) {
    if (filename.empty()) {
        setError(error, "ledger checkpoint filename is empty");
        return false;
    }
    std::ifstream in(filename, std::ios::binary);
    if (!in) {
        setError(error, "could not open ledger checkpoint");
        return false;
    }
    LedgerHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof h);
    if (!in || std::memcmp(h.magic, kLedgerMagic, sizeof h.magic) != 0 || h.version != kLedgerVersion ||
        h.headerSize != sizeof(LedgerHeader)) {
        setError(error, "unsupported or truncated ledger checkpoint header");
        return false;
    }
    uint64_t pixels = 0;
    if (!checkedProduct(h.width, h.height, pixels) || pixels != h.pixelCount ||
        pixels > std::numeric_limits<uint64_t>::max() / 16) {
        setError(error, "ledger checkpoint dimensions are invalid");
        return false;
    }
    uint64_t radianceBytes = pixels * 3 * sizeof(float);
    uint64_t sampleBytes = pixels * sizeof(uint32_t);
    if (h.taskCount > std::numeric_limits<uint64_t>::max() - radianceBytes - sampleBytes ||
        h.payloadBytes != radianceBytes + sampleBytes + h.taskCount) {
        setError(error, "ledger checkpoint payload size is invalid");
        return false;
    }
    in.seekg(0, std::ios::end);
    std::streamoff size = in.tellg();
    if (size < 0 || static_cast<uint64_t>(size) != sizeof(LedgerHeader) + h.payloadBytes) {
        setError(error, "ledger checkpoint file is truncated or has trailing data");
        return false;
    }
    in.seekg(sizeof(LedgerHeader), std::ios::beg);
    LedgerCheckpoint c;
    c.job = LedgerJob{h.jobFingerprint, h.width, h.height, h.tileWidth, h.tileHeight, h.samplesPerPixel,
        h.samplesPerTask};
    c.sequence = h.sequence;
    c.radiance.resize(static_cast<size_t>(3 * pixels));
    c.sampleCounts.resize(static_cast<size_t>(pixels));
    c.completedTasks.resize(static_cast<size_t>(h.taskCount));
    in.read(reinterpret_cast<char*>(c.radiance.data()), static_cast<std::streamsize>(radianceBytes));
    in.read(reinterpret_cast<char*>(c.sampleCounts.data()), static_cast<std::streamsize>(sampleBytes));
    in.read(reinterpret_cast<char*>(c.completedTasks.data()), static_cast<std::streamsize>(h.taskCount));
    if (!in) {
        setError(error, "ledger checkpoint payload is truncated");
        return false;
    }
    if (payloadChecksum(c) != h.payloadChecksum) {
        setError(error, "ledger checkpoint payload checksum does not match");
        return false;
    }
    if (!validShape(c, error)) return false;
    out = std::move(c);
    return true;
}

LedgerCheckpointWriter::LedgerCheckpointWriter(std::string filename)
    : filename_(std::move(filename)), thread_(&LedgerCheckpointWriter::run, this) {}

LedgerCheckpointWriter::~LedgerCheckpointWriter() { shutdown(); }

bool LedgerCheckpointWriter::enqueue(LedgerCheckpoint checkpoint, std::string* error) {
    {
        std::lock_guard<std::mutex> lock(mutex_);
        if (failed_) {
            setError(error, error_);
            return false;
        }
        if (stopping_) {
            setError(error, "ledger checkpoint writer is stopped");
            return false;
        }
        pending_ = std::move(checkpoint);
    }
    workAvailable_.notify_one();
    return true;
}

bool LedgerCheckpointWriter::flush(std::string* error) {
    std::unique_lock<std::mutex> lock(mutex_);
    idle_.wait(lock, [this] { return (!pending_ && !writing_) || failed_; });
    if (failed_) {
        setError(error, error_);
        return false;
    }
    return true;
}

bool LedgerCheckpointWriter::failed(std::string* error) const {
    std::lock_guard<std::mutex> lock(mutex_);
    if (failed_) setError(error, error_);
    return failed_;
}

void LedgerCheckpointWriter::shutdown() {
    {
        std::lock_guard<std::mutex> lock(mutex_);
        if (!thread_.joinable()) return;
        stopping_ = true;
    }
    workAvailable_.notify_one();
    thread_.join();
}

void LedgerCheckpointWriter::run() {
    for (;;) {
        LedgerCheckpoint checkpoint;
        {
            std::unique_lock<std::mutex> lock(mutex_);
            workAvailable_.wait(lock, [this] { return stopping_ || pending_.has_value(); });
            if (!pending_) break;  // stopping with nothing left to write
            checkpoint = std::move(*pending_);
            pending_.reset();
            writing_ = true;
        }
        std::string error;
        bool ok = writeLedgerCheckpoint(filename_, checkpoint, &error);
        {
            std::lock_guard<std::mutex> lock(mutex_);
            writing_ = false;
            if (!ok) {
                failed_ = true;
                error_ = error.empty() ? "ledger checkpoint write failed" : error;
                pending_.reset();
                stopping_ = true;
            }
        }
        idle_.notify_all();
        if (!ok) break;
    }
    idle_.notify_all();
}

}  // namespace wpt
