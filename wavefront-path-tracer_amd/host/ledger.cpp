#include "ledger.hpp"

#include <algorithm>
#include <chrono>
#include <iostream>
#include <limits>

namespace wpt {

LeaseLedger::LeaseLedger(const LedgerJob& job, const std::vector<RenderTask>& tasks)
    : job_(job), radiance_(3 * static_cast<size_t>(job.width) * job.height, 0.0f),
      sampleCounts_(static_cast<size_t>(job.width) * job.height, 0u) {
    for (const RenderTask& t : tasks)
        if (t.jobFingerprint == job_.jobFingerprint) tasks_.push_back(Record{t, LedgerTaskState::Pending, {}, 0});
    rebuildPending();
}

std::optional<LedgerLease> LeaseLedger::leaseNext(const std::string& workerId, uint64_t nowMs, uint64_t leaseMs) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (workerId.empty() || pending_.empty()) return std::nullopt;
    size_t i = pending_.front();
    pending_.pop_front();
    Record& r = tasks_[i];
    r.state = LedgerTaskState::Leased;
    r.workerId = workerId;
    r.expiresAtMs = nowMs + leaseMs;
    return LedgerLease{r.task, workerId, r.expiresAtMs};
}

LedgerCommit LeaseLedger::commit(const LedgerResult& result) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (result.jobFingerprint != job_.jobFingerprint) return LedgerCommit::Rejected;
    size_t i = find(result.taskId);
    if (i == tasks_.size()) return LedgerCommit::Rejected;
    Record& r = tasks_[i];
    if (r.state == LedgerTaskState::Completed) return LedgerCommit::Duplicate;
    const DistributedTile& tile = r.task.tile;
    bool sameTile = result.tile.x == tile.x && result.tile.y == tile.y && result.tile.width == tile.width &&
        result.tile.height == tile.height;
    if (!sameTile || result.sampleCount != r.task.id.sampleCount ||
        result.radiance.size() != 3 * static_cast<size_t>(tile.width) * tile.height)
        return LedgerCommit::Rejected;
    for (uint32_t y = 0; y < tile.height; ++y)
        for (uint32_t x = 0; x < tile.width; ++x) {
            size_t g = static_cast<size_t>(tile.y + y) * job_.width + tile.x + x;
            if (sampleCounts_[g] > std::numeric_limits<uint32_t>::max() - result.sampleCount)
                return LedgerCommit::Rejected;
        }
    for (uint32_t y = 0; y < tile.height; ++y)
        for (uint32_t x = 0; x < tile.width; ++x) {
            size_t l = static_cast<size_t>(y) * tile.width + x;
            size_t g = static_cast<size_t>(tile.y + y) * job_.width + tile.x + x;
            radiance_[3 * g] += result.radiance[3 * l];
            radiance_[3 * g + 1] += result.radiance[3 * l + 1];
            radiance_[3 * g + 2] += result.radiance[3 * l + 2];
            sampleCounts_[g] += result.sampleCount;
        }
    if (r.state == LedgerTaskState::Pending) {
        auto it = std::find(pending_.begin(), pending_.end(), i);
        if (it != pending_.end()) pending_.erase(it);
    }
    r.state = LedgerTaskState::Completed;
    r.workerId.clear();
    r.expiresAtMs = 0;
    ++completed_;
    return LedgerCommit::Accepted;
}

uint32_t LeaseLedger::renewWorkerLeases(const std::string& workerId, uint64_t nowMs, uint64_t leaseMs) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (workerId.empty()) return 0;
    uint32_t n = 0;
    for (Record& r : tasks_)
        if (r.state == LedgerTaskState::Leased && r.workerId == workerId) {
            r.expiresAtMs = nowMs + leaseMs;
            ++n;
        }
    return n;
}

uint32_t LeaseLedger::releaseWorkerLeases(const std::string& workerId) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (workerId.empty()) return 0;
    uint32_t n = 0;
    for (size_t i = 0; i < tasks_.size(); ++i) {
        Record& r = tasks_[i];
        if (r.state == LedgerTaskState::Leased && r.workerId == workerId) {
            r.state = LedgerTaskState::Pending;
            r.workerId.clear();
            r.expiresAtMs = 0;
            pending_.push_back(i);
            ++n;
        }
    }
    return n;
}

uint32_t LeaseLedger::requeueExpired(uint64_t nowMs) {
    std::lock_guard<std::mutex> lock(mutex_);
    uint32_t n = 0;
    for (size_t i = 0; i < tasks_.size(); ++i) {
        Record& r = tasks_[i];
        if (r.state == LedgerTaskState::Leased && r.expiresAtMs <= nowMs) {
            r.state = LedgerTaskState::Pending;
            r.workerId.clear();
            r.expiresAtMs = 0;
            pending_.push_back(i);
            ++n;
        }
    }
    return n;
}

bool LeaseLedger::restore(const LedgerCheckpoint& c, std::string* error) {
    std::lock_guard<std::mutex> lock(mutex_);
    if (!(c.job == job_)) {
        if (error) *error = "ledger checkpoint job configuration does not match the ledger";
        return false;
    }
    if (c.radiance.size() != radiance_.size() || c.sampleCounts.size() != sampleCounts_.size() ||
        c.completedTasks.size() != tasks_.size()) {
        if (error) *error = "ledger checkpoint dimensions or task count do not match the ledger";
        return false;
    }
    for (uint8_t done : c.completedTasks)
        if (done > 1) {
            if (error) *error = "ledger checkpoint completion bitmap is invalid";
            return false;
        }
    radiance_ = c.radiance;
    sampleCounts_ = c.sampleCounts;
    for (size_t i = 0; i < tasks_.size(); ++i)
        tasks_[i].state = c.completedTasks[i] ? LedgerTaskState::Completed : LedgerTaskState::Pending;
    rebuildPending();
    return true;
}

LedgerCheckpoint LeaseLedger::snapshot(uint64_t sequence) const {
    std::lock_guard<std::mutex> lock(mutex_);
    LedgerCheckpoint c;
    c.job = job_;
    c.sequence = sequence;
    c.radiance = radiance_;
    c.sampleCounts = sampleCounts_;
    c.completedTasks.resize(tasks_.size(), 0);
    for (size_t i = 0; i < tasks_.size(); ++i)
        if (tasks_[i].state == LedgerTaskState::Completed) c.completedTasks[i] = 1;
    return c;
}

size_t LeaseLedger::pendingTaskCount() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return pending_.size();
}

size_t LeaseLedger::completedTaskCount() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return completed_;
}

bool LeaseLedger::complete() const {
    std::lock_guard<std::mutex> lock(mutex_);
    return completed_ == tasks_.size();
}

LedgerTaskState LeaseLedger::taskState(const RenderTaskId& id) const {
    std::lock_guard<std::mutex> lock(mutex_);
    size_t i = find(id);
    return i == tasks_.size() ? LedgerTaskState::Pending : tasks_[i].state;
}

size_t LeaseLedger::find(const RenderTaskId& id) const {
    for (size_t i = 0; i < tasks_.size(); ++i)
        if (tasks_[i].task.id == id) return i;
    return tasks_.size();
}

void LeaseLedger::rebuildPending() {
    pending_.clear();
    completed_ = 0;
    for (size_t i = 0; i < tasks_.size(); ++i) {
        if (tasks_[i].state == LedgerTaskState::Completed) {
            ++completed_;
        } else {
            tasks_[i].state = LedgerTaskState::Pending;
            tasks_[i].workerId.clear();
            tasks_[i].expiresAtMs = 0;
            pending_.push_back(i);
        }
    }
}

bool writeLedgerImage(const std::string& filename, const LedgerCheckpoint& c) {
    std::cout << "Raw framebuffer SHA-256: " << sha256Hex(c.radiance.data(), c.radiance.size() * sizeof(float)) << '\n';
    std::vector<float> pixels = c.radiance;
    for (size_t i = 0; i < c.sampleCounts.size(); ++i) {
        uint32_t n = c.sampleCounts[i];
        if (n == 0) continue;
        for (int k = 0; k < 3; ++k) pixels[3 * i + k] = pixels[3 * i + k] / static_cast<float>(n);
    }
    if (!writePpm(filename, pixels, c.job.width, c.job.height)) return false;
    std::cout << "Wrote " << filename << '\n';
    return true;
}

namespace {
uint64_t nowMilliseconds() {
    return static_cast<uint64_t>(std::chrono::duration_cast<std::chrono::milliseconds>(
        std::chrono::steady_clock::now().time_since_epoch()).count());
}
}  // namespace

int runLocalLeases(RenderDriver& driver, const RenderConfig& config, const SceneConfig& scene) {
    const LedgerJob job{driver.fingerprint(), config.width, config.height, config.tileWidth, config.tileHeight,
        config.samplesPerPixel, config.distributedSamplesPerTask};
    std::vector<RenderTask> tasks = createRenderTasks(job.width, job.height, job.tileWidth, job.tileHeight,
        job.samplesPerPixel, job.samplesPerTask, job.jobFingerprint);
    if (tasks.empty()) {
        std::cerr << "Task ledger produced no tasks; check render dimensions and sample settings\n";
        return 1;
    }
    LeaseLedger ledger(job, tasks);
    uint64_t sequence = 0;
    if (config.resumeFromCheckpoint) {
        LedgerCheckpoint saved;
        std::string error;
        if (!readLedgerCheckpoint(config.distributedCheckpointFilename, saved, &error) ||
            !ledger.restore(saved, &error)) {
            std::cerr << "Could not resume ledger checkpoint: " << error << '\n';
            return 1;
        }
        sequence = saved.sequence;
        std::cout << "Resumed ledger checkpoint sequence " << saved.sequence << ": " << ledger.completedTaskCount()
                  << '/' << ledger.taskCount() << " tasks completed\n";
    }
    if (!driver.prepare()) return 1;
    std::cout << "Local task ledger: " << ledger.taskCount() << " tasks of " << job.samplesPerTask
              << " samples, checkpoint " << config.distributedCheckpointFilename << '\n';
    const std::string worker = config.workerId.empty() ? "local-gpu" : config.workerId;
    RenderTaskRenderer renderer(driver);
    LedgerCheckpointWriter writer(config.distributedCheckpointFilename);
    std::string error;
    auto queueSnapshot = [&]() {
        LedgerCheckpoint c = ledger.snapshot(++sequence);
        size_t done = 0;
        for (uint8_t d : c.completedTasks) done += d;
        if (!writer.enqueue(std::move(c), &error)) {
            std::cerr << "Could not queue ledger checkpoint: " << error << '\n';
            return false;
        }
        std::cout << "Queued checkpoint " << sequence << ": " << done << '/' << ledger.taskCount() << " tasks\n";
        return true;
    };
    const uint64_t intervalMs = static_cast<uint64_t>(config.distributedCheckpointIntervalSeconds) * 1000u;
    uint64_t start = nowMilliseconds(), nextCheckpoint = start;
    while (!ledger.complete()) {
        uint64_t now = nowMilliseconds();
        if (writer.failed(&error)) {
            std::cerr << "Ledger checkpoint writer failed: " << error << '\n';
            return 1;
        }
        if (now >= nextCheckpoint) {
            if (!queueSnapshot()) return 1;
            nextCheckpoint = now + intervalMs;
        }
        ledger.requeueExpired(now);
        std::optional<LedgerLease> lease = ledger.leaseNext(worker, now, config.distributedLeaseDurationMs);
        if (!lease) break;  // every remaining task is leased: none can be, with one worker
        RenderTaskResult r = renderer.render(lease->task);
        if (!r.success) {
            ledger.releaseWorkerLeases(worker);
            std::cerr << "Task " << lease->task.id.tileOrdinal << '/' << lease->task.id.sampleStart << " failed\n";
            return 1;
        }
        LedgerResult result{lease->task.jobFingerprint, r.taskId, r.tile, r.sampleCount, std::move(r.radiance)};
        if (ledger.commit(result) == LedgerCommit::Rejected) {
            std::cerr << "Task ledger rejected a result\n";
            return 1;
        }
    }
    if (!ledger.complete()) {
        std::cerr << "Task ledger stopped with unfinished tasks\n";
        return 1;
    }
    if (!writer.flush(&error)) {
        std::cerr << "Could not flush ledger checkpoint: " << error << '\n';
        return 1;
    }
    writer.shutdown();
    LedgerCheckpoint final = ledger.snapshot(++sequence);
    if (!writeLedgerCheckpoint(config.distributedCheckpointFilename, final, &error)) {
        std::cerr << "Could not write final ledger checkpoint: " << error << '\n';
        return 1;
    }
    if (!writeLedgerImage(config.outputFilename, final)) return 1;
    std::cout << "Ledger trace time (first task to final image): " << nowMilliseconds() - start << " ms\n";
    std::cout << "Ledger render complete\n";
    return 0;
}

}  // namespace wpt
