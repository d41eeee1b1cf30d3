// Task ledger: the reference's coordinator state kept on the host of one
// GPU, which leases the job's tasks to local workers, merges their tile
// sums, re-queues leases that expire, and snapshots itself into ledger
// checkpoints (DistributedCoordinator, include/distributed/coordinator.hpp,
// src/distributed/coordinator.cpp:13-246).  No networking is involved: the
// worker is the render loop of this process (runLocalLeases).
#pragma once

#include <cstdint>
#include <deque>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "render_driver.hpp"

namespace wpt {

enum class LedgerTaskState : uint8_t { Pending, Leased, Completed };
enum class LedgerCommit : uint8_t { Accepted, Duplicate, Rejected };

struct LedgerLease {
    RenderTask task;
    std::string workerId;
    uint64_t expiresAtMs = 0;
};

// A finished task: unnormalized RGB sums of its tile (DistributedTaskResult).
struct LedgerResult {
    uint64_t jobFingerprint = 0;
    RenderTaskId taskId{};
    DistributedTile tile{};
    uint32_t sampleCount = 0;
    std::vector<float> radiance;  // 3 floats per tile pixel
};

class LeaseLedger {
public:
    // Tasks of another job (fingerprint) are dropped.
    LeaseLedger(const LedgerJob& job, const std::vector<RenderTask>& tasks);

    // The oldest pending task, leased to workerId until nowMs + leaseMs.
    std::optional<LedgerLease> leaseNext(const std::string& workerId, uint64_t nowMs, uint64_t leaseMs);
    // Merges a result once: a second commit of a task is Duplicate; a result
    // of another job, an unknown task, a different tile or sample count, a
    // radiance buffer of the wrong size or a per-pixel sample count that
    // would overflow is Rejected.
    LedgerCommit commit(const LedgerResult& result);
    uint32_t renewWorkerLeases(const std::string& workerId, uint64_t nowMs, uint64_t leaseMs);
    uint32_t releaseWorkerLeases(const std::string& workerId);
    // Leases whose expiry is at or before nowMs return to the queue.
    uint32_t requeueExpired(uint64_t nowMs);
    // Completed tasks and sums from a checkpoint of the same job; every other
    // task becomes pending.
    bool restore(const LedgerCheckpoint& checkpoint, std::string* error);
    LedgerCheckpoint snapshot(uint64_t sequence) const;

    const LedgerJob& job() const { return job_; }
    const std::vector<float>& radiance() const { return radiance_; }
    const std::vector<uint32_t>& sampleCounts() const { return sampleCounts_; }
    size_t taskCount() const { return tasks_.size(); }
    size_t pendingTaskCount() const;
    size_t completedTaskCount() const;
    bool complete() const;
    // Pending for an unknown task.
    LedgerTaskState taskState(const RenderTaskId& id) const;

private:
    struct Record {
        RenderTask task;
        LedgerTaskState state = LedgerTaskState::Pending;
        std::string workerId;
        uint64_t expiresAtMs = 0;
    };
    size_t find(const RenderTaskId& id) const;
    void rebuildPending();

    LedgerJob job_;
    std::vector<Record> tasks_;
    std::vector<float> radiance_;
    std::vector<uint32_t> sampleCounts_;
    std::deque<size_t> pending_;
    size_t completed_ = 0;
    mutable std::mutex mutex_;
};

// Normalizes per-pixel sums by their sample counts and writes the PPM (a
// pixel without samples stays black); prints the raw sums' SHA-256 first
// (writeDistributedImage, coordinator_app.cpp:19-37).
bool writeLedgerImage(const std::string& filename, const LedgerCheckpoint& checkpoint);

// The coordinator role on one GPU: the task table of the job, a ledger
// (resumed from the ledger checkpoint with --resume), one local worker that
// leases, renders and commits tasks through the driver, checkpoints every
// distributedCheckpointIntervalSeconds and at the end, and the final image.
int runLocalLeases(RenderDriver& driver, const RenderConfig& config, const SceneConfig& scene);

}  // namespace wpt
