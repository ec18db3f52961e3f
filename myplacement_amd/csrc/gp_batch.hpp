// gp_batch.hpp — many placements run in lockstep rounds on one device, every stage of their
// evaluation chains issued as one batched launch.
//
// Each running job is a controller session whose launches are recorded instead of issued (its
// steps suspend at every read-back of the device's scalars).  A round takes the launches every
// job recorded since its last read-back, packs them by stage into one table (one host-to-device
// copy), runs one batched kernel per stage and kernel variant over all jobs, one reduction
// finalize over all jobs' reductions, waits for the completion words, and resumes every job.
// A job's results are the same as when it runs alone: each batched workgroup runs the job's own
// launch's workgroup body.  Jobs join between rounds (continuous batching).

#pragma once

#include "mp_gp_backend.h"

#include <functional>
#include <memory>

namespace mpgp {

// Loads the library's device code on the device now (best effort: failures are left for the first
// job to report), so that a service's first job does not wait for it.
void loadDeviceCodeOn(int device) noexcept;

// Upper bound of the device bytes one global placement run of the design reserves.
uint64_t estimateBytes(const mp_design& design, const mp_gp_options& options);

// add, running and round are called from one thread (the driver's); the callbacks run on the
// driver's own host threads.
class BatchDriver {
public:
    // status, error message (empty on success), result
    using Callback = std::function<void(mp_gp_status, const char*, const mp_gp_result&)>;

    // memory_cap: the device memory the driver keeps for its jobs (blocks of finished jobs are kept
    // for later ones while the total stays below it); 0 = four fifths of the device
    explicit BatchDriver(int device, std::uint64_t memory_cap = 0);
    ~BatchDriver();
    BatchDriver(const BatchDriver&) = delete;
    BatchDriver& operator=(const BatchDriver&) = delete;

    // Starts a job (prepares the design, allocates its device state and records its first
    // launches).  centers: [2M] starting centres, overwritten with the result; it and the design's
    // arrays must stay valid until `done` is called (from add or from a later round).
    void add(const mp_design& design, const mp_gp_options& options, double* centers, Callback done);
    int running() const;
    // Runs one round; jobs that finish call their callback.  A failure of the round itself fails
    // every running job.
    void round();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace mpgp
