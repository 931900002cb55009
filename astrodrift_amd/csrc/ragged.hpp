// ragged.hpp -- descriptors shared by the ragged kernels and their host code.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <vector>

#include "../../include/drift_hip.h"
#include "drift_device.hpp"

struct drift_session;

namespace drift {

// Resident (resumable) state: kResidentWords 8-byte words per orbit, SoA.
constexpr int kResidentWords = 20;
constexpr int kResidentPhaseWord = 19;
constexpr int64_t kOrbitNew = 0;
constexpr int64_t kOrbitRunning = 1;
constexpr int64_t kOrbitDone = 2;

// One job as the kernels see it.
struct RaggedJob {
    const double* state0;       // (n, 6)
    const double* times;        // (nt,)
    double* out;                // (nt, n, 6)
    drift_orbit_stats* stats;   // (n,) or NULL
    double* resident;           // (kResidentWords, n) or NULL = run to completion
    unsigned int* done_count;   // orbits finished, or NULL
    int64_t n;
    int64_t item_begin;         // first work item of this job
    int32_t nt;
    int32_t method;
    int32_t stop_sample;        // pause an orbit once it has emitted this sample
    int32_t pot;                // index into the launch's potentials
    double rtol, atol;          // linear tolerances
    double dt_one_init;         // DOPR54
};

struct RaggedArgs {
    const RaggedJob* jobs;
    int32_t njobs;
    uint32_t max_attempts;
    int64_t items;
    const PotParams* pots;
    const void* const* tables64;
    const void* const* tables32;
    unsigned long long* queue;
    double bound2[6];
    int32_t lds_pot;       // potential whose MW table is staged in LDS, -1 = none
    const void* lds_src;   // that table in global memory (FP64 or the mixed mode's FP32)
    // DOPR54: each work item's initial sub-step, written by the estimate
    // kernel before the integrator runs (items whose job passes an explicit
    // dt_one_init, and orbits that resume, are not written)
    double* dt_est;
    // work item -> orbit within the item's job (cost order inside the jobs
    // that alone fill the device's lanes), or NULL: item - job.item_begin
    const uint32_t* order;
};

// Per-method launch scratch: the device copy of the descriptors (jobs | pots |
// table pointers), their pinned host staging (uploaded with an asynchronous
// copy; upload_done marks the end of the last eager upload) and the work-queue
// counter.
struct RaggedScratch {
    DeviceBuffer desc;
    DeviceBuffer dt_est;  // DOPR54: RaggedArgs::dt_est
    DeviceBuffer sort;    // cost order: keys and orbits (twice each) and the sort's scratch
    // the autotuner's timing of the integrator kernel alone (eager launches)
    bool time_kernel = false;
    hipEvent_t kernel_start = nullptr, kernel_end = nullptr;
    char* pinned = nullptr;
    size_t pinned_bytes = 0;
    hipEvent_t upload_done = nullptr;
    unsigned long long* queue = nullptr;
};

// One ragged launch, prepared on the host: descriptors written to the pinned
// staging, device buffers allocated, grid chosen.  ragged_enqueue then only
// records stream work (an upload, the queue reset, the kernel), so it may run
// under stream capture.
struct RaggedLaunch {
    RaggedArgs args{};
    void (*kernel)(RaggedArgs) = nullptr;
    void (*estimate)(RaggedArgs) = nullptr;  // DOPR54: fills args.dt_est first
    // cost order (drift_ragged_device only): keys from key_kernel for the
    // jobs of at least order_min_orbits orbits, sorted with their orbits
    void (*key_kernel)(RaggedArgs, uint32_t*, uint32_t*, int64_t) = nullptr;
    int64_t order_min_orbits = 0;
    uint32_t *keys_in = nullptr, *keys_out = nullptr, *orbits_in = nullptr, *orbits_out = nullptr;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    int sort_end_bit = 0;
    int64_t blocks = 0;
    int block = 0;
    size_t bytes = 0;  // descriptor bytes to upload
};

// order_min_orbits > 0: cost-order the work items of every job of at least
// that many orbits (eager launches only: the sort is not captured).
int ragged_prepare(drift_session* s, const RaggedJob* host_jobs, int32_t njobs, int method,
                   const std::vector<PotParams>& pots, const std::vector<const void*>& t64,
                   const std::vector<const void*>& t32, const drift_launch_options& opt,
                   RaggedScratch& scratch, RaggedLaunch& launch, int64_t order_min_orbits = 0);
int ragged_enqueue(const RaggedLaunch& launch, RaggedScratch& scratch, hipStream_t stream,
                   bool capturing);
int ragged_launch(drift_session* s, const RaggedJob* host_jobs, int32_t njobs, int method,
                  const std::vector<PotParams>& pots, const std::vector<const void*>& t64,
                  const std::vector<const void*>& t32, const drift_launch_options& opt,
                  hipStream_t stream, RaggedScratch& scratch, int64_t order_min_orbits = 0);
void ragged_scratch_free(RaggedScratch& scratch);
int ragged_clear(double* resident, int64_t n, unsigned int* done_count, hipStream_t stream);
int resolve_pots(drift_session* s, const drift_potential* potentials, size_t npots,
                 std::vector<PotParams>& pots, std::vector<const void*>& t64,
                 std::vector<const void*>& t32);

}  // namespace drift
