// drift_internal.hpp -- declarations shared by the host-side translation units
// of libdrift_hip.so.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/drift_hip.h"
#include "drift_device.hpp"

namespace drift {

struct TableEntry {
    drift_potential key;
    PotParams params;
    double* d_table64 = nullptr;
    float* d_table32 = nullptr;
    double* d_far = nullptr;  // FAR_DOUBLES, global memory
};

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// A moving-origin set uploaded by drift_origins_create.
struct OriginSet {
    int32_t id = 0;
    double* d_tracks = nullptr;  // n_origins x n_segments x ORIGIN_COEFS
    double* d_params = nullptr;  // n_origins x (amp, b^2)
    int32_t n_origins = 0, n_segments = 0;
    double t_start = 0.0, t_end = 0.0;
};

// A launch configuration chosen by drift_tune_device.
struct TunedEntry {
    int method, kind, precision, log2n;
    drift_launch_options options;
};

}  // namespace drift

#include "ragged.hpp"

struct drift_session {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    std::vector<drift::TableEntry> tables;
    drift::DeviceBuffer state0, times, out, stats;
    // ring of work-queue counters for persistent launches (one per launch in
    // flight; launches further apart than the ring reuse a counter only after
    // stream order has retired its previous user)
    unsigned long long* queues = nullptr;
    unsigned queue_next = 0;
    std::vector<bool> queue_owned;  // counters held by captured graphs
    // descriptor buffers of drift_ragged_device, one per method
    drift::RaggedScratch ragged[2];
    // drift_tune_ragged's block size per [precision][method] (0 = untuned: 256)
    int32_t ragged_block[2][2] = {{0, 0}, {0, 0}};
    // cost ordering: per-orbit bin keys, the slot -> orbit map and the bin
    // counters; order_done marks the end of the last launch that read them
    drift::DeviceBuffer order_keys, order_slots, order_bins;
    hipEvent_t order_done = nullptr;
    // autotuner: chosen configurations and the scratch output of its trials
    std::vector<drift::TunedEntry> tuned;
    drift::DeviceBuffer tune_out, tune_stats;
    // output staging (DRIFT_STAGING_STAGED): orbit-blocked sample buffer;
    // staging_done marks the end of the last launch that used it
    drift::DeviceBuffer staging;
    hipEvent_t staging_done = nullptr;
    // time-windowed moving-origin launches: the orbits' resumable state
    // between windows; resume_done marks the end of the last launch using it
    drift::DeviceBuffer resume;
    hipEvent_t resume_done = nullptr;
    // drift_force's device copies of its host buffers
    drift::DeviceBuffer force_io;
    // moving-origin sets (drift_origins_create)
    std::vector<drift::OriginSet> origin_sets;
    int32_t next_origin_id = 1;
    // the integrator shape of the last launch, and the phases of the last
    // host-buffer launch (drift_session_last_launch), timed by phase_events
    int32_t last_grid = 0, last_block = 0;
    drift_launch_report last_launch{};
    bool have_last_launch = false;
    hipEvent_t phase_events[4] = {nullptr, nullptr, nullptr, nullptr};
};

#define DRIFT_HIP_TRY(expr)                                                              \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            drift::set_last_error(std::string("HIP error: ") + hipGetErrorString(_e) +    \
                                  " in " + #expr);                                        \
            return _e == hipErrorOutOfMemory ? DRIFT_ERR_OUT_OF_MEMORY : DRIFT_ERR_DEVICE; \
        }                                                                                 \
    } while (0)

namespace drift {

constexpr unsigned kQueueRing = 64;

// Error helpers (drift_hip.hip).
int fail(int code, const std::string& msg);
int ensure_buffer(DeviceBuffer& b, size_t bytes);
int validate_potential(const drift_potential& p);
// Kernel parameters (and the cached device table for MW) of a potential
// descriptor; NULL selects the unit galpy Kepler force.
int resolve_potential(drift_session* s, const drift_potential* potential, PotParams& params,
                      const void** table64, const void** table32);

// Milky-Way spherical table (spherical_table.cpp).
void build_mw_params(const drift_potential& p, PotParams& out);
void build_mw_table(const drift_potential& p, double* coeffs);
// The mixed mode's float table: TABLE32_FLOATS values, TABLE32_NCOEF per
// TABLE32_STRIDE-float row.
void build_mw_table_f32(const drift_potential& p, float* coeffs);
double mw_table_max_rel_error(const drift_potential& p, const double* coeffs, int samples_per_seg);
void build_mw_far_table(const drift_potential& p, double* coeffs);
double mw_far_table_max_rel_error(const drift_potential& p, const double* coeffs,
                                  int samples_per_seg);
double mw_spherical_exact(const drift_potential& p, double r2);

// Quintic Hermite tracks (origins.cpp).
int origin_tracks_hermite(const double* states, const double* accels, const double* times,
                          size_t nt, size_t n, double* tracks, double* t_start, double* t_end);

// Thread-local error message (drift_last_error).
void set_last_error(const std::string& msg);

}  // namespace drift
