/*
 * drift_hip.h -- C ABI of the MI355X (gfx950) batched Dormand-Prince integrator.
 *
 * Drop-in replacement for the CUDA backend of drift's GPU engine.  The entry
 * points carry the reference kernel ABI (kernels/src/lib.rs:143-216):
 *   state0     n x 6 doubles, AoS rows (x, y, z, vx, vy, vz)
 *   times      nt doubles (the output grid, src/dispatch/mod.rs:14-18)
 *   state_out  nt x n x 6 doubles, time-major (kernels/src/lib.rs:20-33)
 *   rtol, atol NATURAL LOGARITHMS of the tolerances, as galpy takes them
 *              (src/integrators/galpy/mod.rs:17-24)
 *   dt_one_init  DOPR54 initial sub-step; -9999.99 selects the RK4
 *              step-doubling estimate (src/integrators/galpy/mod.rs:8)
 * and the launch contract of the reference (n >= 1, nt >= 2; the reference's
 * nt <= 1024 limit is lifted).  Violations return DRIFT_ERR_INVALID_ARGUMENT
 * with a message from drift_last_error(), never a device fault.
 *
 * Additions over the reference ABI (none changes the meaning of the above):
 *   - a potential descriptor (NULL = the reference's hard-coded unit galpy
 *     Kepler force, src/dispatch/gpu.rs:155-156);
 *   - per-orbit statistics and status (the reference GPU has no error channel);
 *   - a session handle owning the device, stream and cached buffers
 *     (the reference creates a context per launch, gpu/cuda_oxide.rs:16-18);
 *   - device-resident variants that take device pointers and a HIP stream;
 *   - moving origins (point masses and Plummer spheres on tabulated quintic
 *     tracks), the reference's CustomOrigin (shared/src/potential.rs:264-356),
 *     whose recipes never construct it (the FIXMEs of shared/src/model.rs:
 *     193-199, 226-235).
 *
 * All functions return DRIFT_OK (0) or a drift_status code.  Thread safety: a
 * session may be used from one thread at a time; distinct sessions are
 * independent.
 */
#ifndef DRIFT_HIP_H
#define DRIFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 6: drift_launch_shape, drift_session_last_launch, drift_server_busy,
 * drift_galpy_dop853_device_noalloc.
 * 5: drift_tune_ragged, drift_server_stats.  4: moving origins, drift_force,
 * the NONE potential.  3: drift_launch_options.staging.  2: .ordering.  A
 * binding checks drift_get_abi_version() against the header it was built with. */
#define DRIFT_ABI_VERSION 6

/* Status codes.  They map onto the reference's error taxonomy
 * (DispatchError, src/dispatch/mod.rs:36-79; IntegrationError,
 * src/integrators/mod.rs:10-20). */
typedef enum drift_status {
    DRIFT_OK = 0,
    DRIFT_ERR_INVALID_ARGUMENT = 1, /* launch contract violated (kernels/src/lib.rs:129-152) */
    DRIFT_ERR_UNSUPPORTED = 2,      /* UnsupportedConfiguration analogue */
    DRIFT_ERR_DEVICE = 3,           /* DispatchError::Cuda analogue (HIP runtime error) */
    DRIFT_ERR_INTEGRATION = 4,      /* an orbit returned a non-zero DOPR54 error code
                                       (DispatchError::CpuIntegration analogue) */
    DRIFT_ERR_OUT_OF_MEMORY = 5,
    DRIFT_ERR_TIMEOUT = 6            /* drift_server_wait: not finished in time */
} drift_status;

typedef enum drift_method {
    DRIFT_METHOD_DOPR54 = 0, /* galpy bovy_rk.c DOPR54, src/integrators/galpy/dopr54.rs */
    DRIFT_METHOD_DOP853 = 1  /* galpy leung_dop853.c DOP853, src/integrators/galpy/dop853.rs */
} drift_method;

typedef enum drift_potential_kind {
    DRIFT_POTENTIAL_GALPY_KEPLER = 0, /* unit point mass, galpy cylindrical form (dopr54.rs:75-88) */
    DRIFT_POTENTIAL_KEPLER = 1,       /* shared/src/potential.rs:217-239 */
    DRIFT_POTENTIAL_PLUMMER = 2,      /* shared/src/potential.rs:241-262 */
    DRIFT_POTENTIAL_MW = 3,           /* bulge + Miyamoto-Nagai disk + NFW halo,
                                         shared/src/potential.rs:87-215 */
    DRIFT_POTENTIAL_NONE = 4          /* no static force: moving origins only, or free motion */
} drift_potential_kind;

typedef enum drift_precision {
    DRIFT_PRECISION_FP64 = 0,  /* everything in FP64 (default) */
    DRIFT_PRECISION_MIXED = 1  /* stage forces in FP32; state update, error norm and
                                  step control in FP64 (requires rtol >= 1e-7) */
} drift_precision;

/* Potential descriptor (G = 1 natural units).  Unused fields are ignored.
 * MW: bulge rho = bulge_amp r^-alpha exp(-(r/rc)^2) (0 <= alpha < 3),
 *     Miyamoto-Nagai (disk_amp, disk_a, disk_b), NFW (halo_amp, halo_a).
 * origins: 0, or a moving-origin set of the launching session
 * (drift_origins_create) whose force is added to the static kind's
 * (GALPY_KEPLER excepted).  Moving origins need FP64 precision and the
 * single-batch entry points (not drift_ragged_device or the server). */
typedef struct drift_potential {
    int32_t kind;      /* drift_potential_kind */
    int32_t origins;   /* moving-origin set id, 0 = none */
    double amp;        /* Kepler / Plummer amplitude */
    double b;          /* Plummer scale radius */
    double bulge_amp, bulge_alpha, bulge_rc;
    double disk_amp, disk_a, disk_b;
    double halo_amp, halo_a;
} drift_potential;

/* Per-orbit statistics.  status holds the reference DOPR54 error code in its
 * low byte (dopr54.rs:407-418; always 0 for DOP853) plus flags. */
typedef struct drift_orbit_stats {
    uint32_t accepted;    /* accepted steps */
    uint32_t rejected;    /* rejected step attempts */
    uint32_t dense_steps; /* DOP853: accepted steps that evaluated the dense stages */
    int32_t status;
} drift_orbit_stats;

#define DRIFT_STATUS_CODE_MASK 0xff
#define DRIFT_STATUS_NONFINITE 0x100   /* a non-finite state was written */
#define DRIFT_STATUS_STEP_LIMIT 0x200  /* max_attempts reached; later samples are NaN */

/* Launch tuning knobs for the device-resident entry points. */
typedef struct drift_launch_options {
    int32_t precision;  /* drift_precision */
    int32_t block_size; /* threads per block: 64..256 in steps of 64, or 512 for DOP853
                           with a static potential (dense output packed across each
                           wavefront's lanes: accepted steps hand their dense-output
                           inputs to an LDS stash, and a wavefront runs them once most
                           of its lanes have one; bit-identical results); 0 = library
                           default: 512 for persistent DOP853 launches with direct output
                           and a static potential, else 256 */
    int32_t schedule;   /* drift_schedule */
    uint32_t max_attempts; /* per-orbit attempted-step limit; 0 = default (2^22).
                              The reference has none: a DOP853 orbit whose error
                              estimate never drops below 1 retries forever at the
                              step floor. */
    int32_t ordering;   /* drift_ordering: the order in which lanes take orbits */
    int32_t staging;    /* drift_output_staging: how samples reach the time-major output */
} drift_launch_options;

/* How the integrator writes its samples.  The output is time-major: sample s
 * of orbit i is the 48-byte record at ((s * n) + i) * 6, and a lane writes its
 * orbit's records one at a time as the orbit reaches them.  Lanes that carry
 * neighbouring orbits rarely reach a sample together (the cost ordering and
 * the persistent queue pair unrelated orbits), so DIRECT writes land as
 * scattered partial lines.  STAGED writes each orbit's samples in blocks of
 * 16 consecutive samples to a session scratch buffer (whole lines, written in
 * order) and a tiled transpose then writes the time-major output with full
 * lines; it pays one more pass over the output and wins where samples are
 * dense against steps (low tolerances).  Outputs are bit-identical. */
typedef enum drift_output_staging {
    DRIFT_STAGING_AUTO = 0,   /* the autotuner's choice for TUNED launches, else DIRECT */
    DRIFT_STAGING_DIRECT = 1,
    DRIFT_STAGING_STAGED = 2  /* needs n * ceil(nt / 16) * 16 * 48 bytes of session scratch,
                                 which a later launch may reallocate: refused under stream
                                 capture */
} drift_output_staging;

/* Orbit order of a launch.  Outputs never depend on it (orbits are
 * independent); it only changes how evenly the lanes of a wavefront are
 * loaded.  COST sorts the orbits, on the device, by the estimated orbital
 * frequency sqrt(|a|/r) at t0 (1/8-octave buckets, highest first), so that a
 * wavefront's lanes carry orbits of similar step counts and the expensive ones
 * start first.  It uses per-session scratch buffers: launches of one session
 * that use it are serialised with each other, across streams, by an event. */
typedef enum drift_ordering {
    DRIFT_ORDER_AUTO = 0,  /* COST with n >= 16384 orbits (either schedule), else INDEX */
    DRIFT_ORDER_INDEX = 1, /* orbit i is taken in index order */
    DRIFT_ORDER_COST = 2
} drift_ordering;

typedef enum drift_schedule {
    DRIFT_SCHEDULE_PLAIN = 0,      /* one lane per orbit (plain grid) */
    DRIFT_SCHEDULE_PERSISTENT = 1, /* persistent kernel, lanes refilled from a work queue */
    DRIFT_SCHEDULE_TUNED = 2       /* the options drift_tune_device stored in the session for
                                      this (method, potential kind, precision, log2 n); without
                                      one, the method's default (DOP853 PERSISTENT, DOPR54
                                      PLAIN) with AUTO ordering */
} drift_schedule;

typedef struct drift_session drift_session;

/* ---- library / session -------------------------------------------------- */

int drift_get_abi_version(void);

/* Last error message of the calling thread ("" if none). */
const char* drift_last_error(void);

/* Creates a session bound to HIP device `device` with its own stream
 * (replaces the per-launch CudaContext::new(0), gpu/cuda_oxide.rs:16-18). */
int drift_session_create(int device, drift_session** out);
/* Waits for the device (launches on any stream may use the session's
 * buffers), then frees everything the session holds. */
int drift_session_destroy(drift_session* session);

/* ---- host-buffer launches (GpuKernel::GalpyDopr54 / GalpyDop853 through
 *      backend::launch, src/dispatch/gpu/cuda_oxide.rs:9-86) ------------------ */

/* DOPR54.  Copies state0 and times to the device, integrates every orbit,
 * copies the time-major trajectories back into state_out.  stats (nullable)
 * receives n records.  Returns DRIFT_ERR_INTEGRATION if any orbit's DOPR54
 * error code is non-zero (outputs are still written), mirroring
 * integrate_kepler's Err(code) (dopr54.rs:123). */
int drift_galpy_dopr54(drift_session* session, const double* state0, const double* times,
                       double* state_out, size_t n, size_t nt, double rtol, double atol,
                       double dt_one_init, const drift_potential* potential,
                       drift_orbit_stats* stats);

/* DOP853, dense output at every times[i] (dop853.rs:747-763). */
int drift_galpy_dop853(drift_session* session, const double* state0, const double* times,
                       double* state_out, size_t n, size_t nt, double rtol, double atol,
                       const drift_potential* potential, drift_orbit_stats* stats);

/* ---- device-resident launches (inputs already in HBM) -------------------- */

/* Same contracts with device pointers, enqueued on `stream` (a hipStream_t;
 * NULL = the session's stream).  Asynchronous: no host synchronisation, no
 * error-code scan; d_stats (nullable) receives per-orbit records on the
 * device. */
int drift_galpy_dopr54_device(drift_session* session, const double* d_state0,
                              const double* d_times, double* d_state_out, size_t n, size_t nt,
                              double rtol, double atol, double dt_one_init,
                              const drift_potential* potential, drift_orbit_stats* d_stats,
                              const drift_launch_options* options, void* stream);

/* Stream capture: the two _device entry points may be captured into a
 * hipGraph (hipStreamBeginCapture on `stream`) after one eager launch of the
 * same session with the same potential and schedule (which builds the force
 * table and the work-queue counters; capture allocates nothing).  Under
 * capture DRIFT_ORDER_AUTO means INDEX and DRIFT_ORDER_COST is refused,
 * DRIFT_STAGING_AUTO means DIRECT and DRIFT_STAGING_STAGED is refused (both
 * use session scratch that later launches reuse), and a TUNED launch takes
 * the tuner's schedule and block size with index order and direct output
 * (the same results); each
 * captured persistent launch keeps a work-queue counter of the session for
 * good (up to 56 per session).  A graph must not be replayed concurrently
 * with itself. */
int drift_galpy_dop853_device(drift_session* session, const double* d_state0,
                              const double* d_times, double* d_state_out, size_t n, size_t nt,
                              double rtol, double atol, const drift_potential* potential,
                              drift_orbit_stats* d_stats, const drift_launch_options* options,
                              void* stream);

/* DOP853 on device pointers without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as drift_galpy_dop853_device, on the
 * current HIP device and `stream` (NULL = the null stream), with no session:
 * the work-queue counter, the cost ordering's scratch and the Milky-Way force
 * tables live in one fixed library array per GPU (about 9 MB, held until the
 * process ends).  Limits: cost ordering takes at most 2^20 orbits (AUTO
 * orders 16384..2^20 orbits by cost, others by index; larger COST launches
 * are refused); schedule PLAIN or PERSISTENT; staging AUTO or DIRECT; no
 * moving origins.  The array keeps the last MW potential's tables: a call
 * with another MW potential waits for the device and uploads them, which is
 * refused under stream capture.  Calls that overlap on one GPU (other
 * streams, other threads) share the array and race: serialise them. */
int drift_galpy_dop853_device_noalloc(const double* d_state0, const double* d_times,
                                      double* d_state_out, size_t n, size_t nt, double rtol,
                                      double atol, const drift_potential* potential,
                                      drift_orbit_stats* d_stats,
                                      const drift_launch_options* options, void* stream);

/* ---- launch diagnostics (the reference's debug records of
 *      src/dispatch/gpu.rs:157-178 and gpu/cuda_oxide.rs:15-34, 77-84) -------- */

/* The integrator kernel's grid (blocks) and block (threads) for a launch of
 * n orbits: what drift_galpy_*_device would launch with these options
 * (options NULL: the host-buffer entry points' schedule, persistent for DOP853
 * and plain for DOPR54).  The
 * persistent schedule's grid is the device's resident block count, capped by
 * n.  Allocates nothing; the reference's fixed shape is ceil(n / 128) x 128
 * (src/dispatch/gpu.rs:23-29). */
int drift_launch_shape(drift_session* session, int32_t method, const drift_potential* potential,
                       size_t n, const drift_launch_options* options, int32_t* grid,
                       int32_t* block);

/* Phases of the session's last host-buffer launch (drift_galpy_dopr54 /
 * drift_galpy_dop853), for the reference backend's per-phase records.  The
 * device phases are HIP-event times on the session's stream. */
typedef struct drift_launch_report {
    double setup_s;      /* host time before the first copy: device, buffers, force tables */
    double h2d_s;        /* state0 and times to the device */
    double kernel_s;     /* the integration (order kernels included) */
    double d2h_s;        /* trajectories and statistics back to the host */
    uint64_t h2d_bytes;  /* bytes copied to the device */
    uint64_t d2h_bytes;  /* trajectory bytes copied to the host (statistics excluded) */
    int32_t grid, block; /* the integrator's launch shape */
} drift_launch_report;

/* DRIFT_ERR_INVALID_ARGUMENT if the session has made no host-buffer launch. */
int drift_session_last_launch(drift_session* session, drift_launch_report* report);

/* Autotuner (no counterpart in the reference).  Times every candidate
 * {PLAIN, PERSISTENT} x block {64, 128, 256} x ordering {INDEX, COST} on the
 * launch described by the arguments (PLAIN only while n is below twice the
 * device's resident lanes: a prefix hides the plain schedule's tail), then
 * the fastest with STAGED output, each integrated over the first
 * prefix_samples output times only (0 = max(2, nt / 20)) so that every orbit
 * takes part, and
 * returns the fastest in *best (precision and max_attempts as given in
 * *best on entry, or 0) and its time in *best_ms (nullable).  The choice is
 * stored in the session under (method, potential kind, precision, log2 n) for
 * launches with schedule = DRIFT_SCHEDULE_TUNED.  Needs prefix_samples x n x 48
 * bytes of device scratch; synchronises `stream`. */
int drift_tune_device(drift_session* session, int32_t method, const double* d_state0,
                      const double* d_times, size_t n, size_t nt, double rtol, double atol,
                      double dt_one_init, const drift_potential* potential, size_t prefix_samples,
                      drift_launch_options* best, double* best_ms, void* stream);

/* ---- ragged batches -------------------------------------------------------- */

/* One job of a ragged batch: its own orbits, method, potential, time grid and
 * tolerances (device pointers). */
typedef struct drift_job {
    int32_t method;            /* drift_method */
    int32_t potential;         /* index into the potentials array */
    size_t n;                  /* orbits (0 allowed) */
    size_t nt;                 /* output times, >= 2 */
    double rtol, atol;         /* natural logarithms */
    double dt_one_init;        /* DOPR54 initial sub-step; -9999.99 = RK4 estimate */
    double span_hint;          /* |t[nt-1] - t[0]| for scheduling (0 = unknown) */
    const double* d_state0;    /* (n, 6) */
    const double* d_times;     /* (nt,) */
    double* d_state_out;       /* (nt, n, 6) */
    drift_orbit_stats* d_stats;/* (n,) or NULL */
} drift_job;

/* Integrates every job to completion: one persistent launch per method, lanes
 * refilled from a work queue across all jobs of that method (jobs grouped by
 * potential kind, costliest first; the orbits of a job that alone fills the
 * device's lanes in cost order, those of smaller jobs in index order; DOPR54
 * step estimates in a kernel of their own before the integrator).  Device
 * scratch of the session: a few words per work item.  options->schedule is
 * ignored.  Returns after the work is complete (it synchronises `stream`). */
int drift_ragged_device(drift_session* session, const drift_job* jobs, size_t njobs,
                        const drift_potential* potentials, size_t npotentials,
                        const drift_launch_options* options, void* stream);

/* Autotuner for drift_ragged_device: integrates every job over its first
 * prefix_samples outputs (0 = 5% of its grid, at least 2) with each block size
 * in {64, 128, 256}, one method at a time, twice each (the integrator kernel's
 * faster time counts), and stores the fastest per (precision, method) in the
 * session, 128 or 64 only if 3% faster than 256; ragged launches of the session whose
 * options->block_size is 0 use it (256 while untuned).  best_blocks (nullable)
 * receives [DOPR54, DOP853] (0 = no job of that method).  Needs the prefixes'
 * outputs as session scratch and synchronises `stream`.  No counterpart in the
 * reference, which launches each container with a fixed 128-thread block. */
int drift_tune_ragged(drift_session* session, const drift_job* jobs, size_t njobs,
                      const drift_potential* potentials, size_t npotentials,
                      const drift_launch_options* options, size_t prefix_samples,
                      int32_t* best_blocks, void* stream);

/* ---- serving loop (continuous batching) ---------------------------------- */

/* A request: one job submitted from host memory.  No counterpart in the
 * reference; it wraps drift_ragged_device's job with host buffers. */
typedef struct drift_request {
    int32_t method;               /* drift_method */
    int32_t reserved;             /* must be 0 */
    drift_potential potential;
    size_t n, nt;
    double rtol, atol;            /* natural logarithms */
    double dt_one_init;           /* DOPR54; -9999.99 = RK4 estimate */
    const double* state0;         /* host (n, 6), copied at submission */
    const double* times;          /* host (nt,), copied at submission */
    double* state_out;            /* host (nt, n, 6); must stay valid until completion */
    drift_orbit_stats* stats;     /* host (n,) or NULL */
} drift_request;

typedef struct drift_server_options {
    uint64_t hbm_budget_bytes;    /* device bytes for admitted requests; 0 = half the free HBM */
    int32_t chunk_samples;        /* output samples per orbit per chunk launch; 0 = adaptive:
                                     starts at 8 and is doubled or halved after each chunk to
                                     keep a chunk's duration near chunk_target_ms, in [1, 512] */
    int32_t precision;            /* drift_precision for every request */
    int32_t block_size;           /* 0 = 256 */
    uint32_t max_attempts;        /* per-orbit attempt limit; 0 = default */
    double chunk_target_ms;       /* adaptive chunks only; 0 = 10 ms.  A queued request waits
                                     at most about one chunk before it joins the GPU work */
} drift_server_options;

typedef enum drift_request_state {
    DRIFT_REQUEST_QUEUED = 0,
    DRIFT_REQUEST_RUNNING = 1,
    DRIFT_REQUEST_DONE = 2,
    DRIFT_REQUEST_FAILED = 3
} drift_request_state;

typedef struct drift_request_info {
    int32_t state;                /* drift_request_state */
    int32_t status;               /* drift_status of a finished request */
    double submit_s, admit_s, complete_s;  /* seconds since server creation */
    uint64_t accepted_steps;      /* summed over the request's orbits */
    uint64_t chunks;              /* chunk launches the request took part in */
} drift_request_info;

typedef struct drift_server drift_server;

/* Starts a serving loop on `device` (its own session, stream and thread).
 * Requests are admitted first-come first-served while their device footprint
 * (inputs, (nt, n, 6) output, resident integrator state) fits the HBM budget;
 * every chunk launch advances all admitted requests by up to chunk_samples
 * output samples in one persistent ragged launch per method, their orbit
 * state resident in HBM between chunks.  A request larger than the budget is
 * admitted alone. */
int drift_server_create(int device, const drift_server_options* options, drift_server** out);
int drift_server_submit(drift_server* server, const drift_request* request, uint64_t* ticket);
int drift_server_query(drift_server* server, uint64_t ticket, drift_request_info* info);
/* Blocks until the request finishes or timeout_s elapses (timeout_s < 0:
 * forever).  Returns the request's drift_status, or DRIFT_ERR_TIMEOUT. */
int drift_server_wait(drift_server* server, uint64_t ticket, double timeout_s,
                      drift_request_info* info);
/* Chunk launches so far.  A chunk's stream work (per method: descriptor
 * upload, queue reset, ragged kernel; then the finished-orbit counters back to
 * host memory) is captured into a hipGraph once per chunk shape (the admitted
 * set, the kernels' grids and buffers) and replayed while the shape holds;
 * graph_launches counts every launch of a captured graph, graph_captures the
 * captures, eager_chunks the chunks launched without a graph (only after a
 * capture failed).  Any pointer may be NULL.  No counterpart in the
 * reference, which has no serving loop. */
int drift_server_stats(drift_server* server, uint64_t* graph_launches, uint64_t* graph_captures,
                       uint64_t* eager_chunks);
/* Where the server's time went so far: busy_s, the summed chunk launches
 * (from the first enqueue to the chunk stream's synchronisation: the GPU's
 * share of the wall time); copy_s and copy_bytes, the finished requests'
 * trajectories copied to host memory (their rate is the host link's).  Any
 * pointer may be NULL.  No counterpart in the reference. */
int drift_server_busy(drift_server* server, double* busy_s, double* copy_s, uint64_t* copy_bytes);
/* Finishes every submitted request, stops the loop and frees the server. */
int drift_server_destroy(drift_server* server);

/* ---- moving origins (CustomOrigin, shared/src/potential.rs:264-356) ------ */

/* Uploads a set of moving origins to the session and returns its id (> 0) in
 * *origins_id, for drift_potential.origins.  Origin i is a point mass
 * (params[2i + 1] = b = 0, KeplerPotential) or a Plummer sphere (b > 0) of
 * amplitude params[2i] whose centre follows
 *   o_i(t) = quintic_s(tau),  u = (t - t_start) n_segments / (t_end - t_start),
 *   s = floor(u) clamped to [0, n_segments - 1],  tau = u - s,
 * with tracks[((i * n_segments) + s) * 18 + 6 c + k] the tau^(5 - k)
 * coefficient of coordinate c (x, y, z), the order of CustomOrigin's
 * QuinticCoeff a..f (potential.rs:273-289).  The force at x is
 * sum_i F_i(x - o_i(t)), added to the static force.  Host arrays, copied
 * (synchronously); the set lives until drift_origins_destroy or the end of
 * the session. */
int drift_origins_create(drift_session* session, const double* tracks, const double* params,
                         int32_t n_origins, int32_t n_segments, double t_start, double t_end,
                         int32_t* origins_id);
/* Frees a set.  Waits for the device first, so launches already enqueued
 * that use it (on any stream) finish; later launches must not name it. */
int drift_origins_destroy(drift_session* session, int32_t origins_id);

/* Host-only: quintic Hermite tracks of n orbits sampled at nt strictly
 * monotone, uniformly spaced times, matching position, velocity and
 * acceleration at every sample.  states: nt x n x 6 (time-major, as the
 * integrators write it), accels: nt x n x 3.  Writes n x (nt - 1) x 18
 * coefficients in drift_origins_create's layout, segments in increasing time,
 * and the span [*t_start, *t_end].  No device needed. */
int drift_origin_tracks_hermite(const double* states, const double* accels, const double* times,
                                size_t nt, size_t n, double* tracks, double* t_start,
                                double* t_end);

/* ---- force evaluation (test and diagnostics entry) ----------------------- */

/* Accelerations of the device force formulation at n points.
 * d_pos, d_acc: n x 3 doubles on the device; d_t: n times (device), or NULL
 * for t = 0 (only moving origins depend on time). */
int drift_force_device(drift_session* session, const drift_potential* potential,
                       const double* d_pos, const double* d_t, double* d_acc, size_t n,
                       int32_t precision, void* stream);

/* The same in FP64 on host buffers (copied through the session's stream,
 * which is synchronised): pos, acc n x 3, t n times or NULL for t = 0. */
int drift_force(drift_session* session, const drift_potential* potential, const double* pos,
                const double* t, double* acc, size_t n);

/* Host-only check of the Milky-Way spherical force tables: worst relative
 * error of the tabulated S(r^2) against its long-double evaluation over
 * samples_per_segment points per table segment, over both the LDS table
 * (r^2 in [2^-12, 2^12)) and the far table (every positive finite r^2, where S
 * is a normal double).  Also checks that the two tables hold the same
 * coefficients, bit for bit, on the segments they share (a kernel lane reads
 * the far table outside the LDS range); DRIFT_ERR_INVALID_ARGUMENT if not.
 * No device needed. */
int drift_mw_table_check(const drift_potential* potential, int32_t samples_per_segment,
                         double* max_rel_error);

#ifdef __cplusplus
}
#endif

#endif /* DRIFT_HIP_H */
