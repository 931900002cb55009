// drift_hip.hip -- gfx950 kernels and the C ABI of libdrift_hip.so.
//
// Replaces the reference's CUDA device crate (kernels/src/lib.rs) and its
// cuda-oxide / rust-cuda backends (src/dispatch/gpu/*.rs).  Layout in HBM:
//   state0  (n, 6)      AoS rows, read once per orbit (3 x 16-byte loads)
//   times   (nt,)       shared by every orbit; DOP853 blocks copy it to LDS
//                       when nt <= 1,024
//   out     (nt, n, 6)  time-major AoS; lane i writes its 48-byte record of
//                       sample s at ((s * n) + i) * 6 as 3 x 16-byte stores
//   stats   (n,)        drift_orbit_stats, written once per orbit
//   table   MW spherical-force coefficients, staged per block in LDS
// Per-orbit phase-space state, stage derivatives and dense-output
// coefficients live in registers for the whole integration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "drift_internal.hpp"
#include "forces.hpp"
#include "integrators.hpp"

namespace drift {

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------

// Output staging (DRIFT_STAGING_STAGED): samples in blocks of kStageBlock
// consecutive samples per orbit, block b of orbit i at
// ((b * n) + i) * kStageBlock records, so a lane writes its orbit's records in
// order and whole lines; unstage_kernel then writes the time-major output.
constexpr int kStageShift = 4;
constexpr int kStageBlock = 1 << kStageShift;

__device__ __forceinline__ int64_t out_record(int64_t s, int64_t n, int64_t idx, bool staged) {
    if (!staged) return s * n + idx;
    return (((s >> kStageShift) * n + idx) << kStageShift) + (s & (kStageBlock - 1));
}

struct OutSink {
    double* out;
    int64_t n;
    int64_t idx;
    bool staged;
    __device__ __forceinline__ void emit(int32_t s, const double* y) const {
        double2* p = reinterpret_cast<double2*>(out + out_record(s, n, idx, staged) * DIM);
        p[0] = make_double2(y[0], y[1]);
        p[1] = make_double2(y[2], y[3]);
        p[2] = make_double2(y[4], y[5]);
    }
    __device__ __forceinline__ void emit_component(int32_t s, int c, double v) const {
        out[out_record(s, n, idx, staged) * DIM + c] = v;
    }
};

struct PlainArgs {
    const double* state0;
    const double* times;
    double* out;
    drift_orbit_stats* stats;
    int64_t n;
    int32_t nt;
    uint32_t max_attempts;
    double rtol, atol;    // linear tolerances
    double dt_one_init;   // DOPR54 only
    double bound2[6];     // DOPR54 only
    PotParams pot;
    const void* table;    // MW: TABLE_DOUBLES doubles, or TABLE32_FLOATS floats (mixed)
    unsigned long long* queue;  // persistent schedule: next unclaimed orbit slot
    const uint32_t* order;      // slot -> orbit (NULL = identity)
    int32_t staged;             // out is the orbit-blocked staging buffer
};

template <int KIND, class FT>
struct ForceFor;

template <class FT>
struct ForceFor<POT_GALPY_KEPLER, FT> {
    using type = GalpyKeplerForce<FT>;
    __device__ static type make(const PotParams&, const FT*) { return type{}; }
};
template <class FT>
struct ForceFor<POT_KEPLER, FT> {
    using type = KeplerForce<FT>;
    __device__ static type make(const PotParams& p, const FT*) { return type{FT(p.amp)}; }
};
template <class FT>
struct ForceFor<POT_PLUMMER, FT> {
    using type = PlummerForce<FT>;
    __device__ static type make(const PotParams& p, const FT*) { return type{FT(p.amp), FT(p.b2)}; }
};
template <class FT>
struct ForceFor<POT_MW, FT> {
    using type = MwForce<FT>;
    __device__ static type make(const PotParams& p, const FT* table) {
        return type{&p, table, FT(p.disk_amp), FT(p.disk_a), FT(p.disk_b2)};
    }
};

template <class FT>
struct ForceFor<POT_NONE, FT> {
    using type = NoForce;
    __device__ static type make(const PotParams&, const FT*) { return type{}; }
};

// The force of a kernel kind: the static force, or with POT_MOVING set the
// static force plus the moving origins.
template <int KIND, class FT, bool MOVING = (KIND & POT_MOVING) != 0>
struct KindForce : ForceFor<KIND, FT> {};

template <int KIND, class FT>
struct KindForce<KIND, FT, true> {
    using Static = ForceFor<static_kind(KIND), FT>;
    using type = MovingForce<typename Static::type>;
    __device__ static type make(const PotParams& p, const FT* table) {
        return type{Static::make(p, table), p.tracks, p.origin_params, p.n_origins, p.n_segments,
                    p.t_start, p.inv_dt};
    }
};

// The force of a time-windowed moving-origin launch (rows in LDS).
template <int KIND>
struct WindowForce {
    using Static = ForceFor<static_kind(KIND), double>;
    using type = MovingForce<typename Static::type>;
    __device__ static type make(const PotParams& p, const double* table, const double* rows,
                                int32_t win_lo, int32_t win_n) {
        type f{Static::make(p, table), p.tracks, p.origin_params, p.n_origins, p.n_segments,
               p.t_start, p.inv_dt};
        f.win_rows = rows;
        f.win_lo = win_lo;
        f.win_n = win_n;
        return f;
    }
};

// Stages the MW coefficient table into LDS (no-op for the other potentials).
template <int KIND, class FT>
__device__ __forceinline__ const FT* stage_table(const void* gtable, FT* lds) {
    if constexpr (kind_has_table(KIND)) {
        const FT* g = static_cast<const FT*>(gtable);
        constexpr int words = sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS;
        for (int k = threadIdx.x; k < words; k += blockDim.x) lds[k] = g[k];
        __syncthreads();
        return lds;
    } else {
        return nullptr;
    }
}

__device__ __forceinline__ void load_state(const double* state0, int64_t i, double* y) {
    const double2* p = reinterpret_cast<const double2*>(state0 + i * DIM);
    double2 a = p[0], b = p[1], c = p[2];
    y[0] = a.x; y[1] = a.y; y[2] = b.x; y[3] = b.y; y[4] = c.x; y[5] = c.y;
}

constexpr int kDefaultBlock = 256;

// Output sample times in LDS.  A lane reads the time of its next sample once
// per sample it writes.  A lane's global loads and stores complete in order
// (one counter counts both), so waiting for a global load of a sample time
// also waits for the stores of the samples before it, scattered 48-byte
// records under load; an LDS load waits for neither.  DOP853 launches with
// at most kLdsTimes outputs therefore run kernels (LT = true) that copy the
// time grid into LDS and read it there (8 KB per block): 1-2% faster at
// 1,001 outputs on 1M Milky-Way orbits (rtol 1e-7 with staged output, and the
// 1e-12 default), same bits.  Longer grids run the kernels that read it from
// global memory (one kernel with an LDS path and a global fallback ran 3-5%
// slower at 2,001 outputs: the fallback's loads needed waits of their own).
constexpr int kLdsTimes = 1024;

__host__ __device__ constexpr bool lds_times(int64_t nt) { return nt <= kLdsTimes; }

template <bool LT>
__device__ __forceinline__ const double* stage_times(const double* times, int32_t nt, double* l) {
    if constexpr (LT) {
        for (int i = threadIdx.x; i < nt && i < kLdsTimes; i += blockDim.x) l[i] = times[i];
        __syncthreads();
        return l;
    } else {
        return times;
    }
}

// Two wavefronts per SIMD for the DOP853 kernels: with one, the latency of
// the force's LDS table loads and FP64 dependency chains is exposed; forcing
// two costs a small register spill and measured ~10-30% faster on gfx950.
#define DRIFT_TWO_WAVES __attribute__((amdgpu_waves_per_eu(2, 2)))

// The plain DOP853 kernel also exists at three waves per SIMD (WAVES = 3) for
// the FP64 Milky-Way force: 168 VGPRs with 9 spilled, three 256-thread
// blocks' 30 KB tables per CU.  It is launched when the grid fills the
// device's block slots at three waves (3 per CU) at least twice
// (wide_plain): on 256 CUs, from 393,216 orbits, where the plain launch ran
// 3-4.5% faster than at two waves (524,288 to 1,048,576 orbits); below, the
// blocks' last round left slots empty and two waves were faster (65,536
// orbits: 3%, 262,144: 7%).
static bool wide_plain(size_t n, int cu_count) {
    return n >= (size_t)2 * 3 * (size_t)std::max(cu_count, 1) * kDefaultBlock;
}

template <int KIND, class FT, int WAVES = 2, bool LT = false>
__global__ void __launch_bounds__(kDefaultBlock) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES)))
dop853_plain_kernel(PlainArgs a) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    __shared__ double ltimes[LT ? kLdsTimes : 1];
    const double* times = stage_times<LT>(a.times, a.nt, ltimes);
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.n) return;
    const int64_t i = a.order ? (int64_t)a.order[slot] : slot;
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    const double t0 = a.times[0];
    const Dop853Job J = make_dop853_job(times, a.nt, a.rtol, a.atol, a.max_attempts, t0,
                                        a.times[a.nt - 1]);
    OutSink sink{a.out, a.n, i, a.staged != 0};
    double y0[DIM];
    load_state(a.state0, i, y0);
    Dop853Lane L;
    dop853_init(L, y0, t0, J, f, sink);
    while (!dop853_attempt(L, J, f, sink)) {
    }
    if (a.stats) {
        drift_orbit_stats st;
        st.accepted = L.accepted;
        st.rejected = L.rejected;
        st.dense_steps = L.dense;
        st.status = L.status;
        a.stats[i] = st;
    }
}

// Waves per SIMD of the DOPR54 kernels.  Without a table the step loop needs
// under 104 registers (the step estimate runs out of line), so four waves fit
// and hide the latency of the short force evaluations (a point-mass launch
// runs 14-24% faster than at three).  With the Milky-Way table the kernels
// need 154-160 VGPRs: the float table's kernel runs three waves, the FP64
// one two (at three it ran 262,144 orbits 10% slower).  The moving origins'
// registers allow two.
template <int KIND, class FT>
constexpr int dopr54_waves() {
    return (KIND & POT_MOVING) ? 2 : kind_has_table(KIND) ? (sizeof(FT) == 8 ? 2 : 3) : 4;
}

template <int KIND, class FT>
__global__ void __launch_bounds__(kDefaultBlock)
__attribute__((amdgpu_waves_per_eu(dopr54_waves<KIND, FT>(), dopr54_waves<KIND, FT>())))
dopr54_plain_kernel(PlainArgs a) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.n) return;
    const int64_t i = a.order ? (int64_t)a.order[slot] : slot;
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    Dopr54Job J;
    J.times = a.times;
    J.nt = a.nt;
    J.max_attempts = a.max_attempts;
    J.dt = a.times[1] - a.times[0];  // dopr54.rs:713
    J.rtol = a.rtol;
    J.atol = a.atol;
    J.dt_one_init = a.dt_one_init;
#pragma unroll
    for (int k = 0; k < 6; k++) J.bound2[k] = a.bound2[k];
    OutSink sink{a.out, a.n, i, a.staged != 0};
    double y0[DIM];
    load_state(a.state0, i, y0);
    Dopr54Lane L;
    dopr54_init(L, y0, a.times[0], J, f, sink);
    bool done = L.ii >= J.nt - 1;
    while (!done) done = dopr54_attempt(L, J, f, sink);
    if (a.stats) {
        drift_orbit_stats st;
        st.accepted = L.accepted;
        st.rejected = L.rejected;
        st.dense_steps = 0;
        st.status = L.err | L.status;
        a.stats[i] = st;
    }
}

// Persistent schedule: a grid sized to the machine; each lane integrates
// orbits one after another, claiming the next slot from a global queue when
// its orbit finishes.  Requests are aggregated per wavefront with a ballot so
// one atomic serves every lane that needs work, and lanes stay busy until the
// queue drains instead of idling behind the wave's slowest orbit.
__device__ __forceinline__ bool claim_orbit(bool need, unsigned long long* queue, int64_t n,
                                            const uint32_t* order, int64_t& orbit) {
    const unsigned long long want = __ballot(need);
    if (want == 0) return true;
    const int lane = __lane_id();
    const int leader = __ffsll(want) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(queue, (unsigned long long)__popcll(want));
    base = __shfl(base, leader);
    if (!need) return true;
    const unsigned long long below = lane == 0 ? 0ull : (want & (~0ull >> (64 - lane)));
    const unsigned long long slot = base + (unsigned long long)__popcll(below);
    if (slot >= (unsigned long long)n) return false;  // queue drained: lane retires
    if (order) {
        orbit = (int64_t)order[slot];
        // Wait for the order entry here, where it is loaded.  A load that
        // might still be in flight at the caller's loop back edge makes the
        // compiler wait at the top of every iteration. One counter covers loads
        // and stores, so that wait would also drain the wavefront's
        // outstanding sample stores.
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    } else {
        orbit = (int64_t)slot;
    }
    return true;
}

template <int KIND, class FT, bool LT = false>
__global__ void __launch_bounds__(kDefaultBlock) DRIFT_TWO_WAVES dop853_persistent_kernel(PlainArgs a) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    __shared__ double ltimes[LT ? kLdsTimes : 1];
    const double* times = stage_times<LT>(a.times, a.nt, ltimes);
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    const double t0 = a.times[0];
    const Dop853Job J = make_dop853_job(times, a.nt, a.rtol, a.atol, a.max_attempts, t0,
                                        a.times[a.nt - 1]);
    OutSink sink{a.out, a.n, 0, a.staged != 0};
    Dop853Lane L;
    bool need = true;
    for (;;) {
        int64_t orbit = -1;
        if (!claim_orbit(need, a.queue, a.n, a.order, orbit)) break;
        if (need) {
            sink.idx = orbit;
            double y0[DIM];
            load_state(a.state0, orbit, y0);
            dop853_init(L, y0, t0, J, f, sink);
            need = false;
        }
        if (dop853_attempt(L, J, f, sink)) {
            if (a.stats) {
                drift_orbit_stats st;
                st.accepted = L.accepted;
                st.rejected = L.rejected;
                st.dense_steps = L.dense;
                st.status = L.status;
                a.stats[sink.idx] = st;
            }
            need = true;
        }
    }
}

template <int KIND, class FT>
__global__ void __launch_bounds__(kDefaultBlock)
__attribute__((amdgpu_waves_per_eu(dopr54_waves<KIND, FT>(), dopr54_waves<KIND, FT>())))
dopr54_persistent_kernel(PlainArgs a) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    Dopr54Job J;
    J.times = a.times;
    J.nt = a.nt;
    J.max_attempts = a.max_attempts;
    J.dt = a.times[1] - a.times[0];
    J.rtol = a.rtol;
    J.atol = a.atol;
    J.dt_one_init = a.dt_one_init;
#pragma unroll
    for (int k = 0; k < 6; k++) J.bound2[k] = a.bound2[k];
    OutSink sink{a.out, a.n, 0, a.staged != 0};
    Dopr54Lane L;
    bool need = true;
    for (;;) {
        int64_t orbit = -1;
        if (!claim_orbit(need, a.queue, a.n, a.order, orbit)) break;
        bool done = false;
        if (need) {
            sink.idx = orbit;
            double y0[DIM];
            load_state(a.state0, orbit, y0);
            dopr54_init(L, y0, a.times[0], J, f, sink);
            need = false;
            done = L.ii >= J.nt - 1;
        }
        if (!done) done = dopr54_attempt(L, J, f, sink);
        if (done) {
            if (a.stats) {
                drift_orbit_stats st;
                st.accepted = L.accepted;
                st.rejected = L.rejected;
                st.dense_steps = 0;
                st.status = L.err | L.status;
                a.stats[sink.idx] = st;
            }
            need = true;
        }
    }
}

// ---------------------------------------------------------------------------
// Time-windowed launches for moving origins (DOP853, persistent schedule,
// two or more origins).  Lanes on unrelated orbits sit at unrelated times, so
// every origin after the first gathers its 144-byte track row on every force
// evaluation from 64 different lines per load instruction.  A windowed launch
// integrates all orbits up to one output sample, then all orbits up to a later
// one, and so on: one kernel per window, each orbit paused after it emits the
// window's last sample with its state saved (kResumeWords words, SoA) and
// resumed by the next window exactly where it stopped, so outputs and
// statistics are bit-identical to one launch.  Within a window the stage
// times lie between the window's first and last sample times (plus one step),
// and the kernel stages those segments' rows of every origin in LDS; a lane
// whose time falls outside them reads the tracks as before.
// ---------------------------------------------------------------------------
constexpr int kResumeWords = 15;
constexpr int kWindowRowsMax = 96;  // LDS rows a window stages (13.5 KB)

struct WindowArgs {
    double* state;    // kResumeWords x n, word k of orbit i at k * n + i
    int32_t first;    // 1: the first window, orbits start from state0
    int32_t stop;     // pause an orbit once it has emitted sample `stop`
    int32_t s_begin;  // the previous window's last sample (0 for the first)
};

__device__ __forceinline__ double pack_words(int32_t a, int32_t b) {
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)a << 32) | (uint32_t)b));
}
__device__ __forceinline__ void unpack_words(double w, int32_t& a, int32_t& b) {
    const uint64_t u = (uint64_t)__double_as_longlong(w);
    a = (int32_t)(u >> 32);
    b = (int32_t)(uint32_t)u;
}

// word 14: 1 while the orbit runs, 2 once it has finished
__device__ __forceinline__ void resume_save(double* st, int64_t n, int64_t i, const Dop853Lane& L,
                                            int32_t phase) {
#pragma unroll
    for (int k = 0; k < DIM; k++) st[k * n + i] = L.y[k];
#pragma unroll
    for (int k = 0; k < 3; k++) st[(6 + k) * n + i] = L.F1[k];
    st[9 * n + i] = L.h;
    st[10 * n + i] = L.t_cur;
    st[11 * n + i] = pack_words(L.ii, L.reject);
    st[12 * n + i] = pack_words((int32_t)L.accepted, (int32_t)L.rejected);
    st[13 * n + i] = pack_words((int32_t)L.dense, L.status);
    st[14 * n + i] = pack_words(phase, 0);
}

__device__ __forceinline__ int32_t resume_load(const double* st, int64_t n, int64_t i,
                                               Dop853Lane& L) {
    int32_t phase, pad;
    unpack_words(st[14 * n + i], phase, pad);
    if (phase != 1) return phase;
#pragma unroll
    for (int k = 0; k < DIM; k++) L.y[k] = st[k * n + i];
#pragma unroll
    for (int k = 0; k < 3; k++) L.F1[k] = st[(6 + k) * n + i];
    L.h = st[9 * n + i];
    L.t_cur = st[10 * n + i];
    int32_t a, b;
    unpack_words(st[11 * n + i], L.ii, L.reject);
    unpack_words(st[12 * n + i], a, b);
    L.accepted = (uint32_t)a;
    L.rejected = (uint32_t)b;
    unpack_words(st[13 * n + i], a, L.status);
    L.dense = (uint32_t)a;
    return phase;
}

// The track segment of time t, as MovingForce finds it.
__device__ __forceinline__ int64_t track_segment(const PotParams& p, double t) {
    double sf = std::floor((t - p.t_start) * p.inv_dt);
    if (!(sf >= 0.0)) sf = 0.0;
    const double last = (double)(p.n_segments - 1);
    return (int64_t)(sf > last ? last : sf);
}

template <int KIND>
__global__ void __launch_bounds__(kDefaultBlock) DRIFT_TWO_WAVES
dop853_window_kernel(PlainArgs a, WindowArgs w) {
    __shared__ double lds[kind_has_table(KIND) ? TABLE_DOUBLES : 1];
    __shared__ double rows[kWindowRowsMax * ORIGIN_COEFS];
    const double* table = stage_table<KIND, double>(a.table, lds);
    // the segments from the window's first sample time to one past its last
    // (a step may overshoot the last sample), as many as fit
    const int64_t seg_lo = track_segment(a.pot, a.times[w.s_begin]);
    const int64_t seg_hi = track_segment(a.pot, a.times[w.stop]) + 1;
    const int32_t fit = kWindowRowsMax / max(a.pot.n_origins, 1);
    const int32_t win_n = (int32_t)max((int64_t)0, min((int64_t)fit, min(seg_hi, (int64_t)a.pot.n_segments - 1) - seg_lo + 1));
    for (int k = threadIdx.x; k < a.pot.n_origins * win_n * ORIGIN_COEFS; k += blockDim.x) {
        const int r = k / ORIGIN_COEFS, c = k % ORIGIN_COEFS;
        const int i = r / win_n, s = r % win_n;
        rows[k] = a.pot.tracks[((int64_t)i * a.pot.n_segments + seg_lo + s) * ORIGIN_COEFS + c];
    }
    __syncthreads();
    const auto f = WindowForce<KIND>::make(a.pot, table, rows, (int32_t)seg_lo, win_n);
    const double t0 = a.times[0];
    const Dop853Job J = make_dop853_job(a.times, a.nt, a.rtol, a.atol, a.max_attempts, t0,
                                        a.times[a.nt - 1]);
    OutSink sink{a.out, a.n, 0, false};
    Dop853Lane L;
    bool need = true;
    for (;;) {
        int64_t orbit = -1;
        if (!claim_orbit(need, a.queue, a.n, a.order, orbit)) break;
        if (need) {
            sink.idx = orbit;
            need = false;
            if (w.first) {
                double y0[DIM];
                load_state(a.state0, orbit, y0);
                dop853_init(L, y0, t0, J, f, sink);
            } else if (resume_load(w.state, a.n, orbit, L) != 1) {
                need = true;  // finished in an earlier window
                continue;
            } else {
                dop853_refresh_next(L, J);
            }
            if (L.ii >= w.stop && L.ii < J.nt - 1) {
                resume_save(w.state, a.n, orbit, L, 1);
                need = true;
                continue;
            }
        }
        if (dop853_attempt(L, J, f, sink)) {
            if (a.stats) {
                drift_orbit_stats st;
                st.accepted = L.accepted;
                st.rejected = L.rejected;
                st.dense_steps = L.dense;
                st.status = L.status;
                a.stats[sink.idx] = st;
            }
            resume_save(w.state, a.n, sink.idx, L, 2);
            need = true;
        } else if (L.ii >= w.stop) {
            resume_save(w.state, a.n, sink.idx, L, 1);
            need = true;
        }
    }
}

// ---------------------------------------------------------------------------
// Packed dense output (DOP853 in 512-thread blocks).  About half of a
// wavefront's lanes need dense output on a given attempted step (an accepted
// step with a sample in it), and a wavefront runs that branch, three force
// evaluations and the rcont1..8 sums, for all 64 lanes whenever one needs it.
// Here a lane hands its step's dense-output inputs (Dop853Dense, 43 words) to
// a stash in LDS that belongs to its wavefront, and the wavefront runs dense
// output only once enough requests are pending to fill most of its lanes: the
// lanes with a request from this step run their own, the others take requests
// from the stash.  Samples are written by whichever lane runs the request, to
// the request's orbit, with the same arithmetic, so outputs are bit-identical
// to the inline branch.  An orbit may finish with requests still stashed;
// its statistics record is written then, its status OR-ed in (the launch
// zeroes the status words first), and a lane that runs one of its requests
// and writes a non-finite sample ORs the flag in as well.  A wavefront whose
// lanes have all run out of orbits runs what is left in its stash before it
// leaves.  One 512-thread block per CU shares one copy of the Milky-Way
// table; the rest of the LDS holds the eight stashes.
// ---------------------------------------------------------------------------
constexpr int kPackedBlock = 512;
constexpr int kPackedWaves = kPackedBlock / 64;
constexpr int kDenseWords = kDenseRecordWords;  // words of a stashed Dop853Dense
constexpr int kDenseRunAt = 48;  // pending requests from which a wavefront runs them
constexpr int kLdsBytes = 163840;

template <int KIND, class FT>
constexpr int table_words() {
    return kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1;
}

// Stash slots per wavefront: what the LDS holds beside the table (and the
// time grid's copy, LT), at most 56.
template <int KIND, class FT, bool LT = false>
constexpr int packed_slots() {
    constexpr int free_bytes = kLdsBytes - table_words<KIND, FT>() * (int)sizeof(FT) -
                               (LT ? 8 * kLdsTimes : 0) - 256;
    constexpr int d = free_bytes / (kPackedWaves * kDenseWords * 8);
    return d < 56 ? d : 56;
}

// LDS accesses of one wavefront's lanes to each other's words need no wait:
// the LDS serves a wavefront's instructions in order.  Only the compiler
// must keep them in program order across this point (a memory fence here
// would also wait for every outstanding global store of the wavefront).
__device__ __forceinline__ void wave_lds_sync() { __asm__ __volatile__("" ::: "memory"); }

// A stash is field-major in 16-byte words: words 2j and 2j + 1 of slot s at
// double2 index j * D + s, so the lanes of one access touch consecutive
// 16-byte words (the record's layout: dense_pack).
constexpr int kDensePairs = kDenseWords / 2;

template <int D>
__device__ __forceinline__ void stash_put(double2* W, int s, const Dop853Dense& r) {
    DenseWords x;
    dense_pack(r, x);
#pragma unroll
    for (int j = 0; j < kDensePairs; j++) W[j * D + s] = make_double2(x.w[2 * j], x.w[2 * j + 1]);
}

template <int D>
__device__ __forceinline__ void stash_get(const double2* W, int s, Dop853Dense& r) {
    DenseWords x;
#pragma unroll
    for (int j = 0; j < kDensePairs; j++) {
        const double2 v = W[j * D + s];
        x.w[2 * j] = v.x;
        x.w[2 * j + 1] = v.y;
    }
    dense_unpack(x, r);
}

// One wavefront's dense output after an attempted step.  own: this lane's
// step left request R.  p: the wavefront's pending requests (slots 0..p-1 of
// its stash W).  Unless force is set, the requests are stashed while they fit
// and fall short of kDenseRunAt.  Otherwise one pass runs: the lanes with a
// request of their own run it, the other lanes take stashed ones from the top
// (so a pass with force set and no own requests runs up to 64 stashed ones).
// Called by every lane of the wavefront.
template <int D, class F>
__device__ __forceinline__ void dense_round(double2* W, int& p, bool own, Dop853Dense& R,
                                            bool force, const Dop853Job& J, const F& f,
                                            OutSink sink, drift_orbit_stats* stats) {
    const uint64_t ownm = __ballot(own);
    const int k = __popcll(ownm);
    if (k + p == 0) return;
    const uint64_t below = (1ull << __lane_id()) - 1ull;
    if (!force && k + p < kDenseRunAt && k + p <= D) {
        if (own) stash_put<D>(W, p + __popcll(ownm & below), R);
        p += k;
        wave_lds_sync();
        return;
    }
    const uint64_t free_lanes = __ballot(true) & ~ownm;
    const int take = min(p, __popcll(free_lanes));
    bool popped = false;
    if (!own) {
        const int r = __popcll(free_lanes & below);
        if (r < take) {
            stash_get<D>(W, p - 1 - r, R);
            popped = true;
        }
    }
    p -= take;
    wave_lds_sync();
    if (own || popped) {
        sink.idx = R.idx;
        const int32_t st = dop853_run_dense(R, J, f, sink);
        if (st && stats) atomicOr(&stats[R.idx].status, st);
    }
}

// An orbit's statistics record: the counts, and its status OR-ed into the
// zeroed status word (a lane running one of the orbit's stashed requests may
// OR in a non-finite flag before or after).
__device__ __forceinline__ void write_packed_stats(drift_orbit_stats* stats, int64_t i,
                                                   const Dop853Lane& L) {
    if (!stats) return;
    stats[i].accepted = L.accepted;
    stats[i].rejected = L.rejected;
    stats[i].dense_steps = L.dense;
    if (L.status) atomicOr(&stats[i].status, L.status);
}

template <int KIND, class FT>
__global__ void __launch_bounds__(kPackedBlock) DRIFT_TWO_WAVES
dop853_plain_packed_kernel(PlainArgs a) {
    constexpr int D = packed_slots<KIND, FT>();
    __shared__ FT lds[table_words<KIND, FT>()];
    __shared__ double2 stash[kPackedWaves][kDensePairs * D];
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const int wave = (int)(threadIdx.x >> 6);
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    const double t0 = a.times[0];
    const Dop853Job J = make_dop853_job(a.times, a.nt, a.rtol, a.atol, a.max_attempts, t0,
                                        a.times[a.nt - 1]);
    OutSink sink{a.out, a.n, 0, a.staged != 0};
    Dop853Lane L;
    // lanes past n take part in the wavefront's dense output only
    bool live = slot < a.n;
    if (live) {
        sink.idx = a.order ? (int64_t)a.order[slot] : slot;
        double y0[DIM];
        load_state(a.state0, sink.idx, y0);
        dop853_init(L, y0, t0, J, f, sink);
    }
    int p = 0;
    for (;;) {
        // every orbit ends (at the latest after max_attempts); then the
        // wavefront runs what is left in its stash and leaves
        const bool any_live = __ballot(live) != 0;
        if (!any_live && p == 0) break;
        bool fin = false, own = false;
        Dop853Dense R;
        if (live)
            fin = dop853_attempt_deferred(L, J, f, sink, [&](bool has, Dop853Dense& req, bool) {
                own = has;
                if (has) R = req;
            });
        R.idx = sink.idx;
        dense_round<D>(stash[wave], p, own, R, !any_live, J, f, sink, a.stats);
        if (fin) {
            write_packed_stats(a.stats, sink.idx, L);
            live = false;
        }
    }
}

template <int KIND, class FT, bool LT = false>
__global__ void __launch_bounds__(kPackedBlock) DRIFT_TWO_WAVES
dop853_persistent_packed_kernel(PlainArgs a) {
    constexpr int D = packed_slots<KIND, FT, LT>();
    __shared__ FT lds[table_words<KIND, FT>()];
    __shared__ double2 stash[kPackedWaves][kDensePairs * D];
    __shared__ double ltimes[LT ? kLdsTimes : 1];
    const double* times = stage_times<LT>(a.times, a.nt, ltimes);
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const int wave = (int)(threadIdx.x >> 6);
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    const double t0 = a.times[0];
    const Dop853Job J = make_dop853_job(times, a.nt, a.rtol, a.atol, a.max_attempts, t0,
                                        a.times[a.nt - 1]);
    OutSink sink{a.out, a.n, 0, a.staged != 0};
    Dop853Lane L;
    bool need = true, live = true;
    int p = 0;
    for (;;) {
        int64_t orbit = -1;
        if (!claim_orbit(need && live, a.queue, a.n, a.order, orbit)) live = false;
        if (live && need) {
            sink.idx = orbit;
            double y0[DIM];
            load_state(a.state0, orbit, y0);
            dop853_init(L, y0, t0, J, f, sink);
            need = false;
        }
        // every orbit ends (at the latest after max_attempts), so every lane
        // reaches an empty queue; then the wavefront runs what is left in its
        // stash and leaves
        const bool any_live = __ballot(live) != 0;
        if (!any_live && p == 0) break;
        bool fin = false, own = false;
        Dop853Dense R;
        if (live)
            fin = dop853_attempt_deferred(L, J, f, sink, [&](bool has, Dop853Dense& req, bool) {
                own = has;
                if (has) R = req;
            });
        R.idx = sink.idx;
        // Pace the sample stores: the previous round's must be done before
        // this round issues more.  The attempt's arithmetic has hidden most
        // of their latency by now.  Without this wait (and without the one
        // the claim used to cause at the top of the loop), rtol 1e-7 launches
        // with direct output, about three samples per step, ran 2-18% slower.
        // Placed here rather than at the top of the loop, the 1e-12 launch
        // runs 1% faster.
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        dense_round<D>(stash[wave], p, own, R, !any_live, J, f, sink, a.stats);
        if (fin) {
            write_packed_stats(a.stats, sink.idx, L);
            need = true;
        }
    }
}

// Zeroes the status words of n statistics records (the packed kernels OR
// into them).
__global__ void zero_status_kernel(drift_orbit_stats* stats, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) stats[i].status = 0;
}

__global__ void reset_queue_kernel(unsigned long long* queue) { *queue = 0ull; }

// Staged samples -> time-major output.  One block moves a tile of 64 orbits x
// kStageBlock samples: it reads the tile's staged records (contiguous, 48 KB
// for a full tile), and writes each of its samples as 64 consecutive records
// (3 KB) of the output.  The LDS rows are padded by one 16-byte word to spread
// the column reads over the banks.
constexpr int kUnstageOrbits = 64;
constexpr int kUnstageRow = kStageBlock * 3;  // double2 words per orbit in a tile

// Streaming (nontemporal) 16-byte loads and stores: each word of the staged
// buffer is read once and each output word written once, so they bypass the
// caches' retention (3.6% faster than cached accesses on a 25.8 GB transpose).
typedef double v2d_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 stream_load(const double2* p) {
    const v2d_t v = __builtin_nontemporal_load(reinterpret_cast<const v2d_t*>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void stream_store(double2 x, double2* p) {
    v2d_t v;
    v.x = x.x;
    v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<v2d_t*>(p));
}

__global__ void __launch_bounds__(256) unstage_kernel(const double* __restrict__ staged,
                                                      double* __restrict__ out, int64_t n,
                                                      int32_t nt) {
    __shared__ double2 tile[kUnstageOrbits][kUnstageRow + 1];
    const int64_t o0 = (int64_t)blockIdx.x * kUnstageOrbits;
    const int64_t orbits = min((int64_t)kUnstageOrbits, n - o0);
    const int32_t nblocks = (nt + kStageBlock - 1) / kStageBlock;
    for (int32_t b = blockIdx.y; b < nblocks; b += gridDim.y) {
        const double2* src = reinterpret_cast<const double2*>(staged) +
                             ((int64_t)b * n + o0) * kStageBlock * 3;
        const int words = (int)orbits * kUnstageRow;
        for (int q = threadIdx.x; q < words; q += blockDim.x)
            tile[q / kUnstageRow][q % kUnstageRow] = stream_load(src + q);
        __syncthreads();
        const int32_t s0 = b * kStageBlock;
        const int rows = min(kStageBlock, nt - s0);
        const int row_words = (int)orbits * 3;
        double2* dst = reinterpret_cast<double2*>(out);
        for (int q = threadIdx.x; q < rows * row_words; q += blockDim.x) {
            const int r = q / row_words, rem = q % row_words;
            const int o = rem / 3, part = rem % 3;
            stream_store(tile[o][r * 3 + part], dst + (((int64_t)(s0 + r)) * n + o0 + o) * 3 + part);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Cost ordering: a counting sort of the orbits by a bucket of their estimated
// step rate at t0, highest first, so that orbits of similar step counts share
// wavefronts and the long ones start first.  The rate is the orbital
// frequency omega = sqrt(|a| / r).  For the plain schedule, whose wavefronts
// last as long as their slowest orbit, the rate of an orbit in a Milky-Way
// potential also counts its disk crossings: an orbit that leaves the disk
// plane crosses the thin disk at up to v_zmax = sqrt(2 E_z) and takes short
// steps there, so its rate adds a quarter of the crossing rate v_zmax / b (b
// the disk's scale height).  On the 1M-orbit disk workload the mean attempts
// of a 64-orbit wavefront over its slowest orbit's rose from 0.755 (omega
// alone) to 0.833 at rtol 1e-12 and from 0.666 to 0.817 at 1e-7, and the
// plain launch took 134 ms instead of 144 ms.  The persistent schedule
// refills lanes, so only the order of claims matters to it; there omega
// alone groups orbits whose steps emit samples equally often (fewer lanes
// idle in the dense-output branch) and measured 1.5-2.7% faster at rtol 1e-7.
// Buckets are 1/16 octave of the rate.
// ---------------------------------------------------------------------------
constexpr double kCrossingWeight = 0.25;

// Vertical energy E_z = v_z^2 / 2 + Phi(R, z) - Phi(R, 0) of a Milky-Way
// orbit: the Miyamoto-Nagai part in closed form, the spherical part as
// S z^2 / 2 with S from the spherical force's z component -S z (the total
// force d less the disk's).
__device__ __forceinline__ double mw_vertical_energy(const PotParams& p, const double* y,
                                                     const double* d) {
    const double R2 = y[0] * y[0] + y[1] * y[1], z = y[2];
    const double zb = std::sqrt(z * z + p.disk_b2), b = std::sqrt(p.disk_b2);
    const double D = std::sqrt(R2 + (p.disk_a + zb) * (p.disk_a + zb));
    const double D0 = std::sqrt(R2 + (p.disk_a + b) * (p.disk_a + b));
    const double dphi_disk = p.disk_amp * (1.0 / D0 - 1.0 / D);
    const double az_disk = zb > 0.0 ? -p.disk_amp * z * (p.disk_a + zb) / (zb * D * D * D) : 0.0;
    const double dphi_sph = -0.5 * (d[5] - az_disk) * z;
    return 0.5 * y[5] * y[5] + dphi_disk + dphi_sph;
}

template <int KIND>
__device__ __forceinline__ uint32_t cost_bin(const PotParams& p, const double* y, const double* d,
                                             bool crossings) {
    const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
    const double a2 = d[3] * d[3] + d[4] * d[4] + d[5] * d[5];
    double rate = std::sqrt(std::sqrt(a2 / r2));
    if constexpr (static_kind(KIND) == POT_MW) {
        if (crossings && p.disk_b2 > 0.0) {
            const double ez = mw_vertical_energy(p, y, d);
            if (ez > 0.0) rate += kCrossingWeight * std::sqrt(2.0 * ez / p.disk_b2);
        }
    }
    return order_bin_of_rate(rate);
}

__global__ void order_reset_kernel(uint32_t* bins) {
    for (int k = threadIdx.x; k < kOrderBins; k += blockDim.x) bins[k] = 0u;
}

template <int KIND, class FT>
__global__ void __launch_bounds__(kDefaultBlock) order_key_kernel(PlainArgs a, uint16_t* keys,
                                                                   uint32_t* bins,
                                                                   int32_t crossings) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    __shared__ uint32_t local[kOrderBins];
    for (int k = threadIdx.x; k < kOrderBins; k += blockDim.x) local[k] = 0u;
    const FT* table = stage_table<KIND, FT>(a.table, lds);  // ends with a barrier
    if constexpr (!kind_has_table(KIND)) __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n) {
        const auto f = KindForce<KIND, FT>::make(a.pot, table);
        double y[DIM], d[DIM];
        load_state(a.state0, i, y);
        f(a.times[0], y, d);
        const uint32_t b = cost_bin<KIND>(a.pot, y, d, crossings != 0);
        keys[i] = (uint16_t)b;
        atomicAdd(&local[b], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kOrderBins; k += blockDim.x)
        if (local[k]) atomicAdd(&bins[k], local[k]);
}

// Exclusive prefix sum of the bin counts, in place; one block of kOrderBins
// threads.
__global__ void __launch_bounds__(kOrderBins) order_scan_kernel(uint32_t* bins) {
    __shared__ uint32_t s[kOrderBins];
    const int t = threadIdx.x;
    const uint32_t own = bins[t];
    s[t] = own;
    __syncthreads();
    for (int off = 1; off < kOrderBins; off <<= 1) {
        const uint32_t v = t >= off ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    bins[t] = s[t] - own;
}

// Each block places kScatterItems consecutive orbits: ranks within a bin come
// from LDS atomics, and one global atomic per (block, non-empty bin) reserves
// the block's range of that bin, so hot bins see few global atomics.
constexpr int kScatterThreads = kOrderBins;
constexpr int kScatterPerThread = 4;
constexpr int kScatterItems = kScatterThreads * kScatterPerThread;

__global__ void __launch_bounds__(kScatterThreads) order_scatter_kernel(const uint16_t* keys,
                                                                        int64_t n, uint32_t* bins,
                                                                        uint32_t* order) {
    __shared__ uint32_t count[kOrderBins];
    __shared__ uint32_t base[kOrderBins];
    const int t = threadIdx.x;
    count[t] = 0u;
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * kScatterItems;
    uint32_t key[kScatterPerThread], rank[kScatterPerThread];
#pragma unroll
    for (int k = 0; k < kScatterPerThread; ++k) {
        const int64_t i = first + (int64_t)k * kScatterThreads + t;
        if (i < n) {
            key[k] = keys[i];
            rank[k] = atomicAdd(&count[key[k]], 1u);
        }
    }
    __syncthreads();
    if (count[t]) base[t] = atomicAdd(&bins[t], count[t]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScatterPerThread; ++k) {
        const int64_t i = first + (int64_t)k * kScatterThreads + t;
        if (i < n) order[base[key[k]] + rank[k]] = (uint32_t)i;
    }
}

struct ForceArgs {
    const double* pos;
    const double* t;  // per-point times, or NULL for t = 0
    double* acc;
    int64_t n;
    PotParams pot;
    const void* table;
};

template <int KIND, class FT>
__global__ void __launch_bounds__(kDefaultBlock) force_kernel(ForceArgs a) {
    __shared__ FT lds[kind_has_table(KIND) ? (sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS) : 1];
    const FT* table = stage_table<KIND, FT>(a.table, lds);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const auto f = KindForce<KIND, FT>::make(a.pot, table);
    double q[DIM] = {a.pos[3 * i], a.pos[3 * i + 1], a.pos[3 * i + 2], 0.0, 0.0, 0.0};
    double d[DIM];
    f(a.t ? a.t[i] : 0.0, q, d);
    a.acc[3 * i] = d[3];
    a.acc[3 * i + 1] = d[4];
    a.acc[3 * i + 2] = d[5];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {
thread_local std::string g_last_error;
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace drift

namespace drift {

int fail(int code, const std::string& msg) {
    set_last_error(msg);
    return code;
}

int ensure_buffer(DeviceBuffer& b, size_t bytes) {
    if (b.bytes >= bytes) return DRIFT_OK;
    if (b.ptr) DRIFT_HIP_TRY(hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
    DRIFT_HIP_TRY(hipMalloc(&b.ptr, bytes));
    b.bytes = bytes;
    return DRIFT_OK;
}

namespace {

bool finite(double x) { return std::isfinite(x); }

// Whether two descriptors have the same static force (the moving-origin set
// does not enter the static force's tables).
bool same_potential(drift_potential a, drift_potential b) {
    a.origins = 0;
    b.origins = 0;
    return std::memcmp(&a, &b, sizeof(drift_potential)) == 0;
}

}  // namespace

int validate_potential(const drift_potential& p) {
    if (p.origins < 0)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "drift_potential.origins must be 0 or a set id");
    switch (p.kind) {
        case POT_GALPY_KEPLER:
            if (p.origins != 0)
                return fail(DRIFT_ERR_UNSUPPORTED,
                            "moving origins cannot be added to the galpy Kepler force");
            return DRIFT_OK;
        case POT_NONE:
            return DRIFT_OK;
        case POT_KEPLER:
            if (!finite(p.amp)) return fail(DRIFT_ERR_INVALID_ARGUMENT, "Kepler amp must be finite");
            return DRIFT_OK;
        case POT_PLUMMER:
            if (!finite(p.amp) || !finite(p.b) || p.b < 0.0)
                return fail(DRIFT_ERR_INVALID_ARGUMENT,
                            "Plummer amp must be finite and radius finite and >= 0");
            return DRIFT_OK;
        case POT_MW: {
            const double v[] = {p.bulge_amp, p.bulge_alpha, p.bulge_rc, p.disk_amp, p.disk_a,
                                p.disk_b, p.halo_amp, p.halo_a};
            for (double x : v)
                if (!finite(x)) return fail(DRIFT_ERR_INVALID_ARGUMENT, "MW parameters must be finite");
            if (!(p.bulge_alpha >= 0.0 && p.bulge_alpha < 3.0))
                return fail(DRIFT_ERR_UNSUPPORTED, "MW bulge alpha must lie in [0, 3)");
            if (!(p.bulge_rc > 0.0) || !(p.halo_a > 0.0) || p.disk_a < 0.0 || p.disk_b < 0.0)
                return fail(DRIFT_ERR_INVALID_ARGUMENT,
                            "MW scale lengths must be positive (disk a, b may be 0)");
            return DRIFT_OK;
        }
        default:
            return fail(DRIFT_ERR_UNSUPPORTED, "unknown potential kind " + std::to_string(p.kind));
    }
}

// Resolves a potential descriptor into kernel parameters (and, for MW, a
// cached device table).  potential == NULL selects the reference's unit
// galpy Kepler force.
int resolve_potential(drift_session* s, const drift_potential* potential, PotParams& params,
                      const void** table64, const void** table32) {
    drift_potential p{};
    p.kind = POT_GALPY_KEPLER;
    if (potential) p = *potential;
    int rc = validate_potential(p);
    if (rc != DRIFT_OK) return rc;
    *table64 = nullptr;
    *table32 = nullptr;
    params = PotParams{};
    params.kind = p.kind;
    // moving origins (a static NONE kind runs as an empty moving set)
    if (p.origins != 0 || p.kind == POT_NONE) {
        params.kind |= POT_MOVING;
        params.n_segments = 1;
        if (p.origins != 0) {
            const OriginSet* set = nullptr;
            for (const auto& o : s->origin_sets)
                if (o.id == p.origins) set = &o;
            if (!set)
                return fail(DRIFT_ERR_INVALID_ARGUMENT,
                            "drift_potential.origins " + std::to_string(p.origins) +
                                " is not a moving-origin set of this session");
            params.tracks = set->d_tracks;
            params.origin_params = set->d_params;
            params.n_origins = set->n_origins;
            params.n_segments = set->n_segments;
            params.t_start = set->t_start;
            params.inv_dt = (double)set->n_segments / (set->t_end - set->t_start);
        }
    }
    if (p.kind == POT_KEPLER) {
        params.amp = p.amp;
    } else if (p.kind == POT_PLUMMER) {
        params.amp = p.amp;
        params.b2 = p.b * p.b;
    } else if (p.kind == POT_MW) {
        const PotParams moving = params;
        auto with_origins = [&](const PotParams& mw) {
            params = mw;
            params.kind = moving.kind;
            params.tracks = moving.tracks;
            params.origin_params = moving.origin_params;
            params.n_origins = moving.n_origins;
            params.n_segments = moving.n_segments;
            params.t_start = moving.t_start;
            params.inv_dt = moving.inv_dt;
        };
        for (auto& e : s->tables) {
            if (same_potential(e.key, p)) {
                with_origins(e.params);
                *table64 = e.d_table64;
                *table32 = e.d_table32;
                return DRIFT_OK;
            }
        }
        TableEntry e;
        e.key = p;
        e.key.origins = 0;
        build_mw_params(p, e.params);
        std::vector<double> h64(TABLE_DOUBLES), far(FAR_DOUBLES);
        build_mw_table(p, h64.data());
        build_mw_far_table(p, far.data());
        std::vector<float> h32(TABLE32_FLOATS);
        build_mw_table_f32(p, h32.data());
        hipError_t err = hipMalloc(&e.d_table64, TABLE_DOUBLES * sizeof(double));
        if (err == hipSuccess) err = hipMalloc(&e.d_table32, TABLE32_FLOATS * sizeof(float));
        if (err == hipSuccess) err = hipMalloc(&e.d_far, FAR_DOUBLES * sizeof(double));
        if (err == hipSuccess)
            err = hipMemcpy(e.d_table64, h64.data(), TABLE_DOUBLES * sizeof(double),
                            hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(e.d_table32, h32.data(), TABLE32_FLOATS * sizeof(float),
                            hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(e.d_far, far.data(), FAR_DOUBLES * sizeof(double),
                            hipMemcpyHostToDevice);
        if (err != hipSuccess) {
            for (void* q : {(void*)e.d_table64, (void*)e.d_table32, (void*)e.d_far})
                if (q) (void)hipFree(q);
            set_last_error(std::string("HIP error: ") + hipGetErrorString(err) +
                           " uploading the Milky-Way force tables");
            return err == hipErrorOutOfMemory ? DRIFT_ERR_OUT_OF_MEMORY : DRIFT_ERR_DEVICE;
        }
        e.params.far_table = e.d_far;
        s->tables.push_back(e);
        with_origins(e.params);
        *table64 = e.d_table64;
        *table32 = e.d_table32;
    }
    return DRIFT_OK;
}

namespace {

using PlainKernel = void (*)(PlainArgs);

// The DOP853 kernels of a static potential with the time grid in LDS (lt:
// at most kLdsTimes outputs) or read from global memory.
template <int KIND, class FT, bool LT>
PlainKernel pick_dop853(int schedule, bool wide) {
    if constexpr (KIND == POT_MW && sizeof(FT) == 8)
        if (!schedule && wide) return dop853_plain_kernel<KIND, FT, 3, LT>;
    return schedule ? dop853_persistent_kernel<KIND, FT, LT> : dop853_plain_kernel<KIND, FT, 2, LT>;
}

template <int KIND, class FT>
PlainKernel pick_schedule(int method, int schedule, bool wide, bool lt) {
    if (method == DRIFT_METHOD_DOP853) {
        if constexpr ((KIND & POT_MOVING) == 0)
            if (lt) return pick_dop853<KIND, FT, true>(schedule, wide);
        return pick_dop853<KIND, FT, false>(schedule, wide);
    }
    return schedule ? dopr54_persistent_kernel<KIND, FT> : dopr54_plain_kernel<KIND, FT>;
}

// DOP853 in 512-thread blocks: dense output packed across each wavefront's
// lanes (static potentials).  The plain packed kernel reads the time grid
// from global memory.
template <int KIND, class FT>
PlainKernel pick_packed(int schedule, bool lt) {
    if (!schedule) return dop853_plain_packed_kernel<KIND, FT>;
    return lt ? dop853_persistent_packed_kernel<KIND, FT, true>
              : dop853_persistent_packed_kernel<KIND, FT, false>;
}

template <int KIND>
PlainKernel pick_plain(int method, int precision, int schedule, bool packed, bool wide, bool lt) {
    if (packed)
        return precision == DRIFT_PRECISION_MIXED ? pick_packed<KIND, float>(schedule, lt)
                                                  : pick_packed<KIND, double>(schedule, lt);
    return precision == DRIFT_PRECISION_MIXED
               ? pick_schedule<KIND, float>(method, schedule, wide, lt)
               : pick_schedule<KIND, double>(method, schedule, wide, lt);
}

// Moving-origin kinds exist in FP64 only (launch_integrator refuses MIXED),
// without packed dense output and with the time grid in global memory.  lt:
// the launch's grid fits LDS (lds_times(nt)).
PlainKernel kernel_for(int method, int kind, int precision, int schedule, bool packed, bool wide,
                       bool lt) {
    switch (kind) {
        case POT_GALPY_KEPLER:
            return pick_plain<POT_GALPY_KEPLER>(method, precision, schedule, packed, wide, lt);
        case POT_KEPLER:
            return pick_plain<POT_KEPLER>(method, precision, schedule, packed, wide, lt);
        case POT_PLUMMER:
            return pick_plain<POT_PLUMMER>(method, precision, schedule, packed, wide, lt);
        case POT_MW: return pick_plain<POT_MW>(method, precision, schedule, packed, wide, lt);
        case POT_MOVING | POT_KEPLER:
            return pick_schedule<POT_MOVING | POT_KEPLER, double>(method, schedule, wide, false);
        case POT_MOVING | POT_PLUMMER:
            return pick_schedule<POT_MOVING | POT_PLUMMER, double>(method, schedule, wide, false);
        case POT_MOVING | POT_MW:
            return pick_schedule<POT_MOVING | POT_MW, double>(method, schedule, wide, false);
        default:
            return pick_schedule<POT_MOVING | POT_NONE, double>(method, schedule, wide, false);
    }
}

// Threads per block of an integrator launch, and whether it is DOP853's
// packed dense output.  block_size 64..256 (steps of 64) runs the inline
// dense output; 512 (kPackedBlock) the packed one, which exists for DOP853
// with a static potential only; 0 takes the default: packed for persistent
// DOP853 launches with direct output and a static potential, where it
// measured faster (1M Milky-Way orbits: 120.7 against 123.6 ms at rtol 1e-12,
// 64.9 against 68.1 ms in the mixed mode at rtol 1e-7), else kDefaultBlock
// (the plain schedule's 512-thread blocks wait for their slowest wavefront,
// and staged output gains nothing from packing).
// The schedule of a launch that names none (host-buffer entry points,
// drift_launch_shape without options, TUNED without a stored choice): DOP853
// persistent; DOPR54 plain, which ran 4.5-31% faster than its persistent
// schedule on Kepler, Plummer and Milky-Way batches of 1,024 to 262,144
// orbits (FP64 and mixed, rtol 1e-10 and 1e-7) and 8% with one moving origin.
constexpr int default_schedule(int method) {
    return method == DRIFT_METHOD_DOPR54 ? DRIFT_SCHEDULE_PLAIN : DRIFT_SCHEDULE_PERSISTENT;
}

int resolve_block(int method, int kind, int block_size, int schedule, bool staged, int& block,
                  bool& packed) {
    const bool packable = method == DRIFT_METHOD_DOP853 && !(kind & POT_MOVING);
    const bool pack_default = packable && schedule == DRIFT_SCHEDULE_PERSISTENT && !staged;
    block = block_size ? block_size : (pack_default ? kPackedBlock : kDefaultBlock);
    packed = block == kPackedBlock;
    if (packed && !packable)
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "block_size 512 (packed dense output) is for DOP853 with a static potential");
    if (!packed && (block < 64 || block > kDefaultBlock || block % 64 != 0))
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "block_size must be a multiple of 64 in [64, 256], or 512 for DOP853");
    return DRIFT_OK;
}

using WindowKernel = void (*)(PlainArgs, WindowArgs);

WindowKernel window_kernel_for(int kind) {
    switch (static_kind(kind)) {
        case POT_KEPLER: return dop853_window_kernel<POT_MOVING | POT_KEPLER>;
        case POT_PLUMMER: return dop853_window_kernel<POT_MOVING | POT_PLUMMER>;
        case POT_MW: return dop853_window_kernel<POT_MOVING | POT_MW>;
        default: return dop853_window_kernel<POT_MOVING | POT_NONE>;
    }
}

// Output samples per time window: as many as keep the window's segments
// (those its samples span, plus one a step may overshoot into, plus one for a
// window that starts inside a segment) within the rows per origin, taking the
// time grid to span the tracks; at least one.
int32_t window_samples(int32_t rows_per_origin, int32_t n_segments, int32_t nt) {
    const double seg_per_sample = (double)n_segments / (double)std::max(nt - 1, 1);
    const double k = std::floor((double)(rows_per_origin - 2) / seg_per_sample);
    return (int32_t)std::max(1.0, std::min(k, (double)std::max(nt - 1, 1)));
}

using KeyKernel = void (*)(PlainArgs, uint16_t*, uint32_t*, int32_t);

template <int KIND>
KeyKernel pick_key(int precision) {
    return precision == DRIFT_PRECISION_MIXED ? order_key_kernel<KIND, float>
                                              : order_key_kernel<KIND, double>;
}

KeyKernel key_kernel_for(int kind, int precision) {
    switch (kind) {
        case POT_GALPY_KEPLER: return pick_key<POT_GALPY_KEPLER>(precision);
        case POT_KEPLER: return pick_key<POT_KEPLER>(precision);
        case POT_PLUMMER: return pick_key<POT_PLUMMER>(precision);
        case POT_MW: return pick_key<POT_MW>(precision);
        case POT_MOVING | POT_KEPLER: return order_key_kernel<POT_MOVING | POT_KEPLER, double>;
        case POT_MOVING | POT_PLUMMER: return order_key_kernel<POT_MOVING | POT_PLUMMER, double>;
        case POT_MOVING | POT_MW: return order_key_kernel<POT_MOVING | POT_MW, double>;
        default: return order_key_kernel<POT_MOVING | POT_NONE, double>;
    }
}

int floor_log2(size_t n) {
    int k = 0;
    while (n > 1) {
        n >>= 1;
        ++k;
    }
    return k;
}

const TunedEntry* find_tuned(const drift_session* s, int method, int kind, int precision,
                             size_t n) {
    const int l = floor_log2(n);
    for (const TunedEntry& e : s->tuned)
        if (e.method == method && e.kind == kind && e.precision == precision && e.log2n == l)
            return &e;
    return nullptr;
}

bool table_cached(const drift_session* s, const drift_potential& p) {
    for (const auto& e : s->tables)
        if (same_potential(e.key, p)) return true;
    return false;
}

// Picks the next work-queue counter of the ring.  Eager launches rotate over
// the counters no captured graph owns (stream order retires a counter's
// previous user before the ring comes back to it); a captured launch takes
// one for the life of the session, as its graph may be replayed at any time.
// A few counters always stay free for eager launches.
int take_queue_slot(drift_session* s, bool capturing) {
    const unsigned owned = (unsigned)std::count(s->queue_owned.begin(), s->queue_owned.end(), true);
    if (capturing && owned + 8 >= kQueueRing) return -1;
    for (unsigned k = 0; k < kQueueRing; ++k) {
        const unsigned slot = (s->queue_next++) % kQueueRing;
        if (s->queue_owned[slot]) continue;
        if (capturing) s->queue_owned[slot] = true;
        return (int)slot;
    }
    return -1;
}

// Integrator grid for n orbits: one lane per orbit for the plain schedule;
// the device's resident block count (capped by n) for the persistent one.
int integrator_grid(int cu_count, PlainKernel k, int schedule, int block, size_t n,
                    int64_t& blocks) {
    blocks = ((int64_t)n + block - 1) / block;
    if (schedule == DRIFT_SCHEDULE_PERSISTENT) {
        int per_cu = 0;
        DRIFT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0));
        const int64_t resident = (int64_t)std::max(per_cu, 1) * std::max(cu_count, 1);
        blocks = std::min(blocks, resident);
    }
    return DRIFT_OK;
}

// Orbit counts from which DRIFT_ORDER_AUTO sorts a launch (either schedule):
// below it the machine is not full and the counting sort would not pay for
// itself.
constexpr size_t kAutoOrderMinOrbits = 16384;

int launch_cost_order(const PlainArgs& a, int precision, int schedule, uint16_t* keys,
                      uint32_t* bins, uint32_t* slots, hipStream_t st);

// Builds s->order_slots (slot -> orbit, most expensive bin first) on stream
// st.  The launch that reads it records s->order_done after its integration
// kernel; the next build waits for that event, so launches on different
// streams never overwrite a map still in use.
int build_cost_order(drift_session* s, const PlainArgs& a, int precision, int schedule,
                     hipStream_t st) {
    const size_t n = (size_t)a.n;
    int rc;
    if ((rc = ensure_buffer(s->order_keys, n * sizeof(uint16_t))) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->order_slots, n * sizeof(uint32_t))) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->order_bins, kOrderBins * sizeof(uint32_t))) != DRIFT_OK) return rc;
    if (!s->order_done)
        DRIFT_HIP_TRY(hipEventCreateWithFlags(&s->order_done, hipEventDisableTiming));
    else
        DRIFT_HIP_TRY(hipStreamWaitEvent(st, s->order_done, 0));
    return launch_cost_order(a, precision, schedule, static_cast<uint16_t*>(s->order_keys.ptr),
                             static_cast<uint32_t*>(s->order_bins.ptr),
                             static_cast<uint32_t*>(s->order_slots.ptr), st);
}

// The four kernels of the cost ordering on caller-given scratch: keys (n
// u16), bins (kOrderBins u32) and the slot -> orbit map (n u32).  The key
// counts disk crossings for the plain schedule only.
int launch_cost_order(const PlainArgs& a, int precision, int schedule, uint16_t* keys,
                      uint32_t* bins, uint32_t* slots, hipStream_t st) {
    const int32_t crossings = schedule == DRIFT_SCHEDULE_PLAIN ? 1 : 0;
    const size_t n = (size_t)a.n;
    const unsigned blocks = (unsigned)((n + kDefaultBlock - 1) / kDefaultBlock);
    hipLaunchKernelGGL(order_reset_kernel, dim3(1), dim3(256), 0, st, bins);
    hipLaunchKernelGGL(key_kernel_for(a.pot.kind, precision), dim3(blocks), dim3(kDefaultBlock), 0,
                       st, a, keys, bins, crossings);
    hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(kOrderBins), 0, st, bins);
    const unsigned scatter_blocks = (unsigned)((n + kScatterItems - 1) / kScatterItems);
    hipLaunchKernelGGL(order_scatter_kernel, dim3(scatter_blocks), dim3(kScatterThreads), 0, st,
                       (const uint16_t*)keys, a.n, bins, slots);
    DRIFT_HIP_TRY(hipGetLastError());
    return DRIFT_OK;
}


constexpr uint32_t kDefaultMaxAttempts = 1u << 22;
constexpr double kMixedMinRtol = 1e-7;

int launch_integrator(drift_session* s, int method, const double* d_state0, const double* d_times,
                      double* d_out, size_t n, size_t nt, double rtol_ln, double atol_ln,
                      double dt_one_init, const drift_potential* potential,
                      drift_orbit_stats* d_stats, const drift_launch_options* options,
                      hipStream_t stream) {
    if (!s) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n must be >= 1");
    if (nt < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt must be >= 2");
    if (nt > (size_t)INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt too large");
    if (!d_state0 || !d_times || !d_out)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "state0, times and state_out must be non-NULL");
    if ((reinterpret_cast<uintptr_t>(d_state0) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "state0 and state_out must be 16-byte aligned");
    if (!finite(rtol_ln) || !finite(atol_ln))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "rtol and atol (natural logs) must be finite");
    drift_launch_options opt{};
    if (options) opt = *options;
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    if (opt.schedule < DRIFT_SCHEDULE_PLAIN || opt.schedule > DRIFT_SCHEDULE_TUNED)
        return fail(DRIFT_ERR_UNSUPPORTED, "schedule must be PLAIN, PERSISTENT or TUNED");
    if (opt.ordering < DRIFT_ORDER_AUTO || opt.ordering > DRIFT_ORDER_COST)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "ordering must be AUTO, INDEX or COST");
    if (opt.staging < DRIFT_STAGING_AUTO || opt.staging > DRIFT_STAGING_STAGED)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "staging must be AUTO, DIRECT or STAGED");

    PlainArgs a{};
    a.state0 = d_state0;
    a.times = d_times;
    a.out = d_out;
    a.stats = d_stats;
    a.n = (int64_t)n;
    a.nt = (int32_t)nt;
    a.max_attempts = opt.max_attempts ? opt.max_attempts : kDefaultMaxAttempts;
    // glibc exp, as dop853.rs:340-341 applies to the logarithmic tolerances.
    a.rtol = std::exp(rtol_ln);
    a.atol = std::exp(atol_ln);
    a.dt_one_init = dt_one_init;
    dopr54_bounds(a.bound2);
    if (opt.precision == DRIFT_PRECISION_MIXED && a.rtol < kMixedMinRtol * (1.0 - 1e-9))
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "mixed precision (FP32 stage forces) requires rtol >= 1e-7");

    DRIFT_HIP_TRY(hipSetDevice(s->device));
    hipStream_t st = stream ? stream : s->stream;
    // Under stream capture (hipGraph) nothing may allocate or synchronise:
    // the potential's table and the queue ring must exist from an earlier
    // eager launch, and the launch owns its queue counter for good.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    DRIFT_HIP_TRY(hipStreamIsCapturing(st, &capture));
    const bool capturing = capture != hipStreamCaptureStatusNone;
    if (capturing && potential && potential->kind == POT_MW && !table_cached(s, *potential))
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "stream capture: run one eager launch with this potential on this session "
                    "first (it builds and uploads the force table)");
    const void *t64 = nullptr, *t32 = nullptr;
    int rc = resolve_potential(s, potential, a.pot, &t64, &t32);
    if (rc != DRIFT_OK) return rc;
    if ((a.pot.kind & POT_MOVING) && opt.precision == DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "moving origins and the NONE potential need FP64 precision");
    a.table = opt.precision == DRIFT_PRECISION_MIXED ? t32 : t64;
    if (capturing && opt.staging == DRIFT_STAGING_STAGED)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "staged output uses the session's staging buffer, which a later eager launch "
                    "may reallocate, and cannot be captured into a graph; pass DRIFT_STAGING_DIRECT "
                    "(AUTO means DIRECT under capture)");
    if (opt.schedule == DRIFT_SCHEDULE_TUNED) {
        const TunedEntry* e = find_tuned(s, method, a.pot.kind, opt.precision, n);
        opt.schedule = e ? e->options.schedule : default_schedule(method);
        opt.block_size = e ? e->options.block_size : 0;
        opt.ordering = e ? e->options.ordering : DRIFT_ORDER_AUTO;
        if (opt.staging == DRIFT_STAGING_AUTO && e) opt.staging = e->options.staging;
        // the tuner's session-scratch choices have capture-safe equivalents
        // with the same results: index order and direct output
        if (capturing) {
            if (opt.ordering == DRIFT_ORDER_COST) opt.ordering = DRIFT_ORDER_INDEX;
            if (opt.staging == DRIFT_STAGING_STAGED) opt.staging = DRIFT_STAGING_DIRECT;
        }
    }

    if (capturing && opt.ordering == DRIFT_ORDER_COST)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "cost ordering uses per-session scratch and cannot be captured into a graph; "
                    "pass DRIFT_ORDER_INDEX (AUTO means INDEX under capture)");
    int block = 0;
    bool packed = false;
    if ((rc = resolve_block(method, a.pot.kind, opt.block_size, opt.schedule,
                            opt.staging == DRIFT_STAGING_STAGED, block, packed)) != DRIFT_OK)
        return rc;
    PlainKernel k = kernel_for(method, a.pot.kind, opt.precision, opt.schedule, packed,
                               wide_plain(n, s->cu_count), lds_times(nt));
    int64_t blocks = 0;
    if ((rc = integrator_grid(s->cu_count, k, opt.schedule, block, n, blocks)) != DRIFT_OK) return rc;
    if (opt.schedule == 1) {
        if (!s->queues) {
            if (capturing)
                return fail(DRIFT_ERR_INVALID_ARGUMENT,
                            "stream capture: run one eager persistent launch on this session "
                            "first (it allocates the work-queue counters)");
            DRIFT_HIP_TRY(hipMalloc(&s->queues, kQueueRing * sizeof(unsigned long long)));
            s->queue_owned.assign(kQueueRing, false);
        }
        const int slot = take_queue_slot(s, capturing);
        if (slot < 0)
            return fail(DRIFT_ERR_UNSUPPORTED,
                        "too many captured persistent launches on this session (each graph keeps "
                        "its own work-queue counter)");
        a.queue = s->queues + slot;
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, st, a.queue);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    if (blocks > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n too large");
    const bool cost_order =
        !capturing &&
        (opt.ordering == DRIFT_ORDER_COST ||
         (opt.ordering == DRIFT_ORDER_AUTO && n >= kAutoOrderMinOrbits));
    if (cost_order) {
        if (n > (size_t)UINT32_MAX)
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "cost ordering supports n < 2^32");
        if ((rc = build_cost_order(s, a, opt.precision, opt.schedule, st)) != DRIFT_OK) return rc;
        a.order = static_cast<const uint32_t*>(s->order_slots.ptr);
    }
    const bool staged = opt.staging == DRIFT_STAGING_STAGED;
    const size_t stage_blocks = (nt + kStageBlock - 1) / kStageBlock;
    if (staged) {
        const size_t bytes = n * stage_blocks * kStageBlock * DIM * sizeof(double);
        if ((rc = ensure_buffer(s->staging, bytes)) != DRIFT_OK) return rc;
        // the staging buffer is per session: wait for the last launch that used it
        if (!s->staging_done)
            DRIFT_HIP_TRY(hipEventCreateWithFlags(&s->staging_done, hipEventDisableTiming));
        else
            DRIFT_HIP_TRY(hipStreamWaitEvent(st, s->staging_done, 0));
        a.out = static_cast<double*>(s->staging.ptr);
        a.staged = 1;
    }
    // two or more moving origins, persistent: time windows with the tracks'
    // rows in LDS (the resumable state is session scratch, so not under
    // capture)
    const int32_t rows_per_origin = kWindowRowsMax / std::max(a.pot.n_origins, 1);
    const bool windowed = method == DRIFT_METHOD_DOP853 && (a.pot.kind & POT_MOVING) &&
                          a.pot.n_origins >= 2 && rows_per_origin >= 3 &&
                          opt.schedule == DRIFT_SCHEDULE_PERSISTENT && !capturing && !staged;
    if (windowed) {
        if ((rc = ensure_buffer(s->resume, (size_t)kResumeWords * n * sizeof(double))) != DRIFT_OK)
            return rc;
        if (!s->resume_done)
            DRIFT_HIP_TRY(hipEventCreateWithFlags(&s->resume_done, hipEventDisableTiming));
        else
            DRIFT_HIP_TRY(hipStreamWaitEvent(st, s->resume_done, 0));
        WindowKernel wk = window_kernel_for(a.pot.kind);
        if ((rc = integrator_grid(s->cu_count, (PlainKernel)nullptr, DRIFT_SCHEDULE_PLAIN, block, n,
                                  blocks)) != DRIFT_OK)
            return rc;
        int per_cu = 0;
        DRIFT_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wk, block, 0));
        blocks = std::min<int64_t>(blocks, (int64_t)std::max(per_cu, 1) * std::max(s->cu_count, 1));
        const int32_t samples = window_samples(rows_per_origin, a.pot.n_segments, (int32_t)nt);
        WindowArgs w{static_cast<double*>(s->resume.ptr), 1, 0, 0};
        for (;;) {
            w.stop = (int32_t)std::min<int64_t>((int64_t)w.s_begin + samples, (int64_t)nt - 1);
            if (!w.first) {
                hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, st, a.queue);
                DRIFT_HIP_TRY(hipGetLastError());
            }
            hipLaunchKernelGGL(wk, dim3((unsigned)blocks), dim3(block), 0, st, a, w);
            DRIFT_HIP_TRY(hipGetLastError());
            if (w.stop >= (int32_t)nt - 1) break;
            w.first = 0;
            w.s_begin = w.stop;
        }
        s->last_grid = (int32_t)blocks;
        s->last_block = block;
        if (cost_order) DRIFT_HIP_TRY(hipEventRecord(s->order_done, st));
        DRIFT_HIP_TRY(hipEventRecord(s->resume_done, st));
        return DRIFT_OK;
    }
    if (packed && a.stats) {
        hipLaunchKernelGGL(zero_status_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           a.stats, (int64_t)n);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(block), 0, st, a);
    DRIFT_HIP_TRY(hipGetLastError());
    s->last_grid = (int32_t)blocks;
    s->last_block = block;
    if (cost_order) DRIFT_HIP_TRY(hipEventRecord(s->order_done, st));
    if (staged) {
        const int64_t tiles = ((int64_t)n + kUnstageOrbits - 1) / kUnstageOrbits;
        if (tiles > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n too large");
        const unsigned gy = (unsigned)std::min<size_t>(stage_blocks, 65535);
        hipLaunchKernelGGL(unstage_kernel, dim3((unsigned)tiles, gy), dim3(256), 0, st,
                           (const double*)s->staging.ptr, d_out, (int64_t)n, (int32_t)nt);
        DRIFT_HIP_TRY(hipGetLastError());
        DRIFT_HIP_TRY(hipEventRecord(s->staging_done, st));
    }
    return DRIFT_OK;
}

int host_launch(drift_session* s, int method, const double* state0, const double* times,
                double* state_out, size_t n, size_t nt, double rtol_ln, double atol_ln,
                double dt_one_init, const drift_potential* potential, drift_orbit_stats* stats) {
    if (!s) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (!state0 || !times || !state_out)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "state0, times and state_out must be non-NULL");
    if (n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n must be >= 1");
    if (nt < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt must be >= 2");
    const auto setup_started = std::chrono::steady_clock::now();
    DRIFT_HIP_TRY(hipSetDevice(s->device));
    const size_t in_bytes = n * DIM * sizeof(double);
    const size_t out_bytes = nt * n * DIM * sizeof(double);
    int rc;
    if ((rc = ensure_buffer(s->state0, in_bytes)) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->times, nt * sizeof(double))) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->out, out_bytes)) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->stats, n * sizeof(drift_orbit_stats))) != DRIFT_OK) return rc;
    {
        // builds (first use) or finds the force tables before anything is
        // timed on the device
        PotParams pp{};
        const void *t64 = nullptr, *t32 = nullptr;
        if ((rc = resolve_potential(s, potential, pp, &t64, &t32)) != DRIFT_OK) return rc;
    }
    for (hipEvent_t& e : s->phase_events)
        if (!e) DRIFT_HIP_TRY(hipEventCreate(&e));
    s->have_last_launch = false;
    drift_launch_report report{};
    report.setup_s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - setup_started).count();
    report.h2d_bytes = in_bytes + nt * sizeof(double);
    report.d2h_bytes = out_bytes;
    DRIFT_HIP_TRY(hipEventRecord(s->phase_events[0], s->stream));
    DRIFT_HIP_TRY(hipMemcpyAsync(s->state0.ptr, state0, in_bytes, hipMemcpyHostToDevice, s->stream));
    DRIFT_HIP_TRY(hipMemcpyAsync(s->times.ptr, times, nt * sizeof(double), hipMemcpyHostToDevice,
                                 s->stream));
    DRIFT_HIP_TRY(hipEventRecord(s->phase_events[1], s->stream));
    // the method's fastest schedule, cost-ordered from 16384 orbits
    drift_launch_options opt{};
    opt.schedule = default_schedule(method);
    opt.ordering = DRIFT_ORDER_AUTO;
    rc = launch_integrator(s, method, (const double*)s->state0.ptr, (const double*)s->times.ptr,
                           (double*)s->out.ptr, n, nt, rtol_ln, atol_ln, dt_one_init, potential,
                           (drift_orbit_stats*)s->stats.ptr, &opt, s->stream);
    if (rc != DRIFT_OK) return rc;
    DRIFT_HIP_TRY(hipEventRecord(s->phase_events[2], s->stream));
    report.grid = s->last_grid;
    report.block = s->last_block;
    std::vector<drift_orbit_stats> host_stats;
    drift_orbit_stats* st = stats;
    if (!st) {
        host_stats.resize(n);
        st = host_stats.data();
    }
    DRIFT_HIP_TRY(hipMemcpyAsync(state_out, s->out.ptr, out_bytes, hipMemcpyDeviceToHost, s->stream));
    DRIFT_HIP_TRY(hipMemcpyAsync(st, s->stats.ptr, n * sizeof(drift_orbit_stats),
                                 hipMemcpyDeviceToHost, s->stream));
    DRIFT_HIP_TRY(hipEventRecord(s->phase_events[3], s->stream));
    DRIFT_HIP_TRY(hipStreamSynchronize(s->stream));
    {
        float ms[3] = {0.0f, 0.0f, 0.0f};
        for (int k = 0; k < 3; ++k)
            DRIFT_HIP_TRY(hipEventElapsedTime(&ms[k], s->phase_events[k], s->phase_events[k + 1]));
        report.h2d_s = ms[0] * 1e-3;
        report.kernel_s = ms[1] * 1e-3;
        report.d2h_s = ms[2] * 1e-3;
        s->last_launch = report;
        s->have_last_launch = true;
    }
    if (method == DRIFT_METHOD_DOPR54) {
        // integrate_kepler maps a non-zero error code to Err(code)
        // (dopr54.rs:123), reported as DispatchError::CpuIntegration.
        for (size_t i = 0; i < n; ++i) {
            const int code = st[i].status & DRIFT_STATUS_CODE_MASK;
            if (code != 0)
                return fail(DRIFT_ERR_INTEGRATION,
                            "GPU DOPR54 integration failed with error code " + std::to_string(code));
        }
    }
    return DRIFT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// The allocation-free DOP853 entry point.  Every buffer its kernels use beyond
// the caller's inputs and outputs lives in `scratch`, one array per GPU for
// the life of the process: the work-queue counter, the cost ordering's bins,
// keys and slot map (up to kNoAllocMaxOrdered orbits), and the Milky-Way force
// tables of the last MW potential used on that GPU.  Host code never passes
// the array's name to a kernel (on the host it names a stand-in, not device
// memory); it takes each device's address once from hipGetSymbolAddress.
// ---------------------------------------------------------------------------

constexpr size_t kNoAllocMaxOrdered = size_t(1) << 20;

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
constexpr size_t kNaQueue = 0;
constexpr size_t kNaBins = 256;
constexpr size_t kNaKeys = align256(kNaBins + kOrderBins * sizeof(uint32_t));
constexpr size_t kNaSlots = align256(kNaKeys + kNoAllocMaxOrdered * sizeof(uint16_t));
constexpr size_t kNaTable64 = align256(kNaSlots + kNoAllocMaxOrdered * sizeof(uint32_t));
constexpr size_t kNaTable32 = align256(kNaTable64 + TABLE_DOUBLES * sizeof(double));
constexpr size_t kNaFar = align256(kNaTable32 + TABLE32_FLOATS * sizeof(float));
constexpr size_t kNoAllocScratchBytes = align256(kNaFar + FAR_DOUBLES * sizeof(double));

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocDevice {
    bool ready = false;
    unsigned char* base = nullptr;  // this device's copy of `scratch`
    int cu_count = 0;
    bool have_table = false;        // the MW tables of table_key are resident
    drift_potential table_key{};
};

constexpr int kNoAllocMaxDevices = 64;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];
// guards g_noalloc and the host staging of a table being built
std::mutex g_noalloc_mutex;
double g_noalloc_h64[TABLE_DOUBLES];
float g_noalloc_h32[TABLE32_FLOATS];
double g_noalloc_far[FAR_DOUBLES];

// Once per device, with that device current: the device address of its
// scratch array and its CU count.  Caller holds g_noalloc_mutex.
int noalloc_setup(int device, NoAllocDevice*& out) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        return fail(DRIFT_ERR_UNSUPPORTED, "device index beyond the allocation-free entry "
                                           "point's " + std::to_string(kNoAllocMaxDevices) +
                                               " devices");
    NoAllocDevice& d = g_noalloc[device];
    if (!d.ready) {
        void* ptr = nullptr;
        DRIFT_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess)
            cus = 256;
        d.base = static_cast<unsigned char*>(ptr);
        d.cu_count = cus;
        d.ready = true;
    }
    out = &d;
    return DRIFT_OK;
}

// Makes the MW tables of p resident in the device's scratch.  A change of
// potential waits for the device (a call still running may read the old
// tables) and copies synchronously; it is refused under stream capture.
int noalloc_tables(NoAllocDevice& d, const drift_potential& p, bool capturing) {
    drift_potential key = p;
    key.origins = 0;
    if (d.have_table && std::memcmp(&d.table_key, &key, sizeof key) == 0) return DRIFT_OK;
    if (capturing)
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "stream capture: the Milky-Way tables of this potential are not resident; "
                    "make one call with it outside the capture first");
    build_mw_table(p, g_noalloc_h64);
    build_mw_table_f32(p, g_noalloc_h32);
    build_mw_far_table(p, g_noalloc_far);
    d.have_table = false;
    DRIFT_HIP_TRY(hipDeviceSynchronize());
    DRIFT_HIP_TRY(hipMemcpy(d.base + kNaTable64, g_noalloc_h64, sizeof g_noalloc_h64,
                            hipMemcpyHostToDevice));
    DRIFT_HIP_TRY(hipMemcpy(d.base + kNaTable32, g_noalloc_h32, sizeof g_noalloc_h32,
                            hipMemcpyHostToDevice));
    DRIFT_HIP_TRY(hipMemcpy(d.base + kNaFar, g_noalloc_far, sizeof g_noalloc_far,
                            hipMemcpyHostToDevice));
    d.table_key = key;
    d.have_table = true;
    return DRIFT_OK;
}

int noalloc_dop853(const double* d_state0, const double* d_times, double* d_out, size_t n,
                   size_t nt, double rtol_ln, double atol_ln, const drift_potential* potential,
                   drift_orbit_stats* d_stats, const drift_launch_options* options,
                   hipStream_t st) {
    if (n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n must be >= 1");
    if (nt < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt must be >= 2");
    if (nt > (size_t)INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt too large");
    if (!d_state0 || !d_times || !d_out)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "state0, times and state_out must be non-NULL");
    if ((reinterpret_cast<uintptr_t>(d_state0) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "state0 and state_out must be 16-byte aligned");
    if (!finite(rtol_ln) || !finite(atol_ln))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "rtol and atol (natural logs) must be finite");
    drift_launch_options opt{};
    if (options) opt = *options;
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    if (opt.schedule != DRIFT_SCHEDULE_PLAIN && opt.schedule != DRIFT_SCHEDULE_PERSISTENT)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "the allocation-free entry point takes PLAIN or PERSISTENT (tuned choices "
                    "live in a session)");
    if (opt.ordering < DRIFT_ORDER_AUTO || opt.ordering > DRIFT_ORDER_COST)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "ordering must be AUTO, INDEX or COST");
    if (opt.staging != DRIFT_STAGING_AUTO && opt.staging != DRIFT_STAGING_DIRECT)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "the allocation-free entry point writes its output directly (staging needs a "
                    "buffer of the output's size)");
    drift_potential p{};
    p.kind = POT_GALPY_KEPLER;
    if (potential) p = *potential;
    int rc = validate_potential(p);
    if (rc != DRIFT_OK) return rc;
    if (p.origins != 0)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "moving-origin sets belong to a session; use drift_galpy_dop853_device");
    const bool cost_order = opt.ordering == DRIFT_ORDER_COST ||
                            (opt.ordering == DRIFT_ORDER_AUTO && n >= kAutoOrderMinOrbits &&
                             n <= kNoAllocMaxOrdered);
    if (cost_order && n > kNoAllocMaxOrdered)
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "cost ordering in the allocation-free entry point takes at most " +
                        std::to_string(kNoAllocMaxOrdered) +
                        " orbits (its scratch size); pass DRIFT_ORDER_INDEX for more");

    PlainArgs a{};
    a.state0 = d_state0;
    a.times = d_times;
    a.out = d_out;
    a.stats = d_stats;
    a.n = (int64_t)n;
    a.nt = (int32_t)nt;
    a.max_attempts = opt.max_attempts ? opt.max_attempts : kDefaultMaxAttempts;
    a.rtol = std::exp(rtol_ln);  // glibc exp, dop853.rs:340-341
    a.atol = std::exp(atol_ln);
    a.dt_one_init = dp54::initial_step_sentinel;
    dopr54_bounds(a.bound2);
    if (opt.precision == DRIFT_PRECISION_MIXED && a.rtol < kMixedMinRtol * (1.0 - 1e-9))
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "mixed precision (FP32 stage forces) requires rtol >= 1e-7");
    if (p.kind == POT_NONE && opt.precision == DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "the NONE potential needs FP64 precision");

    int device = 0;
    DRIFT_HIP_TRY(hipGetDevice(&device));
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    DRIFT_HIP_TRY(hipStreamIsCapturing(st, &capture));
    const bool capturing = capture != hipStreamCaptureStatusNone;

    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    NoAllocDevice* d = nullptr;
    if ((rc = noalloc_setup(device, d)) != DRIFT_OK) return rc;
    unsigned char* base = d->base;
    a.pot = PotParams{};
    a.pot.kind = p.kind;
    if (p.kind == POT_KEPLER) {
        a.pot.amp = p.amp;
    } else if (p.kind == POT_PLUMMER) {
        a.pot.amp = p.amp;
        a.pot.b2 = p.b * p.b;
    } else if (p.kind == POT_NONE) {
        // free motion: the moving-origin kernel with no origins
        a.pot.kind = POT_MOVING | POT_NONE;
        a.pot.n_segments = 1;
    } else if (p.kind == POT_MW) {
        if ((rc = noalloc_tables(*d, p, capturing)) != DRIFT_OK) return rc;
        build_mw_params(p, a.pot);
        a.pot.far_table = reinterpret_cast<const double*>(base + kNaFar);
        a.table = opt.precision == DRIFT_PRECISION_MIXED
                      ? static_cast<const void*>(base + kNaTable32)
                      : static_cast<const void*>(base + kNaTable64);
    }
    int block = 0;
    bool packed = false;
    if ((rc = resolve_block(DRIFT_METHOD_DOP853, a.pot.kind, opt.block_size, opt.schedule, false,
                            block, packed)) != DRIFT_OK)
        return rc;
    PlainKernel k =
        kernel_for(DRIFT_METHOD_DOP853, a.pot.kind, opt.precision, opt.schedule, packed,
                   wide_plain(n, d->cu_count), lds_times(nt));
    int64_t blocks = 0;
    if ((rc = integrator_grid(d->cu_count, k, opt.schedule, block, n, blocks)) != DRIFT_OK)
        return rc;
    if (blocks > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n too large");
    if (opt.schedule == DRIFT_SCHEDULE_PERSISTENT) {
        a.queue = reinterpret_cast<unsigned long long*>(base + kNaQueue);
        // a fill kernel, not a memset, so the sequence may be captured
        hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, st, a.queue);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    if (cost_order) {
        uint32_t* slots = reinterpret_cast<uint32_t*>(base + kNaSlots);
        if ((rc = launch_cost_order(a, opt.precision, opt.schedule,
                                    reinterpret_cast<uint16_t*>(base + kNaKeys),
                                    reinterpret_cast<uint32_t*>(base + kNaBins), slots, st)) !=
            DRIFT_OK)
            return rc;
        a.order = slots;
    }
    if (packed && a.stats) {
        hipLaunchKernelGGL(zero_status_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           a.stats, (int64_t)n);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(block), 0, st, a);
    DRIFT_HIP_TRY(hipGetLastError());
    return DRIFT_OK;
}

}  // namespace
}  // namespace drift

using namespace drift;

extern "C" {

int drift_get_abi_version(void) { return DRIFT_ABI_VERSION; }

const char* drift_last_error(void) { return g_last_error.c_str(); }

int drift_session_create(int device, drift_session** out) {
    if (!out) return fail(DRIFT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int count = 0;
    DRIFT_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "device " + std::to_string(device) + " out of range (" + std::to_string(count) +
                        " devices)");
    DRIFT_HIP_TRY(hipSetDevice(device));
    drift_session* s = new drift_session();
    s->device = device;
    if (hipDeviceGetAttribute(&s->cu_count, hipDeviceAttributeMultiprocessorCount, device) !=
        hipSuccess)
        s->cu_count = 256;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return fail(DRIFT_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e));
    }
    *out = s;
    return DRIFT_OK;
}

int drift_session_destroy(drift_session* s) {
    if (!s) return DRIFT_OK;
    (void)hipSetDevice(s->device);
    // launches on the caller's streams may still use the session's buffers
    (void)hipDeviceSynchronize();
    for (auto& t : s->tables) {
        (void)hipFree(t.d_table64);
        (void)hipFree(t.d_table32);
        (void)hipFree(t.d_far);
    }
    for (auto& o : s->origin_sets) {
        (void)hipFree(o.d_tracks);
        (void)hipFree(o.d_params);
    }
    for (DeviceBuffer* b : {&s->state0, &s->times, &s->out, &s->stats, &s->order_keys,
                            &s->order_slots, &s->order_bins, &s->tune_out, &s->tune_stats,
                            &s->staging, &s->force_io, &s->resume})
        if (b->ptr) (void)hipFree(b->ptr);
    for (hipEvent_t e : s->phase_events)
        if (e) (void)hipEventDestroy(e);
    if (s->order_done) (void)hipEventDestroy(s->order_done);
    if (s->staging_done) (void)hipEventDestroy(s->staging_done);
    if (s->resume_done) (void)hipEventDestroy(s->resume_done);
    if (s->queues) (void)hipFree(s->queues);
    for (auto& r : s->ragged) ragged_scratch_free(r);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return DRIFT_OK;
}

int drift_galpy_dopr54(drift_session* session, const double* state0, const double* times,
                       double* state_out, size_t n, size_t nt, double rtol, double atol,
                       double dt_one_init, const drift_potential* potential,
                       drift_orbit_stats* stats) {
    return host_launch(session, DRIFT_METHOD_DOPR54, state0, times, state_out, n, nt, rtol, atol,
                       dt_one_init, potential, stats);
}

int drift_galpy_dop853(drift_session* session, const double* state0, const double* times,
                       double* state_out, size_t n, size_t nt, double rtol, double atol,
                       const drift_potential* potential, drift_orbit_stats* stats) {
    return host_launch(session, DRIFT_METHOD_DOP853, state0, times, state_out, n, nt, rtol, atol,
                       dp54::initial_step_sentinel, potential, stats);
}

int drift_galpy_dopr54_device(drift_session* session, const double* d_state0,
                              const double* d_times, double* d_state_out, size_t n, size_t nt,
                              double rtol, double atol, double dt_one_init,
                              const drift_potential* potential, drift_orbit_stats* d_stats,
                              const drift_launch_options* options, void* stream) {
    return launch_integrator(session, DRIFT_METHOD_DOPR54, d_state0, d_times, d_state_out, n, nt,
                             rtol, atol, dt_one_init, potential, d_stats, options,
                             (hipStream_t)stream);
}

int drift_galpy_dop853_device(drift_session* session, const double* d_state0,
                              const double* d_times, double* d_state_out, size_t n, size_t nt,
                              double rtol, double atol, const drift_potential* potential,
                              drift_orbit_stats* d_stats, const drift_launch_options* options,
                              void* stream) {
    return launch_integrator(session, DRIFT_METHOD_DOP853, d_state0, d_times, d_state_out, n, nt,
                             rtol, atol, dp54::initial_step_sentinel, potential, d_stats, options,
                             (hipStream_t)stream);
}

int drift_galpy_dop853_device_noalloc(const double* d_state0, const double* d_times,
                                      double* d_state_out, size_t n, size_t nt, double rtol,
                                      double atol, const drift_potential* potential,
                                      drift_orbit_stats* d_stats,
                                      const drift_launch_options* options, void* stream) {
    return noalloc_dop853(d_state0, d_times, d_state_out, n, nt, rtol, atol, potential, d_stats,
                          options, (hipStream_t)stream);
}

int drift_launch_shape(drift_session* s, int32_t method, const drift_potential* potential,
                       size_t n, const drift_launch_options* options, int32_t* grid,
                       int32_t* block) {
    if (!s || !grid || !block) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (method != DRIFT_METHOD_DOPR54 && method != DRIFT_METHOD_DOP853)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "unknown method");
    if (n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n must be >= 1");
    drift_potential p{};
    p.kind = POT_GALPY_KEPLER;
    if (potential) p = *potential;
    int rc = validate_potential(p);
    if (rc != DRIFT_OK) return rc;
    int kind = p.kind;
    if (p.origins != 0 || p.kind == POT_NONE) kind |= POT_MOVING;
    drift_launch_options opt{};
    if (options)
        opt = *options;
    else
        opt.schedule = default_schedule(method);  // what the host-buffer entry points launch
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    if (opt.schedule < DRIFT_SCHEDULE_PLAIN || opt.schedule > DRIFT_SCHEDULE_TUNED)
        return fail(DRIFT_ERR_UNSUPPORTED, "schedule must be PLAIN, PERSISTENT or TUNED");
    if ((kind & POT_MOVING) && opt.precision == DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "moving origins and the NONE potential need FP64 precision");
    if (opt.schedule == DRIFT_SCHEDULE_TUNED) {
        const TunedEntry* e = find_tuned(s, method, kind, opt.precision, n);
        opt.schedule = e ? e->options.schedule : default_schedule(method);
        opt.block_size = e ? e->options.block_size : 0;
    }
    int b = 0;
    bool packed = false;
    if ((rc = resolve_block(method, kind, opt.block_size, opt.schedule,
                            opt.staging == DRIFT_STAGING_STAGED, b, packed)) != DRIFT_OK)
        return rc;
    DRIFT_HIP_TRY(hipSetDevice(s->device));
    int64_t blocks = 0;
    // (the grid does not depend on where the kernel keeps the time grid: an
    // LDS copy leaves every kernel's resident blocks per CU unchanged)
    rc = integrator_grid(s->cu_count,
                         kernel_for(method, kind, opt.precision, opt.schedule, packed,
                                    wide_plain(n, s->cu_count), false),
                         opt.schedule, b, n, blocks);
    if (rc != DRIFT_OK) return rc;
    if (blocks > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n too large");
    *grid = (int32_t)blocks;
    *block = b;
    return DRIFT_OK;
}

int drift_session_last_launch(drift_session* s, drift_launch_report* report) {
    if (!s || !report) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!s->have_last_launch)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "the session has made no host-buffer launch");
    *report = s->last_launch;
    return DRIFT_OK;
}

int drift_tune_device(drift_session* s, int32_t method, const double* d_state0,
                      const double* d_times, size_t n, size_t nt, double rtol, double atol,
                      double dt_one_init, const drift_potential* potential, size_t prefix_samples,
                      drift_launch_options* best, double* best_ms, void* stream) {
    if (!s) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (!best) return fail(DRIFT_ERR_INVALID_ARGUMENT, "best is NULL");
    if (method != DRIFT_METHOD_DOPR54 && method != DRIFT_METHOD_DOP853)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "unknown method");
    if (n < 1 || nt < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "n must be >= 1 and nt >= 2");
    const size_t k = prefix_samples ? std::min(prefix_samples, nt) : std::max<size_t>(2, nt / 20);
    if (k < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "prefix_samples must be >= 2");
    DRIFT_HIP_TRY(hipSetDevice(s->device));
    int rc;
    if ((rc = ensure_buffer(s->tune_out, k * n * DIM * sizeof(double))) != DRIFT_OK) return rc;
    if ((rc = ensure_buffer(s->tune_stats, n * sizeof(drift_orbit_stats))) != DRIFT_OK) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : s->stream;
    drift_launch_options base{};
    base.precision = best->precision;
    base.max_attempts = best->max_attempts;
    auto trial = [&](const drift_launch_options& o) {
        return launch_integrator(s, method, d_state0, d_times, (double*)s->tune_out.ptr, n, k, rtol,
                                 atol, dt_one_init, potential,
                                 (drift_orbit_stats*)s->tune_stats.ptr, &o, st);
    };
    // untimed launches of both schedules: resolve the potential's table and
    // load the code objects before anything is timed
    if ((rc = trial(base)) != DRIFT_OK) return rc;
    drift_launch_options warm = base;
    warm.schedule = DRIFT_SCHEDULE_PERSISTENT;
    warm.ordering = DRIFT_ORDER_COST;
    if ((rc = trial(warm)) != DRIFT_OK) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DRIFT_HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(DRIFT_ERR_DEVICE, "hipEventCreate failed");
    }
    drift_launch_options chosen = base, best_inline = base;
    float chosen_ms = 0.0f, inline_ms = 0.0f;
    bool have = false, have_inline = false;
    // A trial covers only a prefix of the time grid, where the plain
    // schedule's tail (each wavefront waits for its slowest orbit, the grid
    // for its slowest block) is short; over the whole grid the persistent
    // DOP853 schedule wins whenever its lanes integrate several orbits each.
    // So once n reaches twice the resident lanes (2 waves of 64 per SIMD, 4
    // SIMDs per CU) only persistent DOP853 launches are tried; DOPR54, whose
    // plain schedule is its default, tries both at every size.
    const size_t resident_lanes = (size_t)std::max(s->cu_count, 1) * 4 * 2 * 64;
    const bool try_plain = n < 2 * resident_lanes || method == DRIFT_METHOD_DOPR54;
    // packed dense output: DOP853 with a static potential (no moving origins)
    const bool packable = method == DRIFT_METHOD_DOP853 &&
                          !(potential && (potential->origins != 0 ||
                                          potential->kind == DRIFT_POTENTIAL_NONE));
    for (int schedule : {DRIFT_SCHEDULE_PLAIN, DRIFT_SCHEDULE_PERSISTENT}) {
        if (schedule == DRIFT_SCHEDULE_PLAIN && !try_plain) continue;
        for (int block : {64, 128, 256, kPackedBlock}) {
            if (block == kPackedBlock && !packable) continue;
            for (int ordering : {DRIFT_ORDER_INDEX, DRIFT_ORDER_COST}) {
                drift_launch_options o = base;
                o.schedule = schedule;
                o.block_size = block;
                o.ordering = ordering;
                float ms = 0.0f;
                if (hipEventRecord(e0, st) != hipSuccess) rc = DRIFT_ERR_DEVICE;
                if (rc == DRIFT_OK) rc = trial(o);
                if (rc == DRIFT_OK && hipEventRecord(e1, st) != hipSuccess) rc = DRIFT_ERR_DEVICE;
                if (rc == DRIFT_OK && hipEventSynchronize(e1) != hipSuccess) rc = DRIFT_ERR_DEVICE;
                if (rc == DRIFT_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    rc = DRIFT_ERR_DEVICE;
                if (rc != DRIFT_OK) {
                    (void)hipEventDestroy(e0);
                    (void)hipEventDestroy(e1);
                    if (rc == DRIFT_ERR_DEVICE) set_last_error("autotuner: event timing failed");
                    return rc;
                }
                if (!have || ms < chosen_ms) {
                    chosen = o;
                    chosen_ms = ms;
                    have = true;
                }
                if (block != kPackedBlock && (!have_inline || ms < inline_ms)) {
                    best_inline = o;
                    inline_ms = ms;
                    have_inline = true;
                }
            }
        }
    }
    // Output samples per accepted step over the prefix (every trial takes the
    // same steps per orbit).  The prefix understates what direct output costs
    // over the whole grid: there the lanes' orbits still sit near the same
    // sample, so their scattered records still fill lines together; later
    // they drift apart and each record leaves L2 as a partial line.  Where a
    // step emits several samples that cost dominates (1M Milky-Way orbits,
    // DOP853: 3.1 samples per step at rtol 1e-7, staged 59 ms against 69 ms
    // direct; 0.74 at rtol 1e-12, direct 120 ms against 119 + 19 ms staged),
    // so from kStagedSamplesPerStep on staged output is chosen whatever the
    // prefix timings say.
    constexpr double kStagedSamplesPerStep = 1.5;
    double samples_per_step = 0.0;
    if (rc == DRIFT_OK && have) {
        std::vector<drift_orbit_stats> stats(n);
        if (hipMemcpyAsync(stats.data(), s->tune_stats.ptr, n * sizeof(drift_orbit_stats),
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return fail(DRIFT_ERR_DEVICE, "autotuner: reading the trial statistics failed");
        }
        uint64_t accepted = 0;
        for (const drift_orbit_stats& r : stats) accepted += r.accepted;
        samples_per_step = accepted ? (double)(n * (k - 1)) / (double)accepted : 0.0;
    }
    // then the fastest configuration with inline dense output, with staged
    // output (packing gains nothing there: the staged writes cost the same
    // whichever lane makes them, and packing costs registers)
    if (rc == DRIFT_OK && have) {
        drift_launch_options o = have_inline ? best_inline : chosen;
        o.staging = DRIFT_STAGING_STAGED;
        float ms = 0.0f;
        if ((rc = trial(o)) == DRIFT_OK) {  // untimed: allocates the staging buffer
            if (hipEventRecord(e0, st) != hipSuccess) rc = DRIFT_ERR_DEVICE;
            if (rc == DRIFT_OK) rc = trial(o);
            if (rc == DRIFT_OK && hipEventRecord(e1, st) != hipSuccess) rc = DRIFT_ERR_DEVICE;
            if (rc == DRIFT_OK && hipEventSynchronize(e1) != hipSuccess) rc = DRIFT_ERR_DEVICE;
            if (rc == DRIFT_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                rc = DRIFT_ERR_DEVICE;
        }
        if (rc != DRIFT_OK) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc == DRIFT_ERR_DEVICE) set_last_error("autotuner: event timing failed");
            return rc;
        }
        if (ms < chosen_ms || samples_per_step >= kStagedSamplesPerStep) {
            chosen = o;
            chosen_ms = ms;
        } else {
            chosen.staging = DRIFT_STAGING_DIRECT;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    PotParams pp{};
    const void *t64 = nullptr, *t32 = nullptr;
    if ((rc = resolve_potential(s, potential, pp, &t64, &t32)) != DRIFT_OK) return rc;
    TunedEntry entry{method, pp.kind, base.precision, floor_log2(n), chosen};
    bool replaced = false;
    for (TunedEntry& e : s->tuned) {
        if (e.method == entry.method && e.kind == entry.kind && e.precision == entry.precision &&
            e.log2n == entry.log2n) {
            e = entry;
            replaced = true;
        }
    }
    if (!replaced) s->tuned.push_back(entry);
    *best = chosen;
    if (best_ms) *best_ms = chosen_ms;
    return DRIFT_OK;
}

int drift_origins_create(drift_session* session, const double* tracks, const double* params,
                         int32_t n_origins, int32_t n_segments, double t_start, double t_end,
                         int32_t* origins_id) {
    if (!session || !tracks || !params || !origins_id)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n_origins < 1 || n_segments < 1)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "n_origins and n_segments must be >= 1");
    if (!std::isfinite(t_start) || !std::isfinite(t_end) || !(t_end > t_start) || !std::isfinite(t_end - t_start))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "the track span must be finite with t_end > t_start");
    const size_t ncoef = (size_t)n_origins * (size_t)n_segments * ORIGIN_COEFS;
    for (size_t k = 0; k < ncoef; ++k)
        if (!std::isfinite(tracks[k]))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "track coefficients must be finite");
    std::vector<double> pb(2 * (size_t)n_origins);
    for (int32_t i = 0; i < n_origins; ++i) {
        const double amp = params[2 * i], b = params[2 * i + 1];
        if (!std::isfinite(amp) || !std::isfinite(b) || b < 0.0)
            return fail(DRIFT_ERR_INVALID_ARGUMENT,
                        "origin amp must be finite and b finite and >= 0");
        pb[2 * i] = amp;
        pb[2 * i + 1] = b * b;
    }
    DRIFT_HIP_TRY(hipSetDevice(session->device));
    OriginSet o;
    o.n_origins = n_origins;
    o.n_segments = n_segments;
    o.t_start = t_start;
    o.t_end = t_end;
    hipError_t err = hipMalloc(&o.d_tracks, ncoef * sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&o.d_params, pb.size() * sizeof(double));
    if (err == hipSuccess)
        err = hipMemcpy(o.d_tracks, tracks, ncoef * sizeof(double), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = hipMemcpy(o.d_params, pb.data(), pb.size() * sizeof(double), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        if (o.d_tracks) (void)hipFree(o.d_tracks);
        if (o.d_params) (void)hipFree(o.d_params);
        set_last_error(std::string("HIP error: ") + hipGetErrorString(err) +
                       " uploading moving-origin tracks");
        return err == hipErrorOutOfMemory ? DRIFT_ERR_OUT_OF_MEMORY : DRIFT_ERR_DEVICE;
    }
    o.id = session->next_origin_id++;
    session->origin_sets.push_back(o);
    *origins_id = o.id;
    return DRIFT_OK;
}

int drift_origins_destroy(drift_session* session, int32_t origins_id) {
    if (!session) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    for (size_t k = 0; k < session->origin_sets.size(); ++k) {
        OriginSet& o = session->origin_sets[k];
        if (o.id != origins_id) continue;
        DRIFT_HIP_TRY(hipSetDevice(session->device));
        // launches already enqueued on any stream may still read the set
        DRIFT_HIP_TRY(hipDeviceSynchronize());
        DRIFT_HIP_TRY(hipFree(o.d_tracks));
        DRIFT_HIP_TRY(hipFree(o.d_params));
        session->origin_sets.erase(session->origin_sets.begin() + (ptrdiff_t)k);
        return DRIFT_OK;
    }
    return fail(DRIFT_ERR_INVALID_ARGUMENT,
                "no moving-origin set " + std::to_string(origins_id) + " in this session");
}

int drift_force_device(drift_session* session, const drift_potential* potential,
                       const double* d_pos, const double* d_t, double* d_acc, size_t n,
                       int32_t precision, void* stream) {
    if (!session) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (!d_pos || !d_acc || n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "bad force arguments");
    if (precision != DRIFT_PRECISION_FP64 && precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    DRIFT_HIP_TRY(hipSetDevice(session->device));
    ForceArgs a{};
    a.pos = d_pos;
    a.t = d_t;
    a.acc = d_acc;
    a.n = (int64_t)n;
    const void *t64 = nullptr, *t32 = nullptr;
    int rc = resolve_potential(session, potential, a.pot, &t64, &t32);
    if (rc != DRIFT_OK) return rc;
    const bool mixed = precision == DRIFT_PRECISION_MIXED;
    if ((a.pot.kind & POT_MOVING) && mixed)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "moving origins and the NONE potential need FP64 precision");
    a.table = mixed ? t32 : t64;
    using FK = void (*)(ForceArgs);
    FK k;
    switch (a.pot.kind) {
        case POT_GALPY_KEPLER:
            k = mixed ? force_kernel<POT_GALPY_KEPLER, float> : force_kernel<POT_GALPY_KEPLER, double>;
            break;
        case POT_KEPLER:
            k = mixed ? force_kernel<POT_KEPLER, float> : force_kernel<POT_KEPLER, double>;
            break;
        case POT_PLUMMER:
            k = mixed ? force_kernel<POT_PLUMMER, float> : force_kernel<POT_PLUMMER, double>;
            break;
        case POT_MW:
            k = mixed ? force_kernel<POT_MW, float> : force_kernel<POT_MW, double>;
            break;
        case POT_MOVING | POT_KEPLER: k = force_kernel<POT_MOVING | POT_KEPLER, double>; break;
        case POT_MOVING | POT_PLUMMER: k = force_kernel<POT_MOVING | POT_PLUMMER, double>; break;
        case POT_MOVING | POT_MW: k = force_kernel<POT_MOVING | POT_MW, double>; break;
        default: k = force_kernel<POT_MOVING | POT_NONE, double>; break;
    }
    const int64_t blocks = ((int64_t)n + kDefaultBlock - 1) / kDefaultBlock;
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(kDefaultBlock), 0,
                       stream ? (hipStream_t)stream : session->stream, a);
    DRIFT_HIP_TRY(hipGetLastError());
    return DRIFT_OK;
}

int drift_force(drift_session* session, const drift_potential* potential, const double* pos,
                const double* t, double* acc, size_t n) {
    if (!session) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (!pos || !acc || n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "bad force arguments");
    DRIFT_HIP_TRY(hipSetDevice(session->device));
    const size_t vec = n * 3 * sizeof(double), times = t ? n * sizeof(double) : 0;
    int rc = ensure_buffer(session->force_io, 2 * vec + times);
    if (rc != DRIFT_OK) return rc;
    double* d_pos = static_cast<double*>(session->force_io.ptr);
    double* d_acc = d_pos + 3 * n;
    double* d_t = t ? d_acc + 3 * n : nullptr;
    hipStream_t st = session->stream;
    DRIFT_HIP_TRY(hipMemcpyAsync(d_pos, pos, vec, hipMemcpyHostToDevice, st));
    if (t) DRIFT_HIP_TRY(hipMemcpyAsync(d_t, t, times, hipMemcpyHostToDevice, st));
    rc = drift_force_device(session, potential, d_pos, d_t, d_acc, n, DRIFT_PRECISION_FP64, st);
    if (rc != DRIFT_OK) return rc;
    DRIFT_HIP_TRY(hipMemcpyAsync(acc, d_acc, vec, hipMemcpyDeviceToHost, st));
    DRIFT_HIP_TRY(hipStreamSynchronize(st));
    return DRIFT_OK;
}

int drift_mw_table_check(const drift_potential* potential, int32_t samples_per_segment,
                         double* max_rel_error) {
    if (!potential || !max_rel_error || samples_per_segment < 1)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "bad table-check arguments");
    if (potential->kind != POT_MW) return fail(DRIFT_ERR_INVALID_ARGUMENT, "potential is not MW");
    int rc = validate_potential(*potential);
    if (rc != DRIFT_OK) return rc;
    std::vector<double> coeffs(TABLE_DOUBLES), far(FAR_DOUBLES);
    build_mw_table(*potential, coeffs.data());
    build_mw_far_table(*potential, far.data());
    // a lane outside the LDS table's range reads the far table: on the
    // octaves the two share, the coefficients must be the same bits, so that
    // the range only decides where a lane reads, never what it computes
    const size_t shift = (size_t)(TABLE_E_MIN - FAR_E_MIN) << TABLE_SUBBITS;
    if (std::memcmp(coeffs.data(), far.data() + shift * TABLE_NCOEF,
                    sizeof(double) * TABLE_DOUBLES) != 0)
        return fail(DRIFT_ERR_INVALID_ARGUMENT,
                    "the LDS and far Milky-Way tables differ on a segment they share");
    *max_rel_error = std::max(
        mw_table_max_rel_error(*potential, coeffs.data(), samples_per_segment),
        mw_far_table_max_rel_error(*potential, far.data(), samples_per_segment));
    return DRIFT_OK;
}

}  // extern "C"
