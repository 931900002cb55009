// ragged.hip -- ragged batches: many independent jobs (different particle
// counts, potentials, time grids and tolerances) in one persistent launch per
// integration method, with resumable per-orbit state for chunked execution.
//
// Work items are (job, orbit) pairs numbered job by job.  Lanes claim items
// from a global queue (wavefront-ballot aggregation, as the persistent plain
// schedule), locate their job by binary search over the item offsets, and
// integrate the orbit until it has emitted the job's stop sample or finished.
// A paused orbit's complete integrator state is saved to a per-job resident
// buffer (SoA: word w of orbit i at resident[w * n + i]) and restored bit for
// bit by the next chunk, so chunking never changes a result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "drift_internal.hpp"
#include "forces.hpp"
#include "integrators.hpp"
#include "ragged.hpp"

namespace drift {

// ---------------------------------------------------------------------------
// Device side
// ---------------------------------------------------------------------------

// Force of a lane's current job, dispatched at run time (jobs in one launch
// may use different potentials).  c0..c2 are the kind's constants.  The MW
// table of the launch's most used MW potential is staged in LDS (`lds`,
// in_lds set for its jobs); other MW potentials read theirs from global
// memory.
//
// Jobs are grouped by potential kind, so nearly every wavefront carries one
// kind.  set_wave_force records, after every claim, that kind when all the
// wavefront's active lanes share it (wave_kind, a wavefront-uniform value, -1
// otherwise) and whether all their MW tables are the LDS one (wave_lds); the
// kernels then run the whole attempt with that kind's force type
// (attempt_by_kind), the code of the single-job kernels, and only a wavefront
// that straddles two groups runs the per-lane dispatch of operator().
template <class FT>
struct LaneForce {
    int32_t kind;
    int32_t wave_kind;  // the wavefront's common kind, or -1
    int32_t wave_lds;   // 1 when every active lane's MW table is the LDS one
    bool in_lds;
    FT c0, c1, c2;
    const FT* table;
    const FT* lds;
    const PotParams* p;  // far table (MW outside the LDS range)
    __device__ __forceinline__ void operator()(double t, const double* q, double* a) const {
        switch (kind) {
            case POT_GALPY_KEPLER: GalpyKeplerForce<FT>{}(t, q, a); break;
            case POT_KEPLER: KeplerForce<FT>{c0}(t, q, a); break;
            case POT_PLUMMER: PlummerForce<FT>{c0, c1}(t, q, a); break;
            case POT_NONE: NoForce{}(t, q, a); break;
            default:
                if (in_lds)
                    MwForce<FT>{p, lds, c0, c1, c2}(t, q, a);
                else
                    MwForce<FT>{p, table, c0, c1, c2}(t, q, a);
                break;
        }
    }
};

// Records the wavefront's common kind and table (see LaneForce).
template <class FT>
__device__ __forceinline__ void set_wave_force(LaneForce<FT>& f) {
    const int32_t first = __builtin_amdgcn_readfirstlane(f.kind);
    f.wave_kind = __ballot(f.kind != first) == 0 ? first : -1;
    f.wave_lds = __ballot(!f.in_lds) == 0 ? 1 : 0;
}

// One attempt with the force type of the wavefront's common kind (scalar
// branches: wave_kind and wave_lds are wavefront-uniform), or with the
// per-lane dispatch.  step(force) runs the attempt.
template <class FT, class Step>
__device__ __forceinline__ bool attempt_by_kind(const LaneForce<FT>& f, const Step& step) {
    if (f.wave_kind == POT_MW && f.wave_lds) return step(MwForce<FT>{f.p, f.lds, f.c0, f.c1, f.c2});
    if (f.wave_kind == POT_KEPLER) return step(KeplerForce<FT>{f.c0});
    if (f.wave_kind == POT_PLUMMER) return step(PlummerForce<FT>{f.c0, f.c1});
    if (f.wave_kind == POT_GALPY_KEPLER) return step(GalpyKeplerForce<FT>{});
    return step(f);
}

// Sets f to job potential `pot` (f.lds stays the block's LDS table).
template <class FT>
__device__ __forceinline__ void set_lane_force(LaneForce<FT>& f, const RaggedArgs& a, int32_t pot) {
    const PotParams* p = a.pots + pot;
    f.kind = p->kind;
    f.in_lds = pot == a.lds_pot;
    f.p = p;
    if (p->kind == POT_MW) {
        f.c0 = FT(p->disk_amp);
        f.c1 = FT(p->disk_a);
        f.c2 = FT(p->disk_b2);
    } else {
        f.c0 = FT(p->amp);
        f.c1 = FT(p->b2);
        f.c2 = FT(0);
    }
    if constexpr (sizeof(FT) == 8)
        f.table = static_cast<const FT*>(a.tables64[pot]);
    else
        f.table = static_cast<const FT*>(a.tables32[pot]);
}

// Stages the launch's LDS table (a no-op without one) and returns a lane force
// bound to it.
template <class FT>
__device__ __forceinline__ LaneForce<FT> stage_lane_table(const RaggedArgs& a, FT* lds) {
    if (a.lds_pot >= 0) {
        constexpr int words = sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS;
        const FT* g = static_cast<const FT*>(a.lds_src);
        for (int k = threadIdx.x; k < words; k += blockDim.x) lds[k] = g[k];
        __syncthreads();
    }
    LaneForce<FT> f{};
    f.lds = lds;
    return f;
}

__device__ __forceinline__ int32_t find_job(const RaggedJob* jobs, int32_t njobs,
                                            unsigned long long slot) {
    int32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if ((unsigned long long)jobs[mid].item_begin <= slot) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double pack2(int32_t lo, int32_t hi) {
    return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) |
                                            (unsigned long long)(uint32_t)lo));
}
__device__ __forceinline__ void unpack2(double w, int32_t& lo, int32_t& hi) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(w);
    lo = (int32_t)(uint32_t)(b & 0xffffffffull);
    hi = (int32_t)(uint32_t)(b >> 32);
}

struct Resident {
    double* base;
    int64_t n;
    int64_t i;
    __device__ __forceinline__ double& w(int k) const { return base[(int64_t)k * n + i]; }
};

__device__ __forceinline__ void save(const Resident& r, const Dop853Lane& L, int64_t phase) {
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        r.w(k) = L.y[k];
        if (k < 3) r.w(6 + k) = L.F1[k];
    }
    r.w(12) = L.h;
    r.w(13) = L.t_cur;
    // word 14 is unused (the lane keeps no t_old)
    r.w(15) = pack2(L.ii, L.reject);
    r.w(16) = pack2((int32_t)L.accepted, (int32_t)L.rejected);
    r.w(17) = pack2((int32_t)L.dense, L.status);
    r.w(kResidentPhaseWord) = __longlong_as_double(phase);
}

__device__ __forceinline__ void restore(const Resident& r, Dop853Lane& L) {
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        L.y[k] = r.w(k);
        if (k < 3) L.F1[k] = r.w(6 + k);
    }
    L.h = r.w(12);
    L.t_cur = r.w(13);
    int32_t a, b;
    unpack2(r.w(15), L.ii, L.reject);
    unpack2(r.w(16), a, b);
    L.accepted = (uint32_t)a;
    L.rejected = (uint32_t)b;
    unpack2(r.w(17), a, L.status);
    L.dense = (uint32_t)a;
}

__device__ __forceinline__ void save(const Resident& r, const Dopr54Lane& L, int64_t phase) {
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        r.w(k) = L.y[k];
        r.w(6 + k) = L.a1[k];
    }
    r.w(12) = L.to;
    r.w(13) = L.dt_one;
    r.w(14) = L.init_to;
    r.w(15) = pack2(L.ii, L.err);
    r.w(16) = pack2((int32_t)L.accepted, (int32_t)L.rejected);
    r.w(17) = L.init_dt_one;
    r.w(18) = pack2(L.status, 0);
    r.w(kResidentPhaseWord) = __longlong_as_double(phase);
}

__device__ __forceinline__ void restore(const Resident& r, Dopr54Lane& L) {
#pragma unroll
    for (int k = 0; k < DIM; k++) {
        L.y[k] = r.w(k);
        L.a1[k] = r.w(6 + k);
    }
    L.to = r.w(12);
    L.dt_one = r.w(13);
    L.init_to = r.w(14);
    int32_t a, b;
    unpack2(r.w(15), L.ii, L.err);
    unpack2(r.w(16), a, b);
    L.accepted = (uint32_t)a;
    L.rejected = (uint32_t)b;
    L.init_dt_one = r.w(17);
    unpack2(r.w(18), L.status, a);
}

struct OutSinkR {
    double* out;
    int64_t n;
    int64_t idx;
    __device__ __forceinline__ void emit(int32_t s, const double* y) const {
        double2* p = reinterpret_cast<double2*>(out + ((int64_t)s * n + idx) * DIM);
        p[0] = make_double2(y[0], y[1]);
        p[1] = make_double2(y[2], y[3]);
        p[2] = make_double2(y[4], y[5]);
    }
    __device__ __forceinline__ void emit_component(int32_t s, int c, double v) const {
        out[((int64_t)s * n + idx) * DIM + c] = v;
    }
};

__device__ __forceinline__ bool claim_item(bool need, unsigned long long* queue,
                                           unsigned long long items,
                                           unsigned long long& slot) {
    const unsigned long long want = __ballot(need);
    if (want == 0) return true;
    const int lane = __lane_id();
    const int leader = __ffsll(want) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(queue, (unsigned long long)__popcll(want));
    base = __shfl(base, leader);
    if (!need) return true;
    const unsigned long long below = lane == 0 ? 0ull : (want & (~0ull >> (64 - lane)));
    slot = base + (unsigned long long)__popcll(below);
    return slot < items;
}

// Two wavefronts per SIMD, as for the single-job kernels: the register budget
// drops to 256 and the latency of one wave's force loads hides behind the other.
#define DRIFT_RAGGED_WAVES __attribute__((amdgpu_waves_per_eu(2, 2)))

template <class FT>
__global__ void __launch_bounds__(256) DRIFT_RAGGED_WAVES ragged_dop853_kernel(RaggedArgs a) {
    __shared__ FT lds[sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS];
    LaneForce<FT> f = stage_lane_table<FT>(a, lds);
    Dop853Lane L;
    Dop853Job J;
    OutSinkR sink{nullptr, 0, 0};
    Resident res{nullptr, 0, 0};
    drift_orbit_stats* stats = nullptr;
    unsigned int* done_count = nullptr;
    int32_t stop = 0;
    bool need = true;
    for (;;) {
        unsigned long long slot = 0;
        if (!claim_item(need, a.queue, (unsigned long long)a.items, slot)) break;
        if (need) {
            const RaggedJob& d = a.jobs[find_job(a.jobs, a.njobs, slot)];
            const int64_t orbit = a.order ? (int64_t)a.order[slot]
                                          : (int64_t)(slot - (unsigned long long)d.item_begin);
            res = Resident{d.resident, d.n, orbit};
            const int64_t phase = d.resident ? __double_as_longlong(res.w(kResidentPhaseWord))
                                             : kOrbitNew;
            if (phase == kOrbitDone) continue;
            J = make_dop853_job(d.times, d.nt, d.rtol, d.atol, a.max_attempts, d.times[0],
                                d.times[d.nt - 1]);
            set_lane_force<FT>(f, a, d.pot);
            f.wave_kind = -1;  // the init below dispatches per lane
            sink = OutSinkR{d.out, d.n, orbit};
            stats = d.stats;
            done_count = d.done_count;
            stop = d.stop_sample;
            if (phase == kOrbitRunning) {
                restore(res, L);
                dop853_refresh_next(L, J);
            } else {
                double y0[DIM];
                const double2* p = reinterpret_cast<const double2*>(d.state0 + orbit * DIM);
                double2 u = p[0], v = p[1], w = p[2];
                y0[0] = u.x; y0[1] = u.y; y0[2] = v.x; y0[3] = v.y; y0[4] = w.x; y0[5] = w.y;
                dop853_init(L, y0, J.times[0], J, f, sink);
            }
            need = false;
        }
        set_wave_force(f);
        const bool done = attempt_by_kind(
            f, [&](const auto& force) { return dop853_attempt(L, J, force, sink); });
        if (done || L.ii >= stop) {
            if (res.base) save(res, L, done ? kOrbitDone : kOrbitRunning);
            if (done) {
                if (stats) {
                    drift_orbit_stats st;
                    st.accepted = L.accepted;
                    st.rejected = L.rejected;
                    st.dense_steps = L.dense;
                    st.status = L.status;
                    stats[sink.idx] = st;
                }
                if (done_count) atomicAdd(done_count, 1u);
            }
            need = true;
        }
    }
}

template <class FT>
__global__ void __launch_bounds__(256) DRIFT_RAGGED_WAVES ragged_dopr54_kernel(RaggedArgs a) {
    __shared__ FT lds[sizeof(FT) == 8 ? TABLE_DOUBLES : TABLE32_FLOATS];
    LaneForce<FT> f = stage_lane_table<FT>(a, lds);
    Dopr54Lane L;
    Dopr54Job J;
    OutSinkR sink{nullptr, 0, 0};
    Resident res{nullptr, 0, 0};
    drift_orbit_stats* stats = nullptr;
    unsigned int* done_count = nullptr;
    int32_t stop = 0;
    bool need = true;
#pragma unroll
    for (int k = 0; k < 6; k++) J.bound2[k] = a.bound2[k];
    for (;;) {
        unsigned long long slot = 0;
        if (!claim_item(need, a.queue, (unsigned long long)a.items, slot)) break;
        bool done = false;
        if (need) {
            const RaggedJob& d = a.jobs[find_job(a.jobs, a.njobs, slot)];
            const int64_t orbit = a.order ? (int64_t)a.order[slot]
                                          : (int64_t)(slot - (unsigned long long)d.item_begin);
            res = Resident{d.resident, d.n, orbit};
            const int64_t phase = d.resident ? __double_as_longlong(res.w(kResidentPhaseWord))
                                             : kOrbitNew;
            if (phase == kOrbitDone) continue;
            J.times = d.times;
            J.nt = d.nt;
            J.max_attempts = a.max_attempts;
            J.dt = d.times[1] - d.times[0];
            J.rtol = d.rtol;
            J.atol = d.atol;
            J.dt_one_init = d.dt_one_init;
            set_lane_force<FT>(f, a, d.pot);
            f.wave_kind = -1;  // the init below dispatches per lane
            sink = OutSinkR{d.out, d.n, orbit};
            stats = d.stats;
            done_count = d.done_count;
            stop = d.stop_sample;
            if (phase == kOrbitRunning) {
                restore(res, L);
            } else {
                double y0[DIM];
                const double2* p = reinterpret_cast<const double2*>(d.state0 + orbit * DIM);
                double2 u = p[0], v = p[1], w = p[2];
                y0[0] = u.x; y0[1] = u.y; y0[2] = v.x; y0[3] = v.y; y0[4] = w.x; y0[5] = w.y;
                // the RK4 estimate ran in ragged_dopr54_estimate_kernel
                const double dt_one = J.dt_one_init == dp54::initial_step_sentinel
                                          ? a.dt_est[slot]
                                          : J.dt_one_init;
                dopr54_begin(L, y0, d.times[0], dt_one, J, f, sink);
                done = L.ii >= J.nt - 1;
            }
            need = false;
        }
        set_wave_force(f);
        if (!done)
            done = attempt_by_kind(
                f, [&](const auto& force) { return dopr54_attempt(L, J, force, sink); });
        if (done || L.ii >= stop) {
            if (res.base) save(res, L, done ? kOrbitDone : kOrbitRunning);
            if (done) {
                if (stats) {
                    drift_orbit_stats st;
                    st.accepted = L.accepted;
                    st.rejected = L.rejected;
                    st.dense_steps = 0;
                    st.status = L.err | L.status;
                    stats[sink.idx] = st;
                }
                if (done_count) atomicAdd(done_count, 1u);
            }
            need = true;
        }
    }
}

// Initial DOPR54 sub-steps of a launch's new orbits (the RK4 estimate,
// dopr54.rs:210-368), one lane per work item, before the integrator runs.
// Inside the persistent integrator a lane that starts an orbit would run the
// estimate's twelve force evaluations per iteration alone while the rest of
// its wavefront waits: skipping it took the mix's first 1,024 jobs' DOPR54
// part from 384 ms to 344-362 ms.  MW forces read the table in global memory;
// the LDS copy holds the same bits.
template <class FT>
__global__ void __launch_bounds__(256) ragged_dopr54_estimate_kernel(RaggedArgs a) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.items) return;
    const RaggedJob& d = a.jobs[find_job(a.jobs, a.njobs, (unsigned long long)slot)];
    if (d.dt_one_init != dp54::initial_step_sentinel) return;
    const int64_t orbit = a.order ? (int64_t)a.order[slot] : slot - d.item_begin;
    if (d.resident) {
        const Resident res{d.resident, d.n, orbit};
        if (__double_as_longlong(res.w(kResidentPhaseWord)) != kOrbitNew) return;
    }
    LaneForce<FT> f{};
    set_lane_force<FT>(f, a, d.pot);
    f.in_lds = false;
    f.wave_kind = -1;
    double y0[DIM];
    const double2* p = reinterpret_cast<const double2*>(d.state0 + orbit * DIM);
    const double2 u = p[0], v = p[1], w = p[2];
    y0[0] = u.x; y0[1] = u.y; y0[2] = v.x; y0[3] = v.y; y0[4] = w.x; y0[5] = w.y;
    Dopr54Job J{};
    J.dt = d.times[1] - d.times[0];
    J.rtol = d.rtol;
    J.atol = d.atol;
    a.dt_est[slot] = rk4_estimate_step(y0, J.dt, d.times[0], J, f);
}

// Sort keys of a launch's work items for cost order: (job index, bucket of
// the orbit's step rate at t0, omega = sqrt(|a| / r), highest first) for the
// jobs of at least min_orbits orbits, (job index, 0) for the others, which so
// keep index order (the sort is stable); orbits[] holds each item's orbit.
// Sorted, the items stay grouped by job in job order.  A job that alone fills
// the device's lanes otherwise runs in index order, and lanes that draw its
// long orbits late keep the launch running after the others have finished.
template <class FT>
__global__ void __launch_bounds__(256) ragged_order_key_kernel(RaggedArgs a, uint32_t* keys,
                                                               uint32_t* orbits,
                                                               int64_t min_orbits) {
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.items) return;
    const int32_t jix = find_job(a.jobs, a.njobs, (unsigned long long)slot);
    const RaggedJob& d = a.jobs[jix];
    const int64_t orbit = slot - d.item_begin;
    uint32_t bin = 0;
    if (d.n >= min_orbits) {
        LaneForce<FT> f{};
        set_lane_force<FT>(f, a, d.pot);
        f.in_lds = false;
        f.wave_kind = -1;
        double y[DIM], acc[DIM];
        const double2* p = reinterpret_cast<const double2*>(d.state0 + orbit * DIM);
        const double2 u = p[0], v = p[1], w = p[2];
        y[0] = u.x; y[1] = u.y; y[2] = v.x; y[3] = v.y; y[4] = w.x; y[5] = w.y;
        f(d.times[0], y, acc);
        const double r2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        const double a2 = acc[3] * acc[3] + acc[4] * acc[4] + acc[5] * acc[5];
        bin = order_bin_of_rate(std::sqrt(std::sqrt(a2 / r2)));
    }
    keys[slot] = ((uint32_t)jix << 10) | bin;
    orbits[slot] = (uint32_t)orbit;
}
static_assert(kOrderBins == 1 << 10, "a sort key holds the bucket in its low 10 bits");

__global__ void ragged_reset_queue(unsigned long long* queue) { *queue = 0ull; }

// Marks every orbit of a resident buffer "not started" (a fill kernel, so the
// sequence stays valid inside a captured graph).
__global__ void ragged_clear_phase(double* resident, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) resident[(int64_t)kResidentPhaseWord * n + i] = __longlong_as_double(kOrbitNew);
}

__global__ void ragged_clear_u32(unsigned int* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

int ragged_prepare(drift_session* s, const RaggedJob* host_jobs, int32_t njobs, int method,
                   const std::vector<PotParams>& pots, const std::vector<const void*>& t64,
                   const std::vector<const void*>& t32, const drift_launch_options& opt,
                   RaggedScratch& scratch, RaggedLaunch& launch, int64_t order_min_orbits) {
    launch = RaggedLaunch{};
    if (njobs == 0) return DRIFT_OK;
    const int64_t items = host_jobs[njobs - 1].item_begin + host_jobs[njobs - 1].n;
    if (items == 0) return DRIFT_OK;
    // one contiguous upload: jobs | pots | tables64 | tables32
    const size_t jobs_bytes = sizeof(RaggedJob) * (size_t)njobs;
    const size_t pots_bytes = sizeof(PotParams) * pots.size();
    const size_t tab_bytes = sizeof(void*) * pots.size();
    const size_t total = jobs_bytes + pots_bytes + 2 * tab_bytes;
    int rc = ensure_buffer(scratch.desc, total);
    if (rc != DRIFT_OK) return rc;
    // the pinned staging may still feed the last eager upload
    if (scratch.upload_done) DRIFT_HIP_TRY(hipEventSynchronize(scratch.upload_done));
    if (scratch.pinned_bytes < total) {
        if (scratch.pinned) DRIFT_HIP_TRY(hipHostFree(scratch.pinned));
        scratch.pinned = nullptr;
        scratch.pinned_bytes = 0;
        DRIFT_HIP_TRY(hipHostMalloc((void**)&scratch.pinned, total, hipHostMallocDefault));
        scratch.pinned_bytes = total;
    }
    if (!scratch.upload_done)
        DRIFT_HIP_TRY(hipEventCreateWithFlags(&scratch.upload_done, hipEventDisableTiming));
    char* h = scratch.pinned;
    std::memcpy(h, host_jobs, jobs_bytes);
    std::memcpy(h + jobs_bytes, pots.data(), pots_bytes);
    std::memcpy(h + jobs_bytes + pots_bytes, t64.data(), tab_bytes);
    std::memcpy(h + jobs_bytes + pots_bytes + tab_bytes, t32.data(), tab_bytes);
    if (!scratch.queue) DRIFT_HIP_TRY(hipMalloc(&scratch.queue, sizeof(unsigned long long)));

    char* d = static_cast<char*>(scratch.desc.ptr);
    RaggedArgs& a = launch.args;
    a.jobs = reinterpret_cast<const RaggedJob*>(d);
    a.njobs = njobs;
    a.items = items;
    a.pots = reinterpret_cast<const PotParams*>(d + jobs_bytes);
    a.tables64 = reinterpret_cast<const void* const*>(d + jobs_bytes + pots_bytes);
    a.tables32 = reinterpret_cast<const void* const*>(d + jobs_bytes + pots_bytes + tab_bytes);
    a.queue = scratch.queue;
    a.max_attempts = opt.max_attempts ? opt.max_attempts : (1u << 22);
    dopr54_bounds(a.bound2);
    const bool mixed = opt.precision == DRIFT_PRECISION_MIXED;
    // the MW potential with the most orbits in this launch has its table in LDS
    a.lds_pot = -1;
    {
        std::vector<int64_t> orbits(pots.size(), 0);
        for (int32_t k = 0; k < njobs; ++k) orbits[host_jobs[k].pot] += host_jobs[k].n;
        int64_t best = 0;
        for (size_t k = 0; k < pots.size(); ++k)
            if (pots[k].kind == POT_MW && orbits[k] > best) {
                best = orbits[k];
                a.lds_pot = (int32_t)k;
            }
        if (a.lds_pot >= 0) a.lds_src = mixed ? t32[a.lds_pot] : t64[a.lds_pot];
    }
    if (method == DRIFT_METHOD_DOP853) {
        launch.kernel = mixed ? ragged_dop853_kernel<float> : ragged_dop853_kernel<double>;
    } else {
        launch.kernel = mixed ? ragged_dopr54_kernel<float> : ragged_dopr54_kernel<double>;
        launch.estimate =
            mixed ? ragged_dopr54_estimate_kernel<float> : ragged_dopr54_estimate_kernel<double>;
        rc = ensure_buffer(scratch.dt_est, sizeof(double) * (size_t)items);
        if (rc != DRIFT_OK) return rc;
        a.dt_est = static_cast<double*>(scratch.dt_est.ptr);
    }
    // cost order when a job alone fills the device's lanes (and the job index
    // and the bucket fit a 32-bit key)
    int job_bits = 0;
    while (job_bits < 31 && (int64_t(1) << job_bits) < (int64_t)njobs) job_bits++;
    bool order = order_min_orbits > 0 && job_bits + 10 <= 32;
    if (order) {
        order = false;
        for (int32_t k = 0; k < njobs; ++k) order = order || host_jobs[k].n >= order_min_orbits;
    }
    if (order) {
        launch.key_kernel = mixed ? ragged_order_key_kernel<float> : ragged_order_key_kernel<double>;
        launch.order_min_orbits = order_min_orbits;
        launch.sort_end_bit = job_bits + 10;
        size_t temp = 0;
        DRIFT_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
            nullptr, temp, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
            (uint32_t*)nullptr, items, 0, launch.sort_end_bit));
        const size_t words = (size_t)items * sizeof(uint32_t);
        const size_t arrays = (4 * words + 255) & ~size_t(255);
        rc = ensure_buffer(scratch.sort, arrays + temp);
        if (rc != DRIFT_OK) return rc;
        char* b = static_cast<char*>(scratch.sort.ptr);
        launch.keys_in = reinterpret_cast<uint32_t*>(b);
        launch.keys_out = reinterpret_cast<uint32_t*>(b + words);
        launch.orbits_in = reinterpret_cast<uint32_t*>(b + 2 * words);
        launch.orbits_out = reinterpret_cast<uint32_t*>(b + 3 * words);
        launch.sort_temp = b + arrays;
        launch.sort_temp_bytes = temp;
        a.order = launch.orbits_out;
    }
    launch.block = opt.block_size ? opt.block_size : 256;
    int per_cu = 0;
    DRIFT_HIP_TRY(
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, launch.kernel, launch.block, 0));
    const int64_t resident = (int64_t)std::max(per_cu, 1) * std::max(s->cu_count, 1);
    launch.blocks = std::min<int64_t>(resident, (items + launch.block - 1) / launch.block);
    launch.bytes = total;
    return DRIFT_OK;
}

int ragged_enqueue(const RaggedLaunch& launch, RaggedScratch& scratch, hipStream_t stream,
                   bool capturing) {
    if (!launch.kernel) return DRIFT_OK;  // nothing to integrate
    DRIFT_HIP_TRY(hipMemcpyAsync(scratch.desc.ptr, scratch.pinned, launch.bytes,
                                 hipMemcpyHostToDevice, stream));
    // under capture the caller synchronises before it rewrites the staging
    if (!capturing) DRIFT_HIP_TRY(hipEventRecord(scratch.upload_done, stream));
    hipLaunchKernelGGL(ragged_reset_queue, dim3(1), dim3(1), 0, stream, launch.args.queue);
    DRIFT_HIP_TRY(hipGetLastError());
    if (launch.key_kernel) {
        const int64_t blocks = (launch.args.items + 255) / 256;
        if (blocks > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "too many work items");
        hipLaunchKernelGGL(launch.key_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                           launch.args, launch.keys_in, launch.orbits_in, launch.order_min_orbits);
        DRIFT_HIP_TRY(hipGetLastError());
        size_t temp = launch.sort_temp_bytes;
        DRIFT_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(
            launch.sort_temp, temp, launch.keys_in, launch.keys_out, launch.orbits_in,
            launch.orbits_out, launch.args.items, 0, launch.sort_end_bit, stream));
    }
    if (launch.estimate) {
        const int64_t blocks = (launch.args.items + 255) / 256;
        if (blocks > INT32_MAX) return fail(DRIFT_ERR_INVALID_ARGUMENT, "too many work items");
        hipLaunchKernelGGL(launch.estimate, dim3((unsigned)blocks), dim3(256), 0, stream,
                           launch.args);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    const bool timed = scratch.time_kernel && !capturing;
    if (timed) {
        if (!scratch.kernel_start) DRIFT_HIP_TRY(hipEventCreate(&scratch.kernel_start));
        if (!scratch.kernel_end) DRIFT_HIP_TRY(hipEventCreate(&scratch.kernel_end));
        DRIFT_HIP_TRY(hipEventRecord(scratch.kernel_start, stream));
    }
    hipLaunchKernelGGL(launch.kernel, dim3((unsigned)launch.blocks), dim3(launch.block), 0,
                       stream, launch.args);
    DRIFT_HIP_TRY(hipGetLastError());
    if (timed) DRIFT_HIP_TRY(hipEventRecord(scratch.kernel_end, stream));
    return DRIFT_OK;
}

int ragged_launch(drift_session* s, const RaggedJob* host_jobs, int32_t njobs, int method,
                  const std::vector<PotParams>& pots, const std::vector<const void*>& t64,
                  const std::vector<const void*>& t32, const drift_launch_options& opt,
                  hipStream_t stream, RaggedScratch& scratch, int64_t order_min_orbits) {
    RaggedLaunch launch;
    int rc = ragged_prepare(s, host_jobs, njobs, method, pots, t64, t32, opt, scratch, launch,
                            order_min_orbits);
    if (rc != DRIFT_OK) return rc;
    return ragged_enqueue(launch, scratch, stream, false);
}

void ragged_scratch_free(RaggedScratch& scratch) {
    if (scratch.desc.ptr) (void)hipFree(scratch.desc.ptr);
    scratch.desc.ptr = nullptr;
    scratch.desc.bytes = 0;
    if (scratch.dt_est.ptr) (void)hipFree(scratch.dt_est.ptr);
    scratch.dt_est.ptr = nullptr;
    scratch.dt_est.bytes = 0;
    if (scratch.sort.ptr) (void)hipFree(scratch.sort.ptr);
    scratch.sort.ptr = nullptr;
    scratch.sort.bytes = 0;
    if (scratch.pinned) (void)hipHostFree(scratch.pinned);
    scratch.pinned = nullptr;
    scratch.pinned_bytes = 0;
    if (scratch.upload_done) (void)hipEventDestroy(scratch.upload_done);
    scratch.upload_done = nullptr;
    if (scratch.queue) (void)hipFree(scratch.queue);
    scratch.queue = nullptr;
    if (scratch.kernel_start) (void)hipEventDestroy(scratch.kernel_start);
    if (scratch.kernel_end) (void)hipEventDestroy(scratch.kernel_end);
    scratch.kernel_start = scratch.kernel_end = nullptr;
    scratch.time_kernel = false;
}

int ragged_clear(double* resident, int64_t n, unsigned int* done_count, hipStream_t stream) {
    if (resident && n > 0) {
        hipLaunchKernelGGL(ragged_clear_phase, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream, resident, n);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    if (done_count) {
        hipLaunchKernelGGL(ragged_clear_u32, dim3(1), dim3(64), 0, stream, done_count, 1);
        DRIFT_HIP_TRY(hipGetLastError());
    }
    return DRIFT_OK;
}

// Estimated relative cost of one orbit of a job: DOP853 steps scale as
// rtol^(-1/8), DOPR54 as rtol^(-1/5), both linearly with the span (the
// caller's span_hint; the grid itself is in device memory).
double job_cost(const drift_job& j) {
    const double e = j.method == DRIFT_METHOD_DOP853 ? -1.0 / 8.0 : -1.0 / 5.0;
    return std::exp(e * std::min(j.rtol, j.atol)) * std::fabs(j.span_hint > 0 ? j.span_hint : 1.0);
}

int validate_jobs(const drift_job* jobs, size_t njobs, size_t npots) {
    for (size_t k = 0; k < njobs; ++k) {
        const drift_job& j = jobs[k];
        const std::string where = "job " + std::to_string(k) + ": ";
        if (j.method != DRIFT_METHOD_DOPR54 && j.method != DRIFT_METHOD_DOP853)
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "unknown method");
        if (j.potential < 0 || (size_t)j.potential >= npots)
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "potential index out of range");
        if (j.nt < 2 || j.nt > (size_t)INT32_MAX)
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "nt must be in [2, 2^31)");
        if (j.n > 0 && (!j.d_state0 || !j.d_times || !j.d_state_out))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "NULL buffer");
        if ((reinterpret_cast<uintptr_t>(j.d_state0) | reinterpret_cast<uintptr_t>(j.d_state_out)) &
            15u)
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "buffers must be 16-byte aligned");
        if (!std::isfinite(j.rtol) || !std::isfinite(j.atol))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, where + "rtol/atol (logs) must be finite");
    }
    return DRIFT_OK;
}

int resolve_pots(drift_session* s, const drift_potential* potentials, size_t npots,
                 std::vector<PotParams>& pots, std::vector<const void*>& t64,
                 std::vector<const void*>& t32) {
    pots.resize(npots);
    t64.assign(npots, nullptr);
    t32.assign(npots, nullptr);
    for (size_t k = 0; k < npots; ++k) {
        if (potentials && potentials[k].origins != 0)
            return fail(DRIFT_ERR_UNSUPPORTED,
                        "moving origins are supported by the single-batch entry points only");
        int rc = resolve_potential(s, potentials ? &potentials[k] : nullptr, pots[k], &t64[k],
                                   &t32[k]);
        if (rc != DRIFT_OK) return rc;
        // without origins the only moving kind is NONE's empty set, which the
        // lane force evaluates as the static NONE kind
        pots[k].kind = static_kind(pots[k].kind);
    }
    return DRIFT_OK;
}

// Job order inside a launch: grouped by potential kind (lanes of a wavefront
// then mostly evaluate one force), costliest first within a group so the
// queue's tail is made of short orbits.
std::vector<size_t> ragged_order(const drift_job* jobs, size_t njobs,
                                 const std::vector<PotParams>& pots, int method) {
    std::vector<size_t> idx;
    for (size_t k = 0; k < njobs; ++k)
        if (jobs[k].method == method && jobs[k].n > 0) idx.push_back(k);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        const int ka = pots[jobs[a].potential].kind, kb = pots[jobs[b].potential].kind;
        if (ka != kb) return ka > kb;
        return job_cost(jobs[a]) > job_cost(jobs[b]);
    });
    return idx;
}

}  // namespace drift

using namespace drift;

extern "C" int drift_ragged_device(drift_session* session, const drift_job* jobs, size_t njobs,
                                   const drift_potential* potentials, size_t npotentials,
                                   const drift_launch_options* options, void* stream) {
    if (!session) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (njobs == 0) return DRIFT_OK;
    if (!jobs) return fail(DRIFT_ERR_INVALID_ARGUMENT, "jobs is NULL");
    if (npotentials == 0) return fail(DRIFT_ERR_INVALID_ARGUMENT, "no potentials");
    int rc = validate_jobs(jobs, njobs, npotentials);
    if (rc != DRIFT_OK) return rc;
    drift_launch_options opt{};
    if (options) opt = *options;
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    if (opt.block_size && (opt.block_size < 64 || opt.block_size > 256 || opt.block_size % 64))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 in [64, 256]");
    if (opt.staging == DRIFT_STAGING_STAGED)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "ragged launches write every job's output directly (staging is DIRECT)");
    if (opt.staging < DRIFT_STAGING_AUTO || opt.staging > DRIFT_STAGING_STAGED)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "staging must be AUTO, DIRECT or STAGED");
    // every job is checked before the first kernel is enqueued, so a refusal
    // never leaves a launch of this call running
    if (opt.precision == DRIFT_PRECISION_MIXED)
        for (size_t k = 0; k < njobs; ++k)
            if (jobs[k].n > 0 && std::exp(jobs[k].rtol) < 1e-7 * (1 - 1e-9))
                return fail(DRIFT_ERR_UNSUPPORTED,
                            "job " + std::to_string(k) +
                                ": mixed precision (FP32 stage forces) requires rtol >= 1e-7");
    DRIFT_HIP_TRY(hipSetDevice(session->device));
    std::vector<PotParams> pots;
    std::vector<const void*> t64, t32;
    rc = resolve_pots(session, potentials, npotentials, pots, t64, t32);
    if (rc != DRIFT_OK) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : session->stream;
    bool enqueued = false;
    for (int method : {DRIFT_METHOD_DOP853, DRIFT_METHOD_DOPR54}) {
        std::vector<size_t> order = ragged_order(jobs, njobs, pots, method);
        std::vector<RaggedJob> rj;
        int64_t begin = 0;
        for (size_t k : order) {
            const drift_job& j = jobs[k];
            RaggedJob r{};
            r.state0 = j.d_state0;
            r.times = j.d_times;
            r.out = j.d_state_out;
            r.stats = j.d_stats;
            r.resident = nullptr;
            r.done_count = nullptr;
            r.n = (int64_t)j.n;
            r.item_begin = begin;
            r.nt = (int32_t)j.nt;
            r.method = j.method;
            r.stop_sample = (int32_t)j.nt - 1;
            r.pot = j.potential;
            r.rtol = std::exp(j.rtol);
            r.atol = std::exp(j.atol);
            r.dt_one_init = j.dt_one_init;
            begin += r.n;
            rj.push_back(r);
        }
        drift_launch_options mo = opt;
        if (!mo.block_size) mo.block_size = session->ragged_block[opt.precision][method];
        // cost order inside the jobs that alone fill the device's lanes (2
        // waves of 64 per SIMD, 4 SIMDs per CU)
        const int64_t lanes = (int64_t)std::max(session->cu_count, 1) * 4 * 2 * 64;
        rc = ragged_launch(session, rj.data(), (int32_t)rj.size(), method, pots, t64, t32, mo, st,
                           session->ragged[method], lanes);
        if (rc != DRIFT_OK) {
            // a launch of the other method may already use the outputs and the
            // session's descriptors: let it finish before reporting the error
            if (enqueued) (void)hipStreamSynchronize(st);
            return rc;
        }
        enqueued = enqueued || !rj.empty();
    }
    // the session's descriptor buffers are reused by the next call: wait
    DRIFT_HIP_TRY(hipStreamSynchronize(st));
    return DRIFT_OK;
}

extern "C" int drift_tune_ragged(drift_session* session, const drift_job* jobs, size_t njobs,
                                 const drift_potential* potentials, size_t npotentials,
                                 const drift_launch_options* options, size_t prefix_samples,
                                 int32_t* best_blocks, void* stream) {
    if (!session) return fail(DRIFT_ERR_INVALID_ARGUMENT, "session is NULL");
    if (njobs == 0 || !jobs) return fail(DRIFT_ERR_INVALID_ARGUMENT, "no jobs to tune on");
    int rc = validate_jobs(jobs, njobs, npotentials);
    if (rc != DRIFT_OK) return rc;
    drift_launch_options opt{};
    if (options) opt = *options;
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    // trial jobs: each job's first k outputs, written to the session's scratch
    std::vector<drift_job> trial(jobs, jobs + njobs);
    size_t bytes = 0;
    for (drift_job& j : trial) {
        size_t k = prefix_samples ? prefix_samples : std::max<size_t>(2, j.nt / 20);
        k = std::max<size_t>(2, std::min(k, j.nt));
        j.nt = k;
        j.d_stats = nullptr;
        bytes += j.n * k * DIM * sizeof(double);
    }
    DRIFT_HIP_TRY(hipSetDevice(session->device));
    if ((rc = ensure_buffer(session->tune_out, std::max<size_t>(bytes, 16))) != DRIFT_OK) return rc;
    {
        char* base = static_cast<char*>(session->tune_out.ptr);
        size_t off = 0;
        for (drift_job& j : trial) {
            j.d_state_out = reinterpret_cast<double*>(base + off);
            off += j.n * j.nt * DIM * sizeof(double);
        }
    }
    // untimed pass: resolves the tables and loads the code objects
    drift_launch_options o = opt;
    o.block_size = 256;
    if ((rc = drift_ragged_device(session, trial.data(), njobs, potentials, npotentials, &o,
                                  stream)) != DRIFT_OK)
        return rc;
    for (int method : {DRIFT_METHOD_DOPR54, DRIFT_METHOD_DOP853}) {
        std::vector<drift_job> sub;
        for (const drift_job& j : trial)
            if (j.method == method) sub.push_back(j);
        int32_t best = 0;
        double best_s = INFINITY;
        if (!sub.empty()) {
            // the integrator kernel alone is timed (events around it): the
            // descriptor upload and the DOPR54 step estimates cost the same at
            // every block size, and on a short prefix they would drown the
            // differences
            //
            // Each block size runs twice and keeps its faster time.  A smaller
            // block than 256 must beat 256 by 3%: on the prefix, which lacks
            // the orbits' long middles, 64 once came out ahead by less and the
            // full mix's DOPR54 jobs then ran 25% slower than at 256.
            RaggedScratch& sc = session->ragged[method];
            sc.time_kernel = true;
            for (int block : {256, 128, 64}) {
                o.block_size = block;
                double t = INFINITY;
                for (int rep = 0; rep < 2; ++rep) {
                    // returns once the launch has finished (it synchronises the stream)
                    rc = drift_ragged_device(session, sub.data(), sub.size(), potentials,
                                             npotentials, &o, stream);
                    float ms = 0.0f;
                    if (rc == DRIFT_OK &&
                        hipEventElapsedTime(&ms, sc.kernel_start, sc.kernel_end) != hipSuccess)
                        rc = fail(DRIFT_ERR_DEVICE, "hipEventElapsedTime failed");
                    if (rc != DRIFT_OK) {
                        sc.time_kernel = false;
                        return rc;
                    }
                    t = std::min(t, (double)ms);
                }
                if (block == 256 ? t < best_s : t < 0.97 * best_s) {
                    best_s = t;
                    best = block;
                }
            }
            sc.time_kernel = false;
        }
        session->ragged_block[opt.precision][method] = best;
        if (best_blocks) best_blocks[method] = best;
    }
    return DRIFT_OK;
}
