// serving.cpp -- continuous-batching request loop over the ragged kernels.
//
// A server owns a session (device, stream, MW table cache), a worker thread
// that launches the chunks and a finisher thread that copies finished
// requests out on a second stream.  Submitted requests queue first-come
// first-served; the worker admits them while their device footprint fits the
// HBM budget (beside what the finisher still holds), then runs chunk
// launches: each advances every admitted request by up to chunk_samples
// output samples (one persistent ragged launch per integration method), the
// orbits' integrator state resident in HBM between chunks.  A request whose
// orbits have all finished is handed to the finisher, which copies it to its
// host buffer and releases its device memory while the next chunks run; the
// memory released admits the next requests at the following chunk boundary.
// So requests that arrive while others run join the GPU work within one
// chunk, and every result is independent of what it shared the GPU with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "drift_internal.hpp"
#include "ragged.hpp"

namespace drift {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    drift_request req{};
    std::vector<double> state0, times;
    drift_request_info info{};
    uint64_t bytes = 0;
    int32_t cursor = 0;
    double* d_state0 = nullptr;
    double* d_times = nullptr;
    double* d_out = nullptr;
    double* d_resident = nullptr;
    drift_orbit_stats* d_stats = nullptr;
    unsigned int* d_done = nullptr;
    unsigned int host_done = 0;
};

uint64_t footprint(const drift_request& r) {
    const uint64_t n = r.n, nt = r.nt;
    return 8ull * (6 * n + nt + 6 * n * nt + (uint64_t)kResidentWords * n) + 16ull * n + 256;
}

void free_device(Request& r) {
    for (void* p : {(void*)r.d_state0, (void*)r.d_times, (void*)r.d_out, (void*)r.d_resident,
                    (void*)r.d_stats, (void*)r.d_done})
        if (p) (void)hipFree(p);
    r.d_state0 = r.d_times = r.d_out = r.d_resident = nullptr;
    r.d_stats = nullptr;
    r.d_done = nullptr;
}

}  // namespace
}  // namespace drift

struct drift_server {
    drift_session* session = nullptr;
    drift_server_options opt{};
    std::thread worker;
    // copies finished requests to their host buffers and releases their
    // device memory on its own stream, so chunk launches never wait for it
    std::thread finisher;
    hipStream_t copy_stream = nullptr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_finish;
    std::deque<std::shared_ptr<drift::Request>> pending;
    std::deque<std::shared_ptr<drift::Request>> finishing;
    bool worker_done = false;
    std::map<uint64_t, std::shared_ptr<drift::Request>> requests;
    uint64_t next_ticket = 1;
    uint64_t used_bytes = 0;
    bool stopping = false;
    drift::Clock::time_point t0;
    drift::RaggedScratch scratch[2];
    int32_t chunk_now = 8;  // samples per chunk (adaptive when opt.chunk_samples == 0)
    // A chunk's stream work (per method: descriptor upload, queue reset,
    // ragged kernel; then the finished-orbit counters back to pinned memory),
    // captured into a hipGraph and replayed while the chunk's shape holds:
    // the kernels' arguments, grids and the buffers the copies use.  What
    // changes from chunk to chunk (each job's stop sample) lives in the
    // uploaded descriptors.  A new admitted set, a finished request or a
    // reallocated buffer changes the shape and the next chunk is captured anew.
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<uint64_t> graph_key;
    bool graphs_ok = true;  // false once capture failed: chunks are then launched eagerly
    unsigned int* done_pinned = nullptr;  // finished-orbit count of each active request
    size_t done_capacity = 0;
    std::atomic<uint64_t> graph_launches{0}, graph_captures{0}, eager_chunks{0};
    // drift_server_busy: summed chunk times (launch to stream synchronisation)
    // and result copies to host memory
    std::atomic<uint64_t> busy_ns{0}, copy_ns{0}, copy_bytes{0};
};

namespace drift {
namespace {

double now_s(const drift_server* S) {
    return std::chrono::duration<double>(Clock::now() - S->t0).count();
}

void complete(drift_server* S, Request& r, int state, int status, uint64_t accepted = 0) {
    free_device(r);
    std::vector<double>().swap(r.state0);  // the server's input copies
    std::vector<double>().swap(r.times);
    std::lock_guard<std::mutex> lk(S->mu);
    S->used_bytes -= r.bytes;
    r.info.accepted_steps = accepted;
    r.info.state = state;
    r.info.status = status;
    r.info.complete_s = now_s(S);
    S->cv_done.notify_all();
    S->cv_work.notify_all();  // released memory may admit a queued request
}

// A queued request may start when it fits the budget beside what is held
// (admitted or still being copied out), or when nothing is held at all.
bool admissible(const drift_server* S, const Request& r) {
    return S->used_bytes == 0 || S->used_bytes + r.bytes <= S->opt.hbm_budget_bytes;
}

// Allocates the request's device buffers and uploads its inputs.
int admit(drift_server* S, Request& r, hipStream_t st) {
    const size_t n = r.req.n, nt = r.req.nt;
    DRIFT_HIP_TRY(hipMalloc(&r.d_times, nt * sizeof(double)));
    DRIFT_HIP_TRY(hipMalloc(&r.d_done, sizeof(unsigned int)));
    if (n > 0) {
        DRIFT_HIP_TRY(hipMalloc(&r.d_state0, n * DIM * sizeof(double)));
        DRIFT_HIP_TRY(hipMalloc(&r.d_out, nt * n * DIM * sizeof(double)));
        DRIFT_HIP_TRY(hipMalloc(&r.d_resident, (size_t)kResidentWords * n * sizeof(double)));
        DRIFT_HIP_TRY(hipMalloc(&r.d_stats, n * sizeof(drift_orbit_stats)));
        DRIFT_HIP_TRY(hipMemcpyAsync(r.d_state0, r.state0.data(), n * DIM * sizeof(double),
                                     hipMemcpyHostToDevice, st));
    }
    DRIFT_HIP_TRY(hipMemcpyAsync(r.d_times, r.times.data(), nt * sizeof(double),
                                 hipMemcpyHostToDevice, st));
    int rc = ragged_clear(r.d_resident, (int64_t)n, r.d_done, st);
    if (rc != DRIFT_OK) return rc;
    std::lock_guard<std::mutex> lk(S->mu);
    r.info.state = DRIFT_REQUEST_RUNNING;
    r.info.admit_s = now_s(S);
    return DRIFT_OK;
}

// Copies a finished request to its host buffers.
int finish(Request& r, hipStream_t st, int& status, uint64_t& accepted) {
    const size_t n = r.req.n, nt = r.req.nt;
    status = DRIFT_OK;
    accepted = 0;
    if (n == 0) return DRIFT_OK;
    std::vector<drift_orbit_stats> own;
    drift_orbit_stats* stats = r.req.stats;
    if (!stats) {
        own.resize(n);
        stats = own.data();
    }
    DRIFT_HIP_TRY(hipMemcpyAsync(r.req.state_out, r.d_out, nt * n * DIM * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
    DRIFT_HIP_TRY(hipMemcpyAsync(stats, r.d_stats, n * sizeof(drift_orbit_stats),
                                 hipMemcpyDeviceToHost, st));
    DRIFT_HIP_TRY(hipStreamSynchronize(st));
    uint64_t acc = 0;
    for (size_t i = 0; i < n; ++i) {
        acc += stats[i].accepted;
        if (r.req.method == DRIFT_METHOD_DOPR54 && (stats[i].status & DRIFT_STATUS_CODE_MASK))
            status = DRIFT_ERR_INTEGRATION;
    }
    accepted = acc;
    return DRIFT_OK;
}

// One chunk over every active request.
int run_chunk(drift_server* S, std::vector<std::shared_ptr<Request>>& active, hipStream_t st) {
    // distinct potentials of this chunk
    std::vector<drift_potential> keys;
    std::vector<int32_t> pot_of(active.size());
    for (size_t k = 0; k < active.size(); ++k) {
        const drift_potential& p = active[k]->req.potential;
        auto it = std::find_if(keys.begin(), keys.end(), [&](const drift_potential& q) {
            return std::memcmp(&p, &q, sizeof(p)) == 0;
        });
        pot_of[k] = (int32_t)(it - keys.begin());
        if (it == keys.end()) keys.push_back(p);
    }
    std::vector<PotParams> pots;
    std::vector<const void*> t64, t32;
    int rc = resolve_pots(S->session, keys.data(), keys.size(), pots, t64, t32);
    if (rc != DRIFT_OK) return rc;
    drift_launch_options lo{};
    lo.precision = S->opt.precision;
    lo.block_size = S->opt.block_size;
    lo.max_attempts = S->opt.max_attempts;
    const int32_t chunk = S->opt.chunk_samples > 0 ? S->opt.chunk_samples : S->chunk_now;
    RaggedLaunch launches[2];
    for (int method : {DRIFT_METHOD_DOP853, DRIFT_METHOD_DOPR54}) {
        std::vector<size_t> idx;
        for (size_t k = 0; k < active.size(); ++k)
            if (active[k]->req.method == method && active[k]->req.n > 0) idx.push_back(k);
        std::stable_sort(idx.begin(), idx.end(),
                         [&](size_t a, size_t b) { return pots[pot_of[a]].kind > pots[pot_of[b]].kind; });
        std::vector<RaggedJob> jobs;
        int64_t begin = 0;
        for (size_t k : idx) {
            Request& r = *active[k];
            const int32_t last = (int32_t)r.req.nt - 1;
            r.cursor = std::min(last, r.cursor + chunk);
            RaggedJob j{};
            j.state0 = r.d_state0;
            j.times = r.d_times;
            j.out = r.d_out;
            j.stats = r.d_stats;
            j.resident = r.d_resident;
            j.done_count = r.d_done;
            j.n = (int64_t)r.req.n;
            j.item_begin = begin;
            j.nt = (int32_t)r.req.nt;
            j.method = method;
            j.stop_sample = r.cursor;
            j.pot = pot_of[k];
            j.rtol = std::exp(r.req.rtol);
            j.atol = std::exp(r.req.atol);
            j.dt_one_init = r.req.dt_one_init;
            begin += j.n;
            jobs.push_back(j);
        }
        rc = ragged_prepare(S->session, jobs.data(), (int32_t)jobs.size(), method, pots, t64,
                            t32, lo, S->scratch[method], launches[method]);
        if (rc != DRIFT_OK) return rc;
    }
    if (S->done_capacity < active.size()) {
        if (S->done_pinned) DRIFT_HIP_TRY(hipHostFree(S->done_pinned));
        S->done_pinned = nullptr;
        S->done_capacity = 0;
        const size_t cap = std::max<size_t>(64, 2 * active.size());
        DRIFT_HIP_TRY(hipHostMalloc((void**)&S->done_pinned, cap * sizeof(unsigned int),
                                    hipHostMallocDefault));
        S->done_capacity = cap;
    }
    // the chunk's shape: what the captured work bakes in
    std::vector<uint64_t> key;
    for (int method : {DRIFT_METHOD_DOP853, DRIFT_METHOD_DOPR54}) {
        const RaggedLaunch& L = launches[method];
        const RaggedArgs& a = L.args;
        const RaggedScratch& sc = S->scratch[method];
        for (uint64_t v : {(uint64_t)(uintptr_t)L.kernel, (uint64_t)L.blocks, (uint64_t)L.block,
                           (uint64_t)L.bytes, (uint64_t)(uintptr_t)sc.desc.ptr,
                           (uint64_t)(uintptr_t)sc.pinned, (uint64_t)(uintptr_t)a.jobs,
                           (uint64_t)a.njobs, (uint64_t)a.max_attempts, (uint64_t)a.items,
                           (uint64_t)(uintptr_t)a.pots, (uint64_t)(uintptr_t)a.tables64,
                           (uint64_t)(uintptr_t)a.tables32, (uint64_t)(uintptr_t)a.queue,
                           (uint64_t)(int64_t)a.lds_pot, (uint64_t)(uintptr_t)a.lds_src,
                           (uint64_t)(uintptr_t)a.dt_est, (uint64_t)(uintptr_t)L.estimate})
            key.push_back(v);
    }
    key.push_back((uint64_t)(uintptr_t)S->done_pinned);
    for (auto& r : active) key.push_back((uint64_t)(uintptr_t)r->d_done);
    auto enqueue = [&](bool capturing) -> int {
        for (int method : {DRIFT_METHOD_DOP853, DRIFT_METHOD_DOPR54}) {
            const int erc = ragged_enqueue(launches[method], S->scratch[method], st, capturing);
            if (erc != DRIFT_OK) return erc;
        }
        for (size_t k = 0; k < active.size(); ++k)
            DRIFT_HIP_TRY(hipMemcpyAsync(&S->done_pinned[k], active[k]->d_done,
                                         sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        return DRIFT_OK;
    };
    if (S->graphs_ok && (!S->graph_exec || key != S->graph_key)) {
        if (S->graph_exec) (void)hipGraphExecDestroy(S->graph_exec);
        if (S->graph) (void)hipGraphDestroy(S->graph);
        S->graph_exec = nullptr;
        S->graph = nullptr;
        S->graph_key.clear();
        // thread-local mode: the finisher thread keeps copying and freeing
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            S->graphs_ok = false;
        } else {
            const int crc = enqueue(true);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(st, &g);
            hipGraphExec_t exec = nullptr;
            if (crc == DRIFT_OK && ec == hipSuccess && g &&
                hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) == hipSuccess) {
                S->graph = g;
                S->graph_exec = exec;
                S->graph_key = key;
                S->graph_captures++;
            } else {
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
                S->graphs_ok = false;
            }
        }
    }
    if (S->graphs_ok && S->graph_exec) {
        DRIFT_HIP_TRY(hipGraphLaunch(S->graph_exec, st));
        S->graph_launches++;
    } else {
        if ((rc = enqueue(false)) != DRIFT_OK) return rc;
        S->eager_chunks++;
    }
    // also orders the next chunk's rewrite of the pinned descriptors and counters
    DRIFT_HIP_TRY(hipStreamSynchronize(st));
    for (size_t k = 0; k < active.size(); ++k) active[k]->host_done = S->done_pinned[k];
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto& r : active) r->info.chunks++;
    return DRIFT_OK;
}

void worker_main(drift_server* S) {
    (void)hipSetDevice(S->session->device);
    hipStream_t st = S->session->stream;
    std::vector<std::shared_ptr<Request>> active;
    for (;;) {
        std::vector<std::shared_ptr<Request>> admitted;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (active.empty())
                S->cv_work.wait(lk, [&] {
                    return (S->stopping && S->pending.empty()) ||
                           (!S->pending.empty() && admissible(S, *S->pending.front()));
                });
            if (S->stopping && S->pending.empty() && active.empty()) break;
            while (!S->pending.empty() && admissible(S, *S->pending.front())) {
                auto& r = S->pending.front();
                S->used_bytes += r->bytes;
                admitted.push_back(r);
                S->pending.pop_front();
            }
        }
        for (auto& r : admitted) {
            int rc = admit(S, *r, st);
            if (rc != DRIFT_OK) {
                complete(S, *r, DRIFT_REQUEST_FAILED, rc);
            } else if (r->req.n == 0) {
                complete(S, *r, DRIFT_REQUEST_DONE, DRIFT_OK);
            } else {
                active.push_back(r);
            }
        }
        if (active.empty()) continue;
        const Clock::time_point c0 = Clock::now();
        int rc = run_chunk(S, active, st);
        const double ms = std::chrono::duration<double, std::milli>(Clock::now() - c0).count();
        S->busy_ns += (uint64_t)(ms * 1e6);
        if (S->opt.chunk_samples == 0 && rc == DRIFT_OK) {
            // adaptive chunk length: keep a chunk near the target duration so
            // queued requests join within about one chunk, and chunks stay long
            // enough that the pause/resume at their boundaries is cheap
            if (ms < 0.5 * S->opt.chunk_target_ms) S->chunk_now = std::min(S->chunk_now * 2, 512);
            else if (ms > 2.0 * S->opt.chunk_target_ms) S->chunk_now = std::max(S->chunk_now / 2, 1);
        }
        std::vector<std::shared_ptr<Request>> still;
        for (auto& r : active) {
            if (rc != DRIFT_OK) {
                complete(S, *r, DRIFT_REQUEST_FAILED, rc);
            } else if (r->host_done >= r->req.n) {
                // the chunk's stream is synchronised: the outputs are final
                std::lock_guard<std::mutex> lk(S->mu);
                S->finishing.push_back(r);
                S->cv_finish.notify_one();
            } else {
                still.push_back(r);
            }
        }
        active.swap(still);
    }
}

void finisher_main(drift_server* S) {
    (void)hipSetDevice(S->session->device);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_finish.wait(lk, [&] { return S->worker_done || !S->finishing.empty(); });
            if (S->finishing.empty()) break;  // worker_done and nothing left
            r = S->finishing.front();
            S->finishing.pop_front();
        }
        int status = DRIFT_OK;
        uint64_t accepted = 0;
        const Clock::time_point f0 = Clock::now();
        const int frc = finish(*r, S->copy_stream, status, accepted);
        S->copy_ns += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                          Clock::now() - f0).count();
        S->copy_bytes += (uint64_t)r->req.nt * r->req.n * DIM * sizeof(double);
        complete(S, *r, frc == DRIFT_OK ? DRIFT_REQUEST_DONE : DRIFT_REQUEST_FAILED,
                 frc == DRIFT_OK ? status : frc, accepted);
    }
}

}  // namespace
}  // namespace drift

using namespace drift;

extern "C" {

int drift_server_create(int device, const drift_server_options* options, drift_server** out) {
    if (!out) return fail(DRIFT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    drift_server_options opt{};
    if (options) opt = *options;
    if (opt.precision != DRIFT_PRECISION_FP64 && opt.precision != DRIFT_PRECISION_MIXED)
        return fail(DRIFT_ERR_UNSUPPORTED, "unknown precision mode");
    if (opt.chunk_samples < 0) return fail(DRIFT_ERR_INVALID_ARGUMENT, "chunk_samples < 0");
    if (!(opt.chunk_target_ms >= 0.0) || !std::isfinite(opt.chunk_target_ms))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "chunk_target_ms must be finite and >= 0");
    if (opt.chunk_target_ms == 0.0) opt.chunk_target_ms = 10.0;
    if (opt.block_size && (opt.block_size < 64 || opt.block_size > 256 || opt.block_size % 64))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 in [64, 256]");
    drift_session* sess = nullptr;
    int rc = drift_session_create(device, &sess);
    if (rc != DRIFT_OK) return rc;
    if (opt.hbm_budget_bytes == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
            drift_session_destroy(sess);
            return fail(DRIFT_ERR_DEVICE, "hipMemGetInfo failed");
        }
        opt.hbm_budget_bytes = free_b / 2;
    }
    hipStream_t copy_stream = nullptr;
    if (hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking) != hipSuccess) {
        drift_session_destroy(sess);
        return fail(DRIFT_ERR_DEVICE, "hipStreamCreateWithFlags failed");
    }
    drift_server* S = new drift_server();
    S->session = sess;
    S->opt = opt;
    S->copy_stream = copy_stream;
    S->t0 = Clock::now();
    S->worker = std::thread(worker_main, S);
    S->finisher = std::thread(finisher_main, S);
    *out = S;
    return DRIFT_OK;
}

int drift_server_submit(drift_server* S, const drift_request* request, uint64_t* ticket) {
    if (!S || !request || !ticket) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    const drift_request& q = *request;
    if (q.reserved != 0) return fail(DRIFT_ERR_INVALID_ARGUMENT, "drift_request.reserved must be 0");
    if (q.method != DRIFT_METHOD_DOPR54 && q.method != DRIFT_METHOD_DOP853)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "unknown method");
    if (q.nt < 2 || q.nt > (size_t)INT32_MAX)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "nt must be in [2, 2^31)");
    if (!q.times || (q.n > 0 && (!q.state0 || !q.state_out)))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (!std::isfinite(q.rtol) || !std::isfinite(q.atol))
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "rtol and atol (natural logs) must be finite");
    if (S->opt.precision == DRIFT_PRECISION_MIXED && std::exp(q.rtol) < 1e-7 * (1 - 1e-9))
        return fail(DRIFT_ERR_UNSUPPORTED, "mixed precision requires rtol >= 1e-7");
    if (q.potential.origins != 0)
        return fail(DRIFT_ERR_UNSUPPORTED,
                    "moving origins are supported by the single-batch entry points only");
    int rc = validate_potential(q.potential);
    if (rc != DRIFT_OK) return rc;
    auto r = std::make_shared<Request>();
    r->req = q;
    r->state0.assign(q.state0 ? q.state0 : (const double*)nullptr,
                     q.state0 ? q.state0 + q.n * DIM : (const double*)nullptr);
    r->times.assign(q.times, q.times + q.nt);
    r->bytes = footprint(q);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stopping) return fail(DRIFT_ERR_INVALID_ARGUMENT, "server is stopping");
    r->ticket = S->next_ticket++;
    r->info.state = DRIFT_REQUEST_QUEUED;
    r->info.submit_s = now_s(S);
    S->requests[r->ticket] = r;
    S->pending.push_back(r);
    *ticket = r->ticket;
    S->cv_work.notify_one();
    return DRIFT_OK;
}

int drift_server_query(drift_server* S, uint64_t ticket, drift_request_info* info) {
    if (!S || !info) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->requests.find(ticket);
    if (it == S->requests.end()) return fail(DRIFT_ERR_INVALID_ARGUMENT, "unknown ticket");
    *info = it->second->info;
    return DRIFT_OK;
}

int drift_server_stats(drift_server* S, uint64_t* graph_launches, uint64_t* graph_captures,
                       uint64_t* eager_chunks) {
    if (!S) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL server");
    if (graph_launches) *graph_launches = S->graph_launches.load();
    if (graph_captures) *graph_captures = S->graph_captures.load();
    if (eager_chunks) *eager_chunks = S->eager_chunks.load();
    return DRIFT_OK;
}

int drift_server_busy(drift_server* S, double* busy_s, double* copy_s, uint64_t* copy_bytes) {
    if (!S) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL server");
    if (busy_s) *busy_s = S->busy_ns.load() * 1e-9;
    if (copy_s) *copy_s = S->copy_ns.load() * 1e-9;
    if (copy_bytes) *copy_bytes = S->copy_bytes.load();
    return DRIFT_OK;
}

int drift_server_wait(drift_server* S, uint64_t ticket, double timeout_s,
                      drift_request_info* info) {
    if (!S) return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->requests.find(ticket);
    if (it == S->requests.end()) return fail(DRIFT_ERR_INVALID_ARGUMENT, "unknown ticket");
    auto r = it->second;
    auto finished = [&] {
        return r->info.state == DRIFT_REQUEST_DONE || r->info.state == DRIFT_REQUEST_FAILED;
    };
    if (timeout_s < 0)
        S->cv_done.wait(lk, finished);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) {
        if (info) *info = r->info;
        return fail(DRIFT_ERR_TIMEOUT, "request not finished within the timeout");
    }
    if (info) *info = r->info;
    if (r->info.status != DRIFT_OK)
        return fail(r->info.status, "request " + std::to_string(ticket) + " failed with status " +
                                        std::to_string(r->info.status));
    return DRIFT_OK;
}

int drift_server_destroy(drift_server* S) {
    if (!S) return DRIFT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->worker_done = true;
    }
    S->cv_finish.notify_all();
    if (S->finisher.joinable()) S->finisher.join();
    if (S->copy_stream) (void)hipStreamDestroy(S->copy_stream);
    (void)hipSetDevice(S->session->device);
    (void)hipStreamSynchronize(S->session->stream);
    if (S->graph_exec) (void)hipGraphExecDestroy(S->graph_exec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    if (S->done_pinned) (void)hipHostFree(S->done_pinned);
    for (auto& sc : S->scratch) ragged_scratch_free(sc);
    drift_session_destroy(S->session);
    delete S;
    return DRIFT_OK;
}

}  // extern "C"
