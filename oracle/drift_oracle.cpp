// drift_oracle.cpp -- C entry points of the CPU oracle (test infrastructure).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py, never by the product.
//
// Exposes
//   * the double instantiation of drift_oracle.hpp per orbit and per batch
//     (the batch scatters into the reference's time-major (nt, n, 6) layout,
//     src/dispatch/cpu.rs:132-168, optionally sharded over host threads);
//   * a counting instantiation that tallies source-level FP64 operations by
//     integrator phase, from which the algorithmic FLOP constants are derived;
//   * force, potential, incomplete-gamma and time-grid helpers for the tests.
#include "drift_oracle.hpp"

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

namespace drift_oracle {

// ---------------------------------------------------------------------------
// Counting scalar.  Each arithmetic operator forwards to double arithmetic
// (so the run follows exactly the same path as the double instantiation) and
// increments a per-region, per-kind tally.
// ---------------------------------------------------------------------------
enum OpKind : int { OP_ADD = 0, OP_MUL, OP_DIV, OP_SQRT, OP_TRANS, OP_KINDS };

struct OpTally {
    int region = REG_INIT;
    int64_t ops[REG_COUNT][OP_KINDS] = {};
};
thread_local OpTally g_tally;

struct Counted {
    double v;
    Counted() : v(0.0) {}
    Counted(double x) : v(x) {}  // NOLINT: implicit by design
};
inline void tally(int kind) { g_tally.ops[g_tally.region][kind]++; }

inline Counted operator+(Counted a, Counted b) { tally(OP_ADD); return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { tally(OP_ADD); return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { tally(OP_MUL); return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { tally(OP_DIV); return Counted(a.v / b.v); }
inline Counted operator+(Counted a, double b) { tally(OP_ADD); return Counted(a.v + b); }
inline Counted operator-(Counted a, double b) { tally(OP_ADD); return Counted(a.v - b); }
inline Counted operator*(Counted a, double b) { tally(OP_MUL); return Counted(a.v * b); }
inline Counted operator/(Counted a, double b) { tally(OP_DIV); return Counted(a.v / b); }
inline Counted operator+(double a, Counted b) { tally(OP_ADD); return Counted(a + b.v); }
inline Counted operator-(double a, Counted b) { tally(OP_ADD); return Counted(a - b.v); }
inline Counted operator*(double a, Counted b) { tally(OP_MUL); return Counted(a * b.v); }
inline Counted operator/(double a, Counted b) { tally(OP_DIV); return Counted(a / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline Counted& operator*=(Counted& a, Counted b) { a = a * b; return a; }
inline Counted& operator/=(Counted& a, Counted b) { a = a / b; return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator<(Counted a, double b) { return a.v < b; }
inline bool operator>(Counted a, double b) { return a.v > b; }
inline bool operator<=(Counted a, double b) { return a.v <= b; }
inline bool operator>=(Counted a, double b) { return a.v >= b; }

inline Counted m_sqrt(Counted x) { tally(OP_SQRT); return Counted(::sqrt(x.v)); }
inline Counted m_log(Counted x) { tally(OP_TRANS); return Counted(::log(x.v)); }
inline Counted m_exp(Counted x) { tally(OP_TRANS); return Counted(::exp(x.v)); }
inline Counted m_pow(Counted x, Counted y) { tally(OP_TRANS); return Counted(::pow(x.v, y.v)); }
inline Counted m_fabs(Counted x) { return Counted(::fabs(x.v)); }
inline Counted m_fmax(Counted x, Counted y) { return Counted(::fmax(x.v, y.v)); }
inline Counted m_ceil(Counted x) { return Counted(::ceil(x.v)); }
inline Counted m_floor(Counted x) { return Counted(::floor(x.v)); }
inline Counted m_round(Counted x) { return Counted(::round(x.v)); }
inline double m_value(Counted x) { return x.v; }
inline Counted m_lower_gamma(double a, Counted x) {
    tally(OP_TRANS);
    return Counted(lower_gamma(a, x.v));
}

template <>
struct RegionGuard<Counted> {
    int saved;
    explicit RegionGuard(int r) : saved(g_tally.region) { g_tally.region = r; }
    ~RegionGuard() { g_tally.region = saved; }
};

}  // namespace drift_oracle

using namespace drift_oracle;

namespace {

int integrate_one(int method, const Potential& pot, const double* y_in, const double* t, int nt,
                  double dt_one, double rtol_ln, double atol_ln, double* result, OrbitStats& st) {
    double y[DIM];
    for (int i = 0; i < DIM; i++) y[i] = y_in[i];
    if (method == 0) return dopr54<double>(pot, y, nt, dt_one, t, rtol_ln, atol_ln, result, st);
    dop853<double>(pot, y, nt, t, rtol_ln, atol_ln, result, st);
    return 0;
}

}  // namespace

extern "C" {

int oracle_abi_version(void) { return 1; }

// method: 0 = DOPR54, 1 = DOP853.  result: nt*6 doubles, row i = state at t[i].
// Returns the DOPR54 error code (always 0 for DOP853).
int oracle_orbit(int method, const Potential* pot, const double* y0, const double* t, int nt,
                 double dt_one, double rtol_ln, double atol_ln, double* result, OrbitStats* stats) {
    OrbitStats st;
    int err = integrate_one(method, *pot, y0, t, nt, dt_one, rtol_ln, atol_ln, result, st);
    if (stats) *stats = st;
    return err;
}

// Batch driver: orbits i = 0, stride, 2*stride, ... < n (a fixed, seeded
// subset when stride > 1).  out (optional) is time-major (nt, n, 6) as in
// src/dispatch/cpu.rs:150-153; only integrated orbits are written.
// stats (optional) has one OrbitStats per orbit index.  nthreads <= 1 runs on
// the calling thread, as the reference engine does (src/dispatch/cpu.rs:146).
// Returns the number of orbits integrated.
int64_t oracle_batch(int method, const Potential* pot, int64_t n, const double* state0,
                     const double* t, int nt, double dt_one, double rtol_ln, double atol_ln,
                     double* out, OrbitStats* stats, int nthreads, int64_t stride) {
    if (stride < 1) stride = 1;
    const int64_t count = (n + stride - 1) / stride;
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        std::vector<double> traj((size_t)nt * DIM);
        for (;;) {
            int64_t k = next.fetch_add(1);
            if (k >= count) break;
            int64_t i = k * stride;
            OrbitStats st;
            integrate_one(method, *pot, state0 + i * DIM, t, nt, dt_one, rtol_ln, atol_ln,
                          traj.data(), st);
            if (out) {
                for (int s = 0; s < nt; s++)
                    for (int c = 0; c < DIM; c++)
                        out[((int64_t)s * n + i) * DIM + c] = traj[(size_t)s * DIM + c];
            }
            if (stats) stats[i] = st;
        }
    };
    if (nthreads <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int k = 0; k < nthreads; k++) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }
    return count;
}

// Counting run of one orbit.  ops receives REG_COUNT x OP_KINDS int64 tallies
// (row = Region, column = add/sub, mul, div, sqrt, transcendental).
void oracle_count_ops(int method, const Potential* pot, const double* y0, const double* t, int nt,
                      double dt_one, double rtol_ln, double atol_ln, int64_t* ops,
                      OrbitStats* stats) {
    g_tally = OpTally();
    std::vector<Counted> tc(nt), res((size_t)nt * DIM);
    Counted y[DIM];
    for (int i = 0; i < DIM; i++) y[i] = Counted(y0[i]);
    for (int i = 0; i < nt; i++) tc[i] = Counted(t[i]);
    OrbitStats st;
    if (method == 0)
        dopr54<Counted>(*pot, y, nt, Counted(dt_one), tc.data(), Counted(rtol_ln),
                        Counted(atol_ln), res.data(), st);
    else
        dop853<Counted>(*pot, y, nt, tc.data(), Counted(rtol_ln), Counted(atol_ln), res.data(),
                        st);
    for (int r = 0; r < REG_COUNT; r++)
        for (int k = 0; k < OP_KINDS; k++) ops[r * OP_KINDS + k] = g_tally.ops[r][k];
    if (stats) *stats = st;
}

void oracle_force(const Potential* pot, double x, double y, double z, double* out3) {
    double ax, ay, az;
    force<double>(*pot, x, y, z, ax, ay, az);
    out3[0] = ax;
    out3[1] = ay;
    out3[2] = az;
}

// Static force plus moving origins at time t.
void oracle_force_at(const Potential* pot, double t, double x, double y, double z, double* out3) {
    double ax, ay, az;
    force_at<double>(*pot, t, x, y, z, ax, ay, az);
    out3[0] = ax;
    out3[1] = ay;
    out3[2] = az;
}

double oracle_potential_value(const Potential* pot, double x, double y, double z) {
    return potential_value(*pot, x, y, z);
}

double oracle_lower_gamma(double a, double x) { return lower_gamma(a, x); }
double oracle_upper_gamma(double a, double x) { return upper_gamma(a, x); }

// glibc log, used for LogTolerance::from_linear (src/integrators/galpy/mod.rs:19-24).
double oracle_ln(double x) { return ::log(x); }

void oracle_sample_times(double start, double end, int64_t steps, double* out) {
    for (int64_t i = 0; i < steps; i++) out[i] = grid_sample(start, end, steps, i);
}

}  // extern "C"
