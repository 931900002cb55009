// Host build of the product's DOP853 and DOPR54 step machines
// (astrodrift_amd/csrc/integrators.hpp with the forces of forces.hpp), for CPU
// tests of their arithmetic against the oracle.  Test infrastructure: the library itself
// only runs these functions inside its gfx950 kernels.
#include <cstdint>
#include <vector>

#include "drift_internal.hpp"
#include "forces.hpp"
#include "integrators.hpp"

using namespace drift;

namespace {

struct HostSink {
    double* out;
    int64_t n;
    int64_t idx;
    void emit(int32_t s, const double* y) const {
        for (int c = 0; c < DIM; c++) out[((int64_t)s * n + idx) * DIM + c] = y[c];
    }
};

template <class F>
void run(const F& f, int64_t n, const double* y0, const double* t, int nt, double rtol,
         double atol, double* out, int64_t* stats) {
    const Dop853Job J = make_dop853_job(t, nt, rtol, atol, 1u << 22, t[0], t[nt - 1]);
    for (int64_t i = 0; i < n; i++) {
        HostSink sink{out, n, i};
        Dop853Lane L;
        dop853_init(L, y0 + DIM * i, t[0], J, f, sink);
        while (!dop853_attempt(L, J, f, sink)) {
        }
        stats[4 * i] = L.accepted;
        stats[4 * i + 1] = L.rejected;
        stats[4 * i + 2] = L.dense;
        stats[4 * i + 3] = L.status;
    }
}

template <class F>
void run54(const F& f, int64_t n, const double* y0, const double* t, int nt, double rtol,
           double atol, double dt_one_init, double* out, int64_t* stats) {
    Dopr54Job J;
    J.times = t;
    J.nt = nt;
    J.max_attempts = 1u << 22;
    J.dt = t[1] - t[0];
    J.rtol = rtol;
    J.atol = atol;
    J.dt_one_init = dt_one_init;
    dopr54_bounds(J.bound2);
    for (int64_t i = 0; i < n; i++) {
        HostSink sink{out, n, i};
        Dopr54Lane L;
        dopr54_init(L, y0 + DIM * i, t[0], J, f, sink);
        bool done = L.ii >= J.nt - 1;
        while (!done) done = dopr54_attempt(L, J, f, sink);
        stats[4 * i] = L.accepted;
        stats[4 * i + 1] = L.rejected;
        stats[4 * i + 2] = 0;
        stats[4 * i + 3] = L.err | L.status;
    }
}

}  // namespace

// DOPR54 with the unit galpy Kepler force (p = NULL), Kepler or Plummer.
extern "C" int host_dopr54(const drift_potential* p, int64_t n, const double* y0, const double* t,
                           int nt, double rtol, double atol, double dt_one_init, double* out,
                           int64_t* stats) {
    if (p == nullptr)
        run54(GalpyKeplerForce<double>{}, n, y0, t, nt, rtol, atol, dt_one_init, out, stats);
    else if (p->kind == DRIFT_POTENTIAL_KEPLER)
        run54(KeplerForce<double>{p->amp}, n, y0, t, nt, rtol, atol, dt_one_init, out, stats);
    else if (p->kind == DRIFT_POTENTIAL_PLUMMER)
        run54(PlummerForce<double>{p->amp, p->b * p->b}, n, y0, t, nt, rtol, atol, dt_one_init,
              out, stats);
    else
        return 1;
    return 0;
}

// potential: NULL = the unit galpy Kepler force; otherwise a drift_potential
// (Kepler, Plummer or Milky-Way).  Linear tolerances.
extern "C" int host_dop853(const drift_potential* p, int64_t n, const double* y0, const double* t,
                           int nt, double rtol, double atol, double* out, int64_t* stats) {
    if (p == nullptr) {
        run(GalpyKeplerForce<double>{}, n, y0, t, nt, rtol, atol, out, stats);
        return 0;
    }
    if (p->kind == DRIFT_POTENTIAL_KEPLER) {
        run(KeplerForce<double>{p->amp}, n, y0, t, nt, rtol, atol, out, stats);
    } else if (p->kind == DRIFT_POTENTIAL_PLUMMER) {
        run(PlummerForce<double>{p->amp, p->b * p->b}, n, y0, t, nt, rtol, atol, out, stats);
    } else if (p->kind == DRIFT_POTENTIAL_MW) {
        PotParams pp;
        build_mw_params(*p, pp);
        std::vector<double> tab(TABLE_DOUBLES), far(FAR_DOUBLES);
        build_mw_table(*p, tab.data());
        build_mw_far_table(*p, far.data());
        pp.far_table = far.data();
        run(MwForce<double>{&pp, tab.data(), pp.disk_amp, pp.disk_a, pp.disk_b2}, n, y0, t, nt,
            rtol, atol, out, stats);
    } else {
        return 1;
    }
    return 0;
}
