// origins.cpp -- host-side construction of moving-origin tracks.
//
// A moving origin (CustomOrigin, shared/src/potential.rs:264-356) follows
// one quintic per time segment in each coordinate.  The reference reads such
// tables (CustomOrigin::origins, :290-331) but nothing in it writes one; this
// file builds them from an integrated orbit: on each segment between two
// consecutive samples the quintic Hermite interpolant that matches position,
// velocity and acceleration at both ends, so the track is C2 across samples
// and reproduces any polynomial motion of degree <= 5 exactly.
#include <cmath>
#include <cstdint>
#include <string>

#include "drift_internal.hpp"

namespace drift {

namespace {

// Coefficients (tau^5 .. tau^0) of the quintic p on tau in [0, 1] with
// p(0) = x0, p'(0) = v0, p''(0) = a0, p(1) = x1, p'(1) = v1, p''(1) = a1
// (derivatives with respect to tau).  With the low half fixed by the left end,
// the top three coefficients solve
//   c3 + c4 + c5 = D,  3 c3 + 4 c4 + 5 c5 = E,  6 c3 + 12 c4 + 20 c5 = G,
// D = x1 - x0 - v0 - a0/2, E = v1 - v0 - a0, G = a1 - a0.
void hermite5(double x0, double v0, double a0, double x1, double v1, double a1, double* c) {
    const double D = x1 - x0 - v0 - 0.5 * a0;
    const double E = v1 - v0 - a0;
    const double G = a1 - a0;
    c[0] = 6.0 * D - 3.0 * E + 0.5 * G;    // tau^5
    c[1] = -15.0 * D + 7.0 * E - G;        // tau^4
    c[2] = 10.0 * D - 4.0 * E + 0.5 * G;   // tau^3
    c[3] = 0.5 * a0;                       // tau^2
    c[4] = v0;                             // tau^1
    c[5] = x0;                             // tau^0
}

}  // namespace

int origin_tracks_hermite(const double* states, const double* accels, const double* times,
                          size_t nt, size_t n, double* tracks, double* t_start, double* t_end) {
    if (!states || !accels || !times || !tracks || !t_start || !t_end)
        return fail(DRIFT_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nt < 2) return fail(DRIFT_ERR_INVALID_ARGUMENT, "tracks need at least 2 sample times");
    if (n < 1) return fail(DRIFT_ERR_INVALID_ARGUMENT, "tracks need at least one orbit");
    for (size_t k = 0; k < nt; ++k)
        if (!std::isfinite(times[k]))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "sample times must be finite");
    const bool increasing = times[nt - 1] > times[0];
    for (size_t k = 1; k < nt; ++k)
        if (increasing ? !(times[k] > times[k - 1]) : !(times[k] < times[k - 1]))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "sample times must be strictly monotone");
    const double lo = increasing ? times[0] : times[nt - 1];
    const double hi = increasing ? times[nt - 1] : times[0];
    const size_t nseg = nt - 1;
    const double dt = (hi - lo) / (double)nseg;
    // the tracks are evaluated on equal segments: the samples must lie on them
    for (size_t k = 0; k < nt; ++k) {
        const size_t j = increasing ? k : nt - 1 - k;  // position in increasing time
        const double expect = lo + dt * (double)j;
        if (std::fabs(times[k] - expect) > 1e-9 * (hi - lo))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "sample times must be uniformly spaced");
    }
    for (size_t i = 0; i < n; ++i) {
        for (size_t s = 0; s < nseg; ++s) {
            // segment s spans samples kl -> kr in increasing time
            const size_t kl = increasing ? s : nt - 1 - s;
            const size_t kr = increasing ? s + 1 : nt - 2 - s;
            const double* yl = states + (kl * n + i) * 6;
            const double* yr = states + (kr * n + i) * 6;
            const double* al = accels + (kl * n + i) * 3;
            const double* ar = accels + (kr * n + i) * 3;
            double* c = tracks + (i * nseg + s) * ORIGIN_COEFS;
            for (int d = 0; d < 3; ++d)
                hermite5(yl[d], yl[3 + d] * dt, al[d] * dt * dt, yr[d], yr[3 + d] * dt,
                         ar[d] * dt * dt, c + 6 * d);
        }
    }
    for (size_t k = 0; k < nseg * n * ORIGIN_COEFS; ++k)
        if (!std::isfinite(tracks[k]))
            return fail(DRIFT_ERR_INVALID_ARGUMENT, "states and accelerations must be finite");
    *t_start = lo;
    *t_end = hi;
    return DRIFT_OK;
}

}  // namespace drift

extern "C" int drift_origin_tracks_hermite(const double* states, const double* accels,
                                           const double* times, size_t nt, size_t n,
                                           double* tracks, double* t_start, double* t_end) {
    return drift::origin_tracks_hermite(states, accels, times, nt, n, tracks, t_start, t_end);
}
