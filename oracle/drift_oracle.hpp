// drift_oracle.hpp -- CPU restatement of the reference's galpy DOPR54 / DOP853
// integrators and force formulas.
//
// TEST INFRASTRUCTURE ONLY.  This is the CPU restatement the HIP path is compared
// against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it.  The product (astrodrift_amd/) never links or calls it.
//
// Every routine is a statement-by-statement restatement of the reference Rust
// source it cites, with the same operation order, the same glibc libm calls and
// (when built by oracle/Makefile) -ffp-contract=off, so that a double
// instantiation reproduces the reference CPU engine bit for bit.  The code is
// templated over the scalar type R: R = double is the oracle proper, R = Counted
// (oracle/drift_oracle.cpp) counts the source-level floating-point operations
// of the same run, which is how the algorithmic FLOP constants are derived.
//
// Pins (see tests/test_oracle_*.py):
//   dopr54      -- reference KAT tail bits, tests/galpy_integrator_tests.rs:35-42
//   dop853      -- Kepler closed form, dense output vs. step-end states,
//                  tableau order conditions
//   forces      -- Plummer property tests (shared/tests/plummer_potential_tests.rs),
//                  -grad(Phi) by finite differences, vc^2(1) fractions of the
//                  Milky-Way model, scipy.special.gammainc for the bulge mass.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

namespace drift_oracle {

// ---------------------------------------------------------------------------
// libm access.  For R = double these are the glibc routines the reference
// links explicitly (src/integrators/galpy/dopr54.rs:37-47, dop853.rs:97-103).
// ---------------------------------------------------------------------------
inline double m_sqrt(double x) { return ::sqrt(x); }
inline double m_log(double x) { return ::log(x); }
inline double m_exp(double x) { return ::exp(x); }
inline double m_pow(double x, double y) { return ::pow(x, y); }
inline double m_fabs(double x) { return ::fabs(x); }
inline double m_fmax(double x, double y) { return ::fmax(x, y); }
inline double m_ceil(double x) { return ::ceil(x); }
inline double m_floor(double x) { return ::floor(x); }
inline double m_round(double x) { return ::round(x); }
inline bool m_isnan(double x) { return x != x; }
inline double m_value(double x) { return x; }

// Region tags used by the counting instantiation to attribute operations to
// integrator phases.  For R = double they compile to nothing.
enum Region : int {
    REG_INIT = 0,     // per-orbit set-up (initial step, first RHS)
    REG_ATTEMPT = 1,  // one attempted step (stages, update, error norm, control)
    REG_ACCEPT = 2,   // work done only on accepted steps (DOP853: FSAL + dense coeffs)
    REG_DENSE = 3,    // DOP853 extra dense stages + rcont finalisation
    REG_SAMPLE = 4,   // one emitted output sample
    REG_SUBSTEP = 5,  // DOPR54 sub-step bookkeeping (clamp to interval end)
    REG_RHS = 6,      // inside the force / right-hand side
    REG_COUNT = 7
};

template <class R>
struct RegionGuard {
    explicit RegionGuard(int) {}
};

// ---------------------------------------------------------------------------
// Potentials.  Layout mirrors drift_potential in include/drift_hip.h.
// ---------------------------------------------------------------------------
enum PotentialKind : int32_t {
    POT_GALPY_KEPLER = 0,  // the reference's hard-coded unit galpy Kepler force
    POT_KEPLER = 1,        // shared/src/potential.rs:217-239
    POT_PLUMMER = 2,       // shared/src/potential.rs:241-262
    POT_MW = 3,            // bulge + Miyamoto-Nagai + NFW, shared/src/potential.rs:87-176
    POT_NONE = 4,          // no static force (moving origins only)
};

struct Potential {
    int32_t kind;
    int32_t reserved;
    double amp;  // Kepler / Plummer amplitude (G M)
    double b;    // Plummer scale radius
    double bulge_amp, bulge_alpha, bulge_rc;
    double disk_amp, disk_a, disk_b;
    double halo_amp, halo_a;
    // Moving origins (CustomOrigin, shared/src/potential.rs:264-356), added
    // to the static force; n_origins = 0: none.
    const double* origin_tracks;  // n_origins x n_segments x 18 coefficients
    const double* origin_params;  // n_origins x (amp, b); b = 0: KeplerPotential
    int32_t n_origins, n_segments;
    double t_start, t_end;        // the segments split [t_start, t_end] evenly
};

// Lower incomplete gamma gamma(a, x) (unnormalised), by the power series
// (Abramowitz & Stegun 6.5.29) for x < a + 1 and by the Lentz continued fraction
// for the upper function Gamma(a, x) (A&S 6.5.31; Press et al., Numerical
// Recipes 2nd ed. section 6.2) otherwise.  Pure double arithmetic.
inline double upper_gamma_cf(double a, double x) {
    const double fpmin = 1e-300;
    double bb = x + 1.0 - a;
    double c = 1.0 / fpmin;
    double d = 1.0 / bb;
    double h = d;
    for (int i = 1; i < 1000; ++i) {
        double an = -i * (i - a);
        bb += 2.0;
        d = an * d + bb;
        if (::fabs(d) < fpmin) d = fpmin;
        c = bb + an / c;
        if (::fabs(c) < fpmin) c = fpmin;
        d = 1.0 / d;
        double del = d * c;
        h *= del;
        if (::fabs(del - 1.0) < 1e-17) break;
    }
    return ::exp(-x + a * ::log(x)) * h;
}

inline double lower_gamma_series(double a, double x) {
    double ap = a;
    double del = 1.0 / a;
    double sum = del;
    for (int n = 0; n < 1000; ++n) {
        ap += 1.0;
        del *= x / ap;
        sum += del;
        if (::fabs(del) < ::fabs(sum) * 1e-17) break;
    }
    return sum * ::exp(-x + a * ::log(x));
}

inline double lower_gamma(double a, double x) {
    if (x <= 0.0) return 0.0;
    if (x < a + 1.0) return lower_gamma_series(a, x);
    return ::tgamma(a) - upper_gamma_cf(a, x);
}

inline double upper_gamma(double a, double x) {
    if (x <= 0.0) return ::tgamma(a);
    if (x < a + 1.0) return ::tgamma(a) - lower_gamma_series(a, x);
    return upper_gamma_cf(a, x);
}

inline double m_lower_gamma(double a, double x) { return lower_gamma(a, x); }

// galpy Kepler force, the only force either reference engine evaluates:
// src/integrators/galpy/dopr54.rs:75-88 (identical to dop853.rs:43-57).
template <class R>
inline void galpy_kepler_force(const R& x, const R& y, const R& z, R& ax, R& ay, R& az) {
    R cylindrical_radius = m_sqrt(x * x + y * y);
    R sin_phi = y / cylindrical_radius;
    R cos_phi = x / cylindrical_radius;
    R radius_squared = cylindrical_radius * cylindrical_radius + z * z;
    R radial_force = -cylindrical_radius * m_pow(radius_squared, -1.5);
    double phi_torque = 0.0;
    ax = cos_phi * radial_force - 1.0 / cylindrical_radius * sin_phi * phi_torque;
    ay = sin_phi * radial_force + 1.0 / cylindrical_radius * cos_phi * phi_torque;
    az = -z * m_pow(radius_squared, -1.5);
}

// KeplerPotential::force, shared/src/potential.rs:229-238.
template <class R>
inline void kepler_force(double amp, const R& x, const R& y, const R& z, R& ax, R& ay, R& az) {
    R r2 = x * x + y * y + z * z;
    R r = m_sqrt(r2);
    R inv_r3 = 1.0 / (r2 * r);
    R ar = -amp * inv_r3;
    ax = ar * x;
    ay = ar * y;
    az = ar * z;
}

// PlummerPotential::force, shared/src/potential.rs:254-261.
template <class R>
inline void plummer_force(double amp, double b, const R& x, const R& y, const R& z, R& ax, R& ay,
                          R& az) {
    R r2 = x * x + y * y + z * z;
    R ar = -amp * m_pow(r2 + (b * b), -1.5);
    ax = ar * x;
    ay = ar * y;
    az = ar * z;
}

// MNPotential::force, shared/src/potential.rs:138-151.
template <class R>
inline void mn_force(double amp, double a, double b, const R& x, const R& y, const R& z, R& ax,
                     R& ay, R& az) {
    R r2 = m_pow(x, 2.) + m_pow(y, 2.);
    R r = m_sqrt(r2);
    R z2 = m_pow(z, 2.);
    R b2 = m_pow(R(b), 2.);
    R sqrtz2b2 = m_sqrt(z2 + b2);
    R pyth = m_pow(a + sqrtz2b2, 2.);
    R denom = m_pow(pyth + r2, 3. / 2.);
    R ar = -amp * (r / denom);
    ax = ar * (x / r);
    ay = ar * (y / r);
    az = -amp * (z * (a + sqrtz2b2)) / (sqrtz2b2 * denom);
}

// NFWPotential::force, shared/src/potential.rs:165-175.
template <class R>
inline void nfw_force(double amp, double a, const R& x, const R& y, const R& z, R& ax, R& ay,
                      R& az) {
    R r2 = m_pow(x, 2.) + m_pow(y, 2.) + m_pow(z, 2.);
    R r = m_sqrt(r2);
    R ar = -amp * (m_log(1. + r / a) - r / (a + r)) / r2;
    ax = ar * (x / r);
    ay = ar * (y / r);
    az = ar * (z / r);
}

// Cut-off power-law bulge.  The reference force (shared/src/potential.rs:
// 203-214) reads its radial force from a table (linear interpolation,
// :187-196) whose construction is unfinished (shared/src/model.rs:240-243).
// This restatement keeps the same Cartesian structure and replaces the table
// by the closed form it samples: the enclosed mass of rho = amp r^-alpha
// exp(-(r/rc)^2), M(r) = 2 pi amp rc^(3-alpha) gamma(3/2 - alpha/2, r^2/rc^2),
// and F_r = -M(r)/r^2.
template <class R>
inline void bulge_force(double amp, double alpha, double rc, const R& x, const R& y, const R& z,
                        R& ax, R& ay, R& az) {
    R r2 = m_pow(x, 2.0) + m_pow(y, 2.0) + m_pow(z, 2.0);
    if (m_value(r2) == 0.0) {
        ax = R(0.0);
        ay = R(0.0);
        az = R(0.0);
        return;
    }
    R r = m_sqrt(r2);
    // Parameter-only factor 2 pi amp rc^(3-alpha), hoisted like the table's
    // construction would be.
    const double mass_scale = 6.283185307179586 * amp * ::pow(rc, 3.0 - alpha);
    R mass = mass_scale * m_lower_gamma(1.5 - 0.5 * alpha, r2 / (rc * rc));
    R ar = -mass / r2;
    ax = ar * x / r;
    ay = ar * y / r;
    az = ar * z / r;
}

// BovyPotential::force = bulge + disk + halo summed in that order
// (shared/src/potential.rs:121-123 with the Sum of :64-75).
template <class R>
inline void mw_force(const Potential& p, const R& x, const R& y, const R& z, R& ax, R& ay, R& az) {
    R bx, by, bz, dx, dy, dz, hx, hy, hz;
    bulge_force(p.bulge_amp, p.bulge_alpha, p.bulge_rc, x, y, z, bx, by, bz);
    mn_force(p.disk_amp, p.disk_a, p.disk_b, x, y, z, dx, dy, dz);
    nfw_force(p.halo_amp, p.halo_a, x, y, z, hx, hy, hz);
    R sx = bx + dx, sy = by + dy, sz = bz + dz;
    ax = sx + hx;
    ay = sy + hy;
    az = sz + hz;
}

template <class R>
inline void force(const Potential& p, const R& x, const R& y, const R& z, R& ax, R& ay, R& az) {
    switch (p.kind) {
        case POT_GALPY_KEPLER: galpy_kepler_force(x, y, z, ax, ay, az); break;
        case POT_KEPLER: kepler_force(p.amp, x, y, z, ax, ay, az); break;
        case POT_PLUMMER: plummer_force(p.amp, p.b, x, y, z, ax, ay, az); break;
        case POT_NONE: ax = ay = az = R(0.0); break;
        default: mw_force(p, x, y, z, ax, ay, az); break;
    }
}

// quintic_interp, shared/src/potential.rs:282-288 (c = QuinticCoeff a..f).
template <class R>
inline R quintic_interp(R t, const double* c) {
    R t2 = t * t;
    R t3 = t * t2;
    R t4 = t * t3;
    R t5 = t * t4;
    return c[0] * t5 + c[1] * t4 + c[2] * t3 + c[3] * t2 + c[4] * t + c[5];
}

// CustomOrigin::origins, shared/src/potential.rs:290-331: origin i's position
// at time t.  The reference takes segment floor(t / dt), dt = final_time /
// division (:292-294), evaluates the quintic at t minus the segment INDEX
// (:327-329) and reads row i * n + segment with n = length / division, the
// number of origins (:293-295).  Here the tracks start at t_start (the
// reference's 0), the local time is the segment's own, u - segment with
// u = (t - t_start) / dt, in [0, 1), origin i's segments follow each other
// (row i * n_segments + segment), and a time past the last segment
// extrapolates it where the reference reads past its table.  With t_start = 0,
// dt = 1 and as many origins as segments the two readings are the same.
template <class R>
inline void origin_position(const Potential& p, R t, int32_t i, R* o) {
    const double dt = (p.t_end - p.t_start) / (double)p.n_segments;
    R u = (t - p.t_start) / dt;
    double seg = m_value(m_floor(u));
    if (!(seg >= 0.0)) seg = 0.0;  // `as usize` saturates negative and NaN to 0
    if (seg > (double)(p.n_segments - 1)) seg = (double)(p.n_segments - 1);
    R tau = u - seg;
    const double* row = p.origin_tracks + ((int64_t)i * p.n_segments + (int64_t)seg) * 18;
    for (int k = 0; k < 3; k++) o[k] = quintic_interp(tau, row + 6 * k);
}

// CustomOrigin::force, shared/src/potential.rs:344-356: the origins' forces
// summed in order, each the point-mass (b = 0) or Plummer force of the
// position relative to the origin.
template <class R>
inline void origin_force(const Potential& p, R t, const R& x, const R& y, const R& z, R& ax,
                         R& ay, R& az) {
    R fx = R(0.0), fy = R(0.0), fz = R(0.0);
    for (int32_t i = 0; i < p.n_origins; i++) {
        R o[3];
        origin_position(p, t, i, o);
        const double amp = p.origin_params[2 * i], b = p.origin_params[2 * i + 1];
        R gx, gy, gz;
        if (b == 0.0)
            kepler_force(amp, x - o[0], y - o[1], z - o[2], gx, gy, gz);
        else
            plummer_force(amp, b, x - o[0], y - o[1], z - o[2], gx, gy, gz);
        fx += gx;
        fy += gy;
        fz += gz;
    }
    ax = fx;
    ay = fy;
    az = fz;
}

// Static force plus the moving origins' at time t.
template <class R>
inline void force_at(const Potential& p, R t, const R& x, const R& y, const R& z, R& ax, R& ay,
                     R& az) {
    force(p, x, y, z, ax, ay, az);
    if (p.n_origins > 0) {
        R ox, oy, oz;
        origin_force(p, t, x, y, z, ox, oy, oz);
        ax = ax + ox;
        ay = ay + oy;
        az = az + oz;
    }
}

// Right-hand side in the reference's FuncPtr shape (t, q, a, nargs, args):
// a = (q3, q4, q5, F(t, q0, q1, q2)), src/integrators/galpy/dopr54.rs:59-73.
template <class R>
inline void rhs(const Potential& p, R t, const R* q, R* a) {
    RegionGuard<R> guard(REG_RHS);
    R ax, ay, az;
    force_at(p, t, q[0], q[1], q[2], ax, ay, az);
    a[0] = q[3];
    a[1] = q[4];
    a[2] = q[5];
    a[3] = ax;
    a[4] = ay;
    a[5] = az;
}

// Potential energy per unit mass (not in the reference: its evaluate() is
// commented out, shared/src/potential.rs:29-43).  Used for energy checks.
inline double potential_value(const Potential& p, double x, double y, double z) {
    double r2 = x * x + y * y + z * z;
    double r = ::sqrt(r2);
    switch (p.kind) {
        case POT_GALPY_KEPLER: return -1.0 / r;
        case POT_KEPLER: return -p.amp / r;
        case POT_PLUMMER: return -p.amp / ::sqrt(r2 + p.b * p.b);
        case POT_NONE: return 0.0;
        default: break;
    }
    const double two_pi = 6.283185307179586;
    // bulge: Phi = -M(r)/r - 2 pi amp rc^(2-alpha) Gamma(1 - alpha/2, r^2/rc^2)
    double xr = r2 / (p.bulge_rc * p.bulge_rc);
    double a1 = 1.5 - 0.5 * p.bulge_alpha;
    double a2 = 1.0 - 0.5 * p.bulge_alpha;
    double mass = two_pi * p.bulge_amp * ::pow(p.bulge_rc, 3.0 - p.bulge_alpha) * lower_gamma(a1, xr);
    double phib = -mass / r -
                  two_pi * p.bulge_amp * ::pow(p.bulge_rc, 2.0 - p.bulge_alpha) * upper_gamma(a2, xr);
    double R2 = x * x + y * y;
    double zb = ::sqrt(z * z + p.disk_b * p.disk_b);
    double phid = -p.disk_amp / ::sqrt(R2 + (p.disk_a + zb) * (p.disk_a + zb));
    double phih = -p.halo_amp * ::log(1.0 + r / p.halo_a) / r;
    return phib + phid + phih;
}

// ---------------------------------------------------------------------------
// Per-orbit statistics (the reference keeps none; added without touching the
// arithmetic).
// ---------------------------------------------------------------------------
struct OrbitStats {
    int64_t attempts;     // attempted steps (accepted + rejected)
    int64_t accepted;     // accepted steps
    int64_t rejected;     // rejected steps
    int64_t rhs_evals;    // force evaluations
    int64_t samples;      // output samples written (excluding t[0])
    int64_t dense_steps;  // DOP853: accepted steps that computed dense coefficients
    int32_t status;       // DOPR54 error code (0 = ok); DOP853 always 0
    int32_t reserved;
};

constexpr int DIM = 6;

// ---------------------------------------------------------------------------
// DOPR54 -- src/integrators/galpy/dopr54.rs (galpy bovy_rk.c)
// ---------------------------------------------------------------------------
constexpr double MAX_STEPCHANGE_POWERTWO = 3.0;   // dopr54.rs:32
constexpr double MIN_STEPCHANGE_POWERTWO = -3.0;  // dopr54.rs:33
constexpr double MAX_STEPREDUCE = 10000.0;        // dopr54.rs:34
constexpr double MAX_DT_REDUCE = 10000.0;         // dopr54.rs:35
constexpr double INITIAL_STEP_SENTINEL = -9999.99;  // src/integrators/galpy/mod.rs:8

// rk4_onestep, dopr54.rs:135-208.
template <class R>
inline void rk4_onestep(const Potential& p, const R* yn, R* yn1, R tn, R dt, R* ynk, R* a,
                        OrbitStats& st) {
    rhs(p, tn, yn, a);
    st.rhs_evals++;
    for (int i = 0; i < DIM; i++) yn1[i] += dt * a[i] / 6.0;
    for (int i = 0; i < DIM; i++) ynk[i] = yn[i] + dt * a[i] / 2.0;
    rhs(p, tn + dt / 2.0, ynk, a);
    st.rhs_evals++;
    for (int i = 0; i < DIM; i++) yn1[i] += dt * a[i] / 3.0;
    for (int i = 0; i < DIM; i++) ynk[i] = yn[i] + dt * a[i] / 2.0;
    rhs(p, tn + dt / 2.0, ynk, a);
    st.rhs_evals++;
    for (int i = 0; i < DIM; i++) yn1[i] += dt * a[i] / 3.0;
    for (int i = 0; i < DIM; i++) ynk[i] = yn[i] + dt * a[i];
    rhs(p, tn + dt, ynk, a);
    st.rhs_evals++;
    for (int i = 0; i < DIM; i++) yn1[i] += dt * a[i] / 6.0;
}

// rk4_estimate_step, dopr54.rs:210-368.
template <class R>
inline R rk4_estimate_step(const Potential& p, const R* yo, R dt, const R* t, R rtol, R atol,
                           OrbitStats& st) {
    R err = R(2.0);
    R max_val;
    R to = t[0];
    R init_dt = dt;
    R yn[DIM], y1[DIM], y21[DIM], y2[DIM], ynk[DIM], a[DIM], scale[DIM];

    max_val = m_log(m_fabs(yo[0]));
    for (int i = 1; i < DIM; i++) {
        R v = m_log(m_fabs(yo[i]));
        if (v > max_val) max_val = v;
    }
    R c = m_fmax(atol, rtol + max_val);
    R s = m_log(m_exp(atol - c) + m_exp(rtol + max_val - c)) + c;
    for (int i = 0; i < DIM; i++) scale[i] = s;

    while (err > 1.0) {
        for (int i = 0; i < DIM; i++) yn[i] = yo[i];
        for (int i = 0; i < DIM; i++) y1[i] = yo[i];
        for (int i = 0; i < DIM; i++) y21[i] = yo[i];
        rk4_onestep(p, yn, y1, to, dt, ynk, a, st);
        rk4_onestep(p, yn, y21, to, dt / 2.0, ynk, a, st);
        for (int i = 0; i < DIM; i++) y2[i] = y21[i];
        rk4_onestep(p, y21, y2, to + dt / 2.0, dt / 2.0, ynk, a, st);
        err = R(0.0);
        for (int i = 0; i < DIM; i++) {
            R diff = y1[i] - y2[i];
            R term = m_exp(2.0 * m_log(m_fabs(diff)) - 2.0 * scale[i]);
            err += term;
        }
        err = m_sqrt(err / (double)DIM);
        R factor = m_ceil(m_pow(err, 1.0 / 5.0));
        if (factor > 1.0 && init_dt / dt * factor < MAX_DT_REDUCE) {
            dt /= factor;
        } else {
            break;
        }
    }
    return dt;
}

// dopr54_actualstep, dopr54.rs:462-659.  Returns the new dt_one.
template <class R>
inline R dopr54_actualstep(const Potential& p, R* yo, R dt, R& to, R rtol, R atol, R* a1, R* a,
                           R* k1, R* k2, R* k3, R* k4, R* k5, R* k6, R* yn1, R* yerr, R* ynk,
                           unsigned char accept, OrbitStats& st) {
    const double C2 = 0.2;
    const double C3 = 0.3;
    const double C4 = 0.8;
    const double C5 = 8.0 / 9.0;
    const double A21 = 0.2;
    const double A31 = 3.0 / 40.0;
    const double A41 = 44.0 / 45.0;
    const double A51 = 19372.0 / 6561.0;
    const double A61 = 9017.0 / 3168.0;
    const double A71 = 35.0 / 384.0;
    const double A32 = 9.0 / 40.0;
    const double A42 = -56.0 / 15.0;
    const double A52 = -25360.0 / 2187.0;
    const double A62 = -355.0 / 33.0;
    const double A43 = 32.0 / 9.0;
    const double A53 = 64448.0 / 6561.0;
    const double A63 = 46732.0 / 5247.0;
    const double A73 = 500.0 / 1113.0;
    const double A54 = -212.0 / 729.0;
    const double A64 = 49.0 / 176.0;
    const double A74 = 125.0 / 192.0;
    const double A65 = -5103.0 / 18656.0;
    const double A75 = -2187.0 / 6784.0;
    const double A76 = 11.0 / 84.0;
    const double B1 = 35.0 / 384.0;
    const double B3 = 500.0 / 1113.0;
    const double B4 = 125.0 / 192.0;
    const double B5 = -2187.0 / 6784.0;
    const double B6 = 11.0 / 84.0;
    const double BE1 = B1 - 5179.0 / 57600.0;
    const double BE3 = B3 - 7571.0 / 16695.0;
    const double BE4 = B4 - 393.0 / 640.0;
    const double BE5 = B5 + 92097.0 / 339200.0;
    const double BE6 = B6 - 187.0 / 2100.0;
    const double BE7 = -1.0 / 40.0;

    for (int i = 0; i < DIM; i++) yn1[i] = yo[i];
    for (int i = 0; i < DIM; i++) a[i] = a1[i];
    for (int i = 0; i < DIM; i++) {
        k1[i] = dt * a[i];
        yn1[i] += B1 * k1[i];
        yerr[i] = BE1 * k1[i];
        ynk[i] = yo[i] + A21 * k1[i];
    }
    rhs(p, to + C2 * dt, ynk, a);
    for (int i = 0; i < DIM; i++) {
        k2[i] = dt * a[i];
        ynk[i] = yo[i] + A31 * k1[i] + A32 * k2[i];
    }
    rhs(p, to + C3 * dt, ynk, a);
    for (int i = 0; i < DIM; i++) {
        k3[i] = dt * a[i];
        yn1[i] += B3 * k3[i];
        yerr[i] += BE3 * k3[i];
        ynk[i] = yo[i] + A41 * k1[i] + A42 * k2[i] + A43 * k3[i];
    }
    rhs(p, to + C4 * dt, ynk, a);
    for (int i = 0; i < DIM; i++) {
        k4[i] = dt * a[i];
        yn1[i] += B4 * k4[i];
        yerr[i] += BE4 * k4[i];
        ynk[i] = yo[i] + A51 * k1[i] + A52 * k2[i] + A53 * k3[i] + A54 * k4[i];
    }
    rhs(p, to + C5 * dt, ynk, a);
    for (int i = 0; i < DIM; i++) {
        k5[i] = dt * a[i];
        yn1[i] += B5 * k5[i];
        yerr[i] += BE5 * k5[i];
        ynk[i] = yo[i] + A61 * k1[i] + A62 * k2[i] + A63 * k3[i] + A64 * k4[i] + A65 * k5[i];
    }
    rhs(p, to + dt, ynk, a);
    for (int i = 0; i < DIM; i++) {
        k6[i] = dt * a[i];
        yn1[i] += B6 * k6[i];
        yerr[i] += BE6 * k6[i];
        ynk[i] = yo[i] + A71 * k1[i] + A73 * k3[i] + A74 * k4[i] + A75 * k5[i] + A76 * k6[i];
    }
    rhs(p, to + dt, ynk, a);
    st.rhs_evals += 6;
    for (int i = 0; i < DIM; i++) yerr[i] += BE7 * dt * a[i];

    R max_val = m_log(m_fabs(yo[0]));
    for (int i = 1; i < DIM; i++) {
        R v = m_log(m_fabs(yo[i]));
        if (v > max_val) max_val = v;
    }
    R c = m_fmax(atol, rtol + max_val);
    R s = m_log(m_exp(atol - c) + m_exp(rtol + max_val - c)) + c;

    R err = R(0.0);
    for (int i = 0; i < DIM; i++) err += m_exp(2.0 * m_log(m_fabs(yerr[i])) - 2.0 * s);
    err = m_sqrt(err / (double)DIM);

    R corr = 0.85 * m_pow(err, -0.2);
    R powertwo = m_round(m_log(corr) / m_log(R(2.0)));
    if (powertwo > MAX_STEPCHANGE_POWERTWO) {
        powertwo = R(MAX_STEPCHANGE_POWERTWO);
    } else if (powertwo < MIN_STEPCHANGE_POWERTWO) {
        powertwo = R(MIN_STEPCHANGE_POWERTWO);
    }

    st.attempts++;
    if (powertwo >= 0.0 || accept != 0) {
        for (int i = 0; i < DIM; i++) {
            a1[i] = a[i];
            yo[i] = yn1[i];
        }
        to += dt;
        st.accepted++;
    } else {
        st.rejected++;
    }
    return dt * m_pow(R(2.0), powertwo);
}

// dopr54_onestep, dopr54.rs:370-460.
template <class R>
inline void dopr54_onestep(const Potential& p, R* yo, R dt, R& to, R& dt_one, R rtol, R atol,
                           R* a1, R* a, R* k1, R* k2, R* k3, R* k4, R* k5, R* k6, R* yn1, R* yerr,
                           R* ynk, int& err, OrbitStats& st) {
    R init_dt_one = dt_one;
    R init_to = to;
    unsigned char accept;
    while ((dt >= 0.0 && to < init_to + dt) || (dt < 0.0 && to > init_to + dt)) {
        RegionGuard<R> guard(REG_SUBSTEP);
        accept = 0;
        if (init_dt_one / dt_one > MAX_STEPREDUCE || m_isnan(m_value(dt_one))) {
            dt_one = init_dt_one / MAX_STEPREDUCE;
            accept = 1;
            if (err % 2 == 0) err += 1;
        }
        if (dt >= 0.0 && dt_one > (init_to + dt - to)) dt_one = init_to + dt - to;
        if (dt < 0.0 && dt_one < (init_to + dt - to)) dt_one = init_to + dt - to;
        RegionGuard<R> guard2(REG_ATTEMPT);
        dt_one = dopr54_actualstep(p, yo, dt_one, to, rtol, atol, a1, a, k1, k2, k3, k4, k5, k6,
                                   yn1, yerr, ynk, accept, st);
    }
}

// dopr54, dopr54.rs:670-778, with integrate_kepler's Result mapping
// (:91-124).  result holds nt*6 doubles (orbit-major rows).  Returns err.
template <class R>
inline int dopr54(const Potential& p, R* yo, int nt, R dt_one, const R* t, R rtol, R atol,
                  R* result, OrbitStats& st) {
    std::memset(&st, 0, sizeof(st));
    R a[DIM], a1[DIM], k1[DIM], k2[DIM], k3[DIM], k4[DIM], k5[DIM], k6[DIM], yn[DIM], yn1[DIM],
        yerr[DIM], ynk[DIM];
    int err = 0;
    R to;
    R dt;
    {
        RegionGuard<R> guard(REG_INIT);
        for (int i = 0; i < DIM; i++) result[i] = yo[i];
        for (int i = 0; i < DIM; i++) yn[i] = yo[i];
        dt = t[1] - t[0];
        if (m_value(dt_one) == INITIAL_STEP_SENTINEL)
            dt_one = rk4_estimate_step(p, yo, dt, t, rtol, atol, st);
        to = t[0];
        rhs(p, to, yn, a1);
        st.rhs_evals++;
    }
    for (int ii = 0; ii < nt - 1; ii++) {
        {
            RegionGuard<R> g1(REG_SUBSTEP);
            dopr54_onestep(p, yn, dt, to, dt_one, rtol, atol, a1, a, k1, k2, k3, k4, k5, k6, yn1,
                           yerr, ynk, err, st);
        }
        RegionGuard<R> guard(REG_SAMPLE);
        for (int i = 0; i < DIM; i++) result[(ii + 1) * DIM + i] = yn[i];
        st.samples++;
    }
    st.status = err;
    return err;
}

// ---------------------------------------------------------------------------
// DOP853 -- src/integrators/galpy/dop853.rs (galpy leung_dop853.c, Hairer &
// Wanner's DOP853).  Tableau dop853.rs:105-264.
// ---------------------------------------------------------------------------
namespace dop853c {
constexpr double UROUND = 2.3e-16;
constexpr double C2 = 0.526001519587677318785587544488e-1;
constexpr double C3 = 0.789002279381515978178381316732e-1;
constexpr double C4 = 0.118350341907227396726757197510;
constexpr double C5 = 0.281649658092772603273242802490;
constexpr double C6 = 0.333333333333333333333333333333;
constexpr double C7 = 0.25;
constexpr double C8 = 0.307692307692307692307692307692;
constexpr double C9 = 0.651282051282051282051282051282;
constexpr double C10 = 0.6;
constexpr double C11 = 0.857142857142857142857142857142;
constexpr double C14 = 0.1;
constexpr double C15 = 0.2;
constexpr double C16 = 0.777777777777777777777777777778;
constexpr double A21 = 5.26001519587677318785587544488e-2;
constexpr double A31 = 1.97250569845378994544595329183e-2;
constexpr double A32 = 5.91751709536136983633785987549e-2;
constexpr double A41 = 2.95875854768068491816892993775e-2;
constexpr double A43 = 8.87627564304205475450678981324e-2;
constexpr double A51 = 2.41365134159266685502369798665e-1;
constexpr double A53 = -8.84549479328286085344864962717e-1;
constexpr double A54 = 9.24834003261792003115737966543e-1;
constexpr double A61 = 3.7037037037037037037037037037e-2;
constexpr double A64 = 1.70828608729473871279604482173e-1;
constexpr double A65 = 1.25467687566822425016691814123e-1;
constexpr double A71 = 3.7109375e-2;
constexpr double A74 = 1.70252211019544039314978060272e-1;
constexpr double A75 = 6.02165389804559606850219397283e-2;
constexpr double A76 = -1.7578125e-2;
constexpr double A81 = 3.70920001185047927108779319836e-2;
constexpr double A84 = 1.70383925712239993810214054705e-1;
constexpr double A85 = 1.07262030446373284651809199168e-1;
constexpr double A86 = -1.53194377486244017527936158236e-2;
constexpr double A87 = 8.27378916381402288758473766002e-3;
constexpr double A91 = 6.24110958716075717114429577812e-1;
constexpr double A94 = -3.36089262944694129406857109825e0;
constexpr double A95 = -8.68219346841726006818189891453e-1;
constexpr double A96 = 2.75920996994467083049415600797e1;
constexpr double A97 = 2.01540675504778934086186788979e1;
constexpr double A98 = -4.34898841810699588477366255144e1;
constexpr double A101 = 4.77662536438264365890433908527e-1;
constexpr double A104 = -2.48811461997166764192642586468e0;
constexpr double A105 = -5.90290826836842996371446475743e-1;
constexpr double A106 = 2.12300514481811942347288949897e1;
constexpr double A107 = 1.52792336328824235832596922938e1;
constexpr double A108 = -3.32882109689848629194453265587e1;
constexpr double A109 = -2.03312017085086261358222928593e-2;
constexpr double A111 = -9.3714243008598732571704021658e-1;
constexpr double A114 = 5.18637242884406370830023853209e0;
constexpr double A115 = 1.09143734899672957818500254654e0;
constexpr double A116 = -8.14978701074692612513997267357e0;
constexpr double A117 = -1.85200656599969598641566180701e1;
constexpr double A118 = 2.27394870993505042818970056734e1;
constexpr double A119 = 2.49360555267965238987089396762e0;
constexpr double A1110 = -3.0467644718982195003823669022e0;
constexpr double A121 = 2.27331014751653820792359768449e0;
constexpr double A124 = -1.05344954667372501984066689879e1;
constexpr double A125 = -2.00087205822486249909675718444e0;
constexpr double A126 = -1.79589318631187989172765950534e1;
constexpr double A127 = 2.79488845294199600508499808837e1;
constexpr double A128 = -2.85899827713502369474065508674e0;
constexpr double A129 = -8.87285693353062954433549289258e0;
constexpr double A1210 = 1.23605671757943030647266201528e1;
constexpr double A1211 = 6.43392746015763530355970484046e-1;
constexpr double A141 = 5.61675022830479523392909219681e-2;
constexpr double A147 = 2.53500210216624811088794765333e-1;
constexpr double A148 = -2.46239037470802489917441475441e-1;
constexpr double A149 = -1.24191423263816360469010140626e-1;
constexpr double A1410 = 1.5329179827876569731206322685e-1;
constexpr double A1411 = 8.20105229563468988491666602057e-3;
constexpr double A1412 = 7.56789766054569976138603589584e-3;
constexpr double A1413 = -8.298e-3;
constexpr double A151 = 3.18346481635021405060768473261e-2;
constexpr double A156 = 2.83009096723667755288322961402e-2;
constexpr double A157 = 5.35419883074385676223797384372e-2;
constexpr double A158 = -5.49237485713909884646569340306e-2;
constexpr double A1511 = -1.08347328697249322858509316994e-4;
constexpr double A1512 = 3.82571090835658412954920192323e-4;
constexpr double A1513 = -3.40465008687404560802977114492e-4;
constexpr double A1514 = 1.41312443674632500278074618366e-1;
constexpr double A161 = -4.28896301583791923408573538692e-1;
constexpr double A166 = -4.69762141536116384314449447206e0;
constexpr double A167 = 7.68342119606259904184240953878e0;
constexpr double A168 = 4.06898981839711007970213554331e0;
constexpr double A169 = 3.56727187455281109270669543021e-1;
constexpr double A1613 = -1.39902416515901462129418009734e-3;
constexpr double A1614 = 2.9475147891527723389556272149e0;
constexpr double A1615 = -9.15095847217987001081870187138e0;
constexpr double B1 = 5.42937341165687622380535766363e-2;
constexpr double B6 = 4.45031289275240888144113950566;
constexpr double B7 = 1.89151789931450038304281599044;
constexpr double B8 = -5.8012039600105847814672114227;
constexpr double B9 = 3.1116436695781989440891606237e-1;
constexpr double B10 = -1.52160949662516078556178806805e-1;
constexpr double B11 = 2.01365400804030348374776537501e-1;
constexpr double B12 = 4.47106157277725905176885569043e-2;
constexpr double BHH1 = 0.244094488188976377952755905512;
constexpr double BHH2 = 0.733846688281611857341361741547;
constexpr double BHH3 = 0.220588235294117647058823529412e-1;
constexpr double D41 = -0.84289382761090128651353491142e1;
constexpr double D46 = 0.56671495351937776962531783590;
constexpr double D47 = -0.30689499459498916912797304727e1;
constexpr double D48 = 0.23846676565120698287728149680e1;
constexpr double D49 = 0.21170345824450282767155149946e1;
constexpr double D410 = -0.87139158377797299206789907490;
constexpr double D411 = 0.22404374302607882758541771650e1;
constexpr double D412 = 0.63157877876946881815570249290;
constexpr double D413 = -0.88990336451333310820698117400e-1;
constexpr double D414 = 0.18148505520854727256656404962e2;
constexpr double D415 = -0.91946323924783554000451984436e1;
constexpr double D416 = -0.44360363875948939664310572000e1;
constexpr double D51 = 0.10427508642579134603413151009e2;
constexpr double D56 = 0.24228349177525818288430175319e3;
constexpr double D57 = 0.16520045171727028198505394887e3;
constexpr double D58 = -0.37454675472269020279518312152e3;
constexpr double D59 = -0.22113666853125306036270938578e2;
constexpr double D510 = 0.77334326684722638389603898808e1;
constexpr double D511 = -0.30674084731089398182061213626e2;
constexpr double D512 = -0.93321305264302278729567221706e1;
constexpr double D513 = 0.15697238121770843886131091075e2;
constexpr double D514 = -0.31139403219565177677282850411e2;
constexpr double D515 = -0.93529243588444783865713862664e1;
constexpr double D516 = 0.35816841486394083752465898540e2;
constexpr double D61 = 0.19985053242002433820987653617e2;
constexpr double D66 = -0.38703730874935176555105901742e3;
constexpr double D67 = -0.18917813819516756882830838328e3;
constexpr double D68 = 0.52780815920542364900561016686e3;
constexpr double D69 = -0.11573902539959630126141871134e2;
constexpr double D610 = 0.68812326946963000169666922661e1;
constexpr double D611 = -0.10006050966910838403183860980e1;
constexpr double D612 = 0.77771377980534432092869265740;
constexpr double D613 = -0.27782057523535084065932004339e1;
constexpr double D614 = -0.60196695231264120758267380846e2;
constexpr double D615 = 0.84320405506677161018159903784e2;
constexpr double D616 = 0.11992291136182789328035130030e2;
constexpr double D71 = -0.25693933462703749003312586129e2;
constexpr double D76 = -0.15418974869023643374053993627e3;
constexpr double D77 = -0.23152937917604549567536039109e3;
constexpr double D78 = 0.35763911791061412378285349910e3;
constexpr double D79 = 0.93405324183624310003907691704e2;
constexpr double D710 = -0.37458323136451633156875139351e2;
constexpr double D711 = 0.10409964950896230045147246184e3;
constexpr double D712 = 0.29840293426660503123344363579e2;
constexpr double D713 = -0.43533456590011143754432175058e2;
constexpr double D714 = 0.96324553959188282948394950600e2;
constexpr double D715 = -0.39177261675615439165231486172e2;
constexpr double D716 = -0.14972683625798562581422125276e3;
constexpr double ER1 = 0.1312004499419488073250102996e-1;
constexpr double ER6 = -0.1225156446376204440720569753e1;
constexpr double ER7 = -0.4957589496572501915214079952;
constexpr double ER8 = 0.1664377182454986536961530415e1;
constexpr double ER9 = -0.3503288487499736816886487290;
constexpr double ER10 = 0.3341791187130174790297318841;
constexpr double ER11 = 0.8192320648511571246570742613e-1;
constexpr double ER12 = -0.2235530786388629525884427845e-1;
}  // namespace dop853c

// c_min / c_max / custom_sign, dop853.rs:286-299.
template <class R>
inline R c_min(const R& a, const R& b) { return a < b ? a : b; }
template <class R>
inline R c_max(const R& a, const R& b) { return a > b ? a : b; }
template <class R>
inline R custom_sign(const R& x, const R& y) { return y > 0.0 ? m_fabs(x) : -m_fabs(x); }

// dop853, dop853.rs:316-783.  y0 is updated in place; result holds nt*6.
// The dense-output stages run on every accepted step, as in the reference.
template <class R>
inline void dop853(const Potential& p, R* y0, int nt, const R* t, R rtol_ln, R atol_ln, R* result,
                   OrbitStats& st) {
    using namespace dop853c;
    std::memset(&st, 0, sizeof(st));
    R rtol, atol, hmax, pos_neg, h;
    R facold;
    const double safe = 0.9;
    const double beta = 0.0;
    const double expo1 = 1.0 / 8.0 - beta * 0.2;
    const double facc1 = 1.0 / 0.333;
    const double facc2 = 1.0 / 6.0;
    R yy1[DIM], yy_temp[DIM], k1[DIM], k2[DIM], k3[DIM], k4[DIM], k5[DIM], k6[DIM], k7[DIM],
        k8[DIM], k9[DIM], k10[DIM], rcont1[DIM], rcont2[DIM], rcont3[DIM], rcont4[DIM],
        rcont5[DIM], rcont6[DIM], rcont7[DIM], rcont8[DIM];
    for (int i = 0; i < DIM; i++) {
        yy1[i] = yy_temp[i] = k1[i] = k2[i] = k3[i] = k4[i] = k5[i] = k6[i] = k7[i] = k8[i] =
            k9[i] = k10[i] = R(0.0);
        rcont1[i] = rcont2[i] = rcont3[i] = rcont4[i] = rcont5[i] = rcont6[i] = rcont7[i] =
            rcont8[i] = R(0.0);
    }
    int result_offset;
    {
        RegionGuard<R> guard(REG_INIT);
        rtol = m_exp(rtol_ln);
        atol = m_exp(atol_ln);
        facold = R(1.0e-4);
        hmax = t[nt - 1] - t[0];
        pos_neg = custom_sign(R(1.0), hmax);

        for (int i = 0; i < DIM; i++) result[i] = y0[i];
        result_offset = DIM;

        rhs(p, t[0], y0, k1);
        R dnf = R(0.0);
        R dny = R(0.0);
        for (int i = 0; i < DIM; i++) {
            R sk = atol + rtol * m_fabs(y0[i]);
            R sqr = k1[i] / sk;
            dnf += sqr * sqr;
            sqr = y0[i] / sk;
            dny += sqr * sqr;
        }
        h = custom_sign(c_min(m_sqrt(dny / dnf) * 0.01, m_fabs(hmax)), pos_neg);
        for (int i = 0; i < DIM; i++) k3[i] = y0[i] + h * k1[i];
        rhs(p, t[0] + h, k3, k2);
        R der2 = R(0.0);
        for (int i = 0; i < DIM; i++) {
            R sk = atol + rtol * m_fabs(y0[i]);
            R sqr = (k2[i] - k1[i]) / sk;
            der2 += sqr * sqr;
        }
        der2 = m_sqrt(der2) / h;
        R der12 = c_max(m_fabs(der2), m_sqrt(dnf));
        R h1 = m_pow(0.01 / der12, 1.0 / 8.0);
        h = custom_sign(c_min(100.0 * m_fabs(h), c_min(m_fabs(h1), m_fabs(hmax))), pos_neg);
        st.rhs_evals += 2;
    }

    int reject = 0;
    R t_current = t[0];
    R t_old = t[0];
    R t_old_older;
    int finished_user_t_ii = 0;

    while (finished_user_t_ii < nt - 1) {
        RegionGuard<R> guard(REG_ATTEMPT);
        h = pos_neg * c_max(m_fabs(h), R(1e3 * UROUND));

        for (int i = 0; i < DIM; i++) yy1[i] = y0[i] + h * A21 * k1[i];
        rhs(p, t_current + C2 * h, yy1, k2);
        for (int i = 0; i < DIM; i++) yy1[i] = y0[i] + h * (A31 * k1[i] + A32 * k2[i]);
        rhs(p, t_current + C3 * h, yy1, k3);
        for (int i = 0; i < DIM; i++) yy1[i] = y0[i] + h * (A41 * k1[i] + A43 * k3[i]);
        rhs(p, t_current + C4 * h, yy1, k4);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A51 * k1[i] + A53 * k3[i] + A54 * k4[i]);
        rhs(p, t_current + C5 * h, yy1, k5);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A61 * k1[i] + A64 * k4[i] + A65 * k5[i]);
        rhs(p, t_current + C6 * h, yy1, k6);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A71 * k1[i] + A74 * k4[i] + A75 * k5[i] + A76 * k6[i]);
        rhs(p, t_current + C7 * h, yy1, k7);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] +
                     h * (A81 * k1[i] + A84 * k4[i] + A85 * k5[i] + A86 * k6[i] + A87 * k7[i]);
        rhs(p, t_current + C8 * h, yy1, k8);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A91 * k1[i] + A94 * k4[i] + A95 * k5[i] + A96 * k6[i] +
                                  A97 * k7[i] + A98 * k8[i]);
        rhs(p, t_current + C9 * h, yy1, k9);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A101 * k1[i] + A104 * k4[i] + A105 * k5[i] + A106 * k6[i] +
                                  A107 * k7[i] + A108 * k8[i] + A109 * k9[i]);
        rhs(p, t_current + C10 * h, yy1, k10);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A111 * k1[i] + A114 * k4[i] + A115 * k5[i] + A116 * k6[i] +
                                  A117 * k7[i] + A118 * k8[i] + A119 * k9[i] + A1110 * k10[i]);
        rhs(p, t_current + C11 * h, yy1, k2);
        for (int i = 0; i < DIM; i++)
            yy1[i] = y0[i] + h * (A121 * k1[i] + A124 * k4[i] + A125 * k5[i] + A126 * k6[i] +
                                  A127 * k7[i] + A128 * k8[i] + A129 * k9[i] + A1210 * k10[i] +
                                  A1211 * k2[i]);
        t_old_older = t_old;
        t_old = t_current;
        t_current = t_current + h;
        rhs(p, t_current, yy1, k3);
        st.rhs_evals += 11;

        for (int i = 0; i < DIM; i++) {
            k4[i] = B1 * k1[i] + B6 * k6[i] + B7 * k7[i] + B8 * k8[i] + B9 * k9[i] + B10 * k10[i] +
                    B11 * k2[i] + B12 * k3[i];
            k5[i] = y0[i] + h * k4[i];
        }
        R err = R(0.0);
        R err2 = R(0.0);
        for (int i = 0; i < DIM; i++) {
            R sk = atol + rtol * c_max(m_fabs(y0[i]), m_fabs(k5[i]));
            R erri = k4[i] - BHH1 * k1[i] - BHH2 * k9[i] - BHH3 * k3[i];
            R sqr = erri / sk;
            err2 += sqr * sqr;
            erri = ER1 * k1[i] + ER6 * k6[i] + ER7 * k7[i] + ER8 * k8[i] + ER9 * k9[i] +
                   ER10 * k10[i] + ER11 * k2[i] + ER12 * k3[i];
            sqr = erri / sk;
            err += sqr * sqr;
        }
        R deno = err + 0.01 * err2;
        if (deno <= 0.0) deno = R(1.0);
        err = m_fabs(h) * err * m_sqrt(1.0 / (deno * (double)DIM));

        R fac11 = m_pow(err, expo1);
        R fac = fac11 / m_pow(facold, beta);
        fac = c_max(R(facc2), c_min(R(facc1), fac / safe));
        R hnew = h / fac;
        st.attempts++;

        if (err <= 1.0) {
            st.accepted++;
            {
                RegionGuard<R> g2(REG_ACCEPT);
                facold = c_max(err, R(1.0e-4));
                rhs(p, t_current, k5, k4);
                st.rhs_evals += 1;
                for (int i = 0; i < DIM; i++) {
                    rcont1[i] = y0[i];
                    R ydiff = k5[i] - y0[i];
                    rcont2[i] = ydiff;
                    R bspl = h * k1[i] - ydiff;
                    rcont3[i] = bspl;
                    rcont4[i] = ydiff - h * k4[i] - bspl;
                }
            }
            {
                RegionGuard<R> g3(REG_DENSE);
                st.dense_steps++;
                for (int i = 0; i < DIM; i++) {
                    rcont5[i] = D41 * k1[i] + D46 * k6[i] + D47 * k7[i] + D48 * k8[i] +
                                D49 * k9[i] + D410 * k10[i] + D411 * k2[i] + D412 * k3[i];
                    rcont6[i] = D51 * k1[i] + D56 * k6[i] + D57 * k7[i] + D58 * k8[i] +
                                D59 * k9[i] + D510 * k10[i] + D511 * k2[i] + D512 * k3[i];
                    rcont7[i] = D61 * k1[i] + D66 * k6[i] + D67 * k7[i] + D68 * k8[i] +
                                D69 * k9[i] + D610 * k10[i] + D611 * k2[i] + D612 * k3[i];
                    rcont8[i] = D71 * k1[i] + D76 * k6[i] + D77 * k7[i] + D78 * k8[i] +
                                D79 * k9[i] + D710 * k10[i] + D711 * k2[i] + D712 * k3[i];
                }
                for (int i = 0; i < DIM; i++)
                    yy1[i] = y0[i] + h * (A141 * k1[i] + A147 * k7[i] + A148 * k8[i] +
                                          A149 * k9[i] + A1410 * k10[i] + A1411 * k2[i] +
                                          A1412 * k3[i] + A1413 * k4[i]);
                rhs(p, t_old + C14 * h, yy1, k10);
                for (int i = 0; i < DIM; i++)
                    yy1[i] = y0[i] + h * (A151 * k1[i] + A156 * k6[i] + A157 * k7[i] +
                                          A158 * k8[i] + A1511 * k2[i] + A1512 * k3[i] +
                                          A1513 * k4[i] + A1514 * k10[i]);
                rhs(p, t_old + C15 * h, yy1, k2);
                for (int i = 0; i < DIM; i++)
                    yy1[i] = y0[i] + h * (A161 * k1[i] + A166 * k6[i] + A167 * k7[i] +
                                          A168 * k8[i] + A169 * k9[i] + A1613 * k4[i] +
                                          A1614 * k10[i] + A1615 * k2[i]);
                rhs(p, t_old + C16 * h, yy1, k3);
                st.rhs_evals += 3;
                for (int i = 0; i < DIM; i++) {
                    rcont5[i] = h * (rcont5[i] + D413 * k4[i] + D414 * k10[i] + D415 * k2[i] +
                                     D416 * k3[i]);
                    rcont6[i] = h * (rcont6[i] + D513 * k4[i] + D514 * k10[i] + D515 * k2[i] +
                                     D516 * k3[i]);
                    rcont7[i] = h * (rcont7[i] + D613 * k4[i] + D614 * k10[i] + D615 * k2[i] +
                                     D616 * k3[i]);
                    rcont8[i] = h * (rcont8[i] + D713 * k4[i] + D714 * k10[i] + D715 * k2[i] +
                                     D716 * k3[i]);
                }
            }
            for (int i = 0; i < DIM; i++) {
                k1[i] = k4[i];
                y0[i] = k5[i];
            }
            while (finished_user_t_ii < nt - 1 &&
                   pos_neg * t[finished_user_t_ii + 1] < pos_neg * t_current) {
                RegionGuard<R> g4(REG_SAMPLE);
                R s = (t[finished_user_t_ii + 1] - t_old) / h;
                R s1 = 1.0 - s;
                for (int i = 0; i < DIM; i++)
                    yy_temp[i] =
                        rcont1[i] +
                        s * (rcont2[i] +
                             s1 * (rcont3[i] +
                                   s * (rcont4[i] +
                                        s1 * (rcont5[i] +
                                              s * (rcont6[i] + s1 * (rcont7[i] + s * rcont8[i]))))));
                for (int i = 0; i < DIM; i++) result[result_offset + i] = yy_temp[i];
                result_offset += DIM;
                finished_user_t_ii += 1;
                st.samples++;
            }
            hnew = m_fabs(hnew) > m_fabs(hmax) ? pos_neg * hmax : hnew;
            if (reject != 0) hnew = pos_neg * c_min(m_fabs(hnew), m_fabs(h));
            reject = 0;
        } else {
            st.rejected++;
            hnew = h / c_min(R(facc1), fac11 / safe);
            reject = 1;
            t_current = t_old;
            t_old = t_old_older;
        }
        h = hnew;
    }
}

// OutputGrid::sample, shared/src/config.rs:57-66.
inline double grid_sample(double start, double end, int64_t steps, int64_t index) {
    if (index == steps - 1) return end;
    return start + (end - start) * (double)index / (double)(steps - 1);
}

}  // namespace drift_oracle
