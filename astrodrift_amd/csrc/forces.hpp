// forces.hpp -- right-hand sides a = (v, F(x)) for the gfx950 integrators.
//
// Each force is a small value type whose operator() fills a[6] from the time t
// and q[6], in the shape of the reference's right-hand side FuncPtr(t, q, a)
// (src/integrators/galpy/dop853.rs:23-31); the integrator cores are templated
// on it so the force is inlined into the step with no indirect calls, and the
// stage times are dead code for the static forces, which ignore t.  FT is the arithmetic type of the force itself:
// double for the FP64 path, float for the mixed-precision path (positions are
// rounded to float, accelerations widened back to double).
#pragma once

#include "drift_device.hpp"

namespace drift {

template <class FT>
struct GalpyKeplerForce {
    // a point mass: few registers (the DOPR54 step estimate runs out of line)
    static constexpr bool kFewRegisters = true;
    DRIFT_HD DRIFT_FI void operator()(double, const double* q, double* a) const {
        FT ax, ay, az;
        galpy_kepler_acc<FT>(FT(q[0]), FT(q[1]), FT(q[2]), ax, ay, az);
        a[0] = q[3];
        a[1] = q[4];
        a[2] = q[5];
        a[3] = double(ax);
        a[4] = double(ay);
        a[5] = double(az);
    }
};

template <class FT>
struct KeplerForce {
    // a point mass: few registers (the DOPR54 step estimate runs out of line)
    static constexpr bool kFewRegisters = true;
    FT amp;
    DRIFT_HD DRIFT_FI void operator()(double, const double* q, double* a) const {
        FT ax, ay, az;
        kepler_acc<FT>(amp, FT(q[0]), FT(q[1]), FT(q[2]), ax, ay, az);
        a[0] = q[3];
        a[1] = q[4];
        a[2] = q[5];
        a[3] = double(ax);
        a[4] = double(ay);
        a[5] = double(az);
    }
};

template <class FT>
struct PlummerForce {
    // a point mass: few registers (the DOPR54 step estimate runs out of line)
    static constexpr bool kFewRegisters = true;
    FT amp, b2;
    DRIFT_HD DRIFT_FI void operator()(double, const double* q, double* a) const {
        FT ax, ay, az;
        plummer_acc<FT>(amp, b2, FT(q[0]), FT(q[1]), FT(q[2]), ax, ay, az);
        a[0] = q[3];
        a[1] = q[4];
        a[2] = q[5];
        a[3] = double(ax);
        a[4] = double(ay);
        a[5] = double(az);
    }
};

// Milky-Way model: Miyamoto-Nagai disk in closed form plus the tabulated
// spherical part (bulge + NFW halo), see spherical_table.cpp.  `table` points
// at the LDS table: TABLE_DOUBLES double coefficients, or for FT = float the
// TABLE32_FLOATS of the cubic float table.
template <class FT>
struct MwForce {
    const PotParams* p;  // far-table pointer (outside the LDS table's range)
    const FT* table;
    FT disk_amp, disk_a, disk_b2;
    DRIFT_HD DRIFT_FI void operator()(double, const double* q, double* a) const {
        const FT x = FT(q[0]), y = FT(q[1]), z = FT(q[2]);
        // Miyamoto-Nagai disk (shared/src/potential.rs:138-151), in the
        // factored form a_R = f R, a_z = f z (a + zb) / zb with
        // f = -amp / D^3, D^2 = R^2 + (a + zb)^2, zb = sqrt(z^2 + b^2)
        const FT R2 = x * x + y * y;
        const FT zb2 = std::fma(z, z, disk_b2);
        FT izb, iD;
        if constexpr (sizeof(FT) == 8) izb = rsqrt_pos(zb2); else izb = rsqrt_f(zb2);
        const FT azb = disk_a + zb2 * izb;
        const FT D2 = azb * azb + R2;
        if constexpr (sizeof(FT) == 8) iD = rsqrt_pos(D2); else iD = rsqrt_f(D2);
        const FT f = -disk_amp * (iD * iD * iD);
        // spherical part (bulge + halo), -S r_vec, table lookup on r2 in double
        double r2;
        if constexpr (sizeof(FT) == 8) r2 = z * z + R2; else r2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
        int seg;
        double s;
        const bool in_lds = table_locate(r2, seg, s);
        // every lane evaluates the LDS table (a lane outside its range reads
        // the last segment), and only a lane outside it branches to the far
        // table, so the common path carries no divergent branch
        const uint32_t last_row = TABLE_SEGMENTS - 1;
        const uint32_t row = (uint32_t)seg < last_row ? (uint32_t)seg : last_row;
        FT S;
        if constexpr (sizeof(FT) == 8)
            S = horner_table<FT>(table + row * TABLE_NCOEF, FT(s));
        else
            S = horner_table<FT, TABLE32_NCOEF>(table + row * TABLE32_STRIDE, FT(s));
        if (!in_lds) S = FT(far_spherical(*p, r2));
        // both parts are multiples of (x, y) in the plane: a = (f - S) (x, y)
        const FT g = f - S;
        a[0] = q[3];
        a[1] = q[4];
        a[2] = q[5];
        a[3] = double(g * x);
        a[4] = double(g * y);
        a[5] = double((f * (azb * izb) - S) * z);
    }
};

// No static force: the background of a potential made of moving origins only
// (or free motion).
struct NoForce {
    // a point mass: few registers (the DOPR54 step estimate runs out of line)
    static constexpr bool kFewRegisters = true;
    DRIFT_HD DRIFT_FI void operator()(double, const double* q, double* a) const {
        a[0] = q[3];
        a[1] = q[4];
        a[2] = q[5];
        a[3] = 0.0;
        a[4] = 0.0;
        a[5] = 0.0;
    }
};

// Static force plus moving origins, after CustomOrigin::force
// (shared/src/potential.rs:334-356): the force of a point mass (b = 0,
// KeplerPotential) or Plummer sphere centred on each origin's position o_i(t),
// a = sum_i F_i(x - o_i(t)) summed in origin order, then added to the static
// force.  The
// origins follow quintic tracks over n_segments equal time segments from
// t_start (CustomOrigin::origins, :290-331): segment s = floor(u) of
// u = (t - t_start) inv_dt, clamped to [0, n_segments - 1] so a time outside
// the tabulated span extrapolates its end segment, at local time tau = u - s.
// FP64 only.
template <class Base>
struct MovingForce {
    Base base;
    const double* tracks;
    const double* origin_params;
    int32_t n_origins, n_segments;
    double t_start, inv_dt;
    // Origin 0's track row of the segment this lane last evaluated: a lane's
    // stage times stay inside one segment for many force evaluations, so the
    // row is fetched again only when the segment changes, and the fetch is
    // issued before the static force, whose evaluation hides its latency.
    mutable int64_t row_seg = -1;
    mutable double row[ORIGIN_COEFS];
    // A window of the tracks staged in LDS by a time-windowed launch: the
    // rows of segments win_lo .. win_lo + win_n - 1 of every origin, origin
    // i's segment win_lo + s at win_rows[(i * win_n + s) * ORIGIN_COEFS]
    // (win_n = 0: no window).  A stage time inside it reads its rows there,
    // one outside it from the tracks; the arithmetic is the same.
    const double* win_rows = nullptr;
    int32_t win_lo = 0, win_n = 0;

    DRIFT_HD DRIFT_FI static double quintic(const double* p, double tau) {
        // (tau^5 .. tau^0) coefficients
        return fma(fma(fma(fma(fma(p[0], tau, p[1]), tau, p[2]), tau, p[3]), tau, p[4]), tau, p[5]);
    }

    DRIFT_HD DRIFT_FI void add_origin(const double* c, int32_t i, const double* q, double tau,
                                      double& ax, double& ay, double& az) const {
        const double* pb =
            static_cast<const double*>(__builtin_assume_aligned(origin_params + 2 * i, 16));
        double fx, fy, fz;
        plummer_acc<double>(pb[0], pb[1], q[0] - quintic(c, tau), q[1] - quintic(c + 6, tau),
                            q[2] - quintic(c + 12, tau), fx, fy, fz);
        ax += fx;
        ay += fy;
        az += fz;
    }

    DRIFT_HD DRIFT_FI void operator()(double t, const double* q, double* a) const {
        const double u = (t - t_start) * inv_dt;
        double sf = std::floor(u);
        if (!(sf >= 0.0)) sf = 0.0;  // also a NaN time: its tau stays NaN
        const double last = (double)(n_segments - 1);
        if (sf > last) sf = last;
        const double tau = u - sf;
        const int64_t seg = (int64_t)sf;
        const int64_t ws = seg - win_lo;
        const bool in_win = (uint64_t)ws < (uint64_t)win_n;
        if (n_origins > 0 && seg != row_seg) {
            // rows are 144 bytes from a 256-byte aligned base: 16-byte loads
            // (LDS and global memory on separate paths, no generic loads)
            if (in_win) {
                const double* c = static_cast<const double*>(
                    __builtin_assume_aligned(win_rows + ws * ORIGIN_COEFS, 16));
#pragma unroll
                for (int k = 0; k < ORIGIN_COEFS; k++) row[k] = c[k];
            } else {
                const double* c = static_cast<const double*>(
                    __builtin_assume_aligned(tracks + seg * ORIGIN_COEFS, 16));
#pragma unroll
                for (int k = 0; k < ORIGIN_COEFS; k++) row[k] = c[k];
            }
            row_seg = seg;
        }
        base(t, q, a);
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (n_origins > 0) {
            const double* pb = static_cast<const double*>(__builtin_assume_aligned(origin_params, 16));
            plummer_acc<double>(pb[0], pb[1], q[0] - quintic(row, tau), q[1] - quintic(row + 6, tau),
                                q[2] - quintic(row + 12, tau), ax, ay, az);
        }
        if (in_win) {
            for (int32_t i = 1; i < n_origins; ++i)
                add_origin(static_cast<const double*>(__builtin_assume_aligned(
                               win_rows + ((int64_t)i * win_n + ws) * ORIGIN_COEFS, 16)),
                           i, q, tau, ax, ay, az);
        } else {
            for (int32_t i = 1; i < n_origins; ++i)
                add_origin(static_cast<const double*>(__builtin_assume_aligned(
                               tracks + ((int64_t)i * n_segments + seg) * ORIGIN_COEFS, 16)),
                           i, q, tau, ax, ay, az);
        }
        a[3] += ax;
        a[4] += ay;
        a[5] += az;
    }
};

}  // namespace drift
