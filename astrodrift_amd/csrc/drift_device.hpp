// drift_device.hpp -- scalar building blocks shared by the gfx950 kernels.
//
// Everything here is per-lane scalar code: one orbit per lane, its phase-space
// state held in registers.  Functions are DRIFT_HD so that the host-side table
// builder and the test harness can compile the same formulas with g++.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DRIFT_HD __host__ __device__
#define DRIFT_FI __forceinline__
#else
#define DRIFT_HD
#define DRIFT_FI inline
#endif

namespace drift {

constexpr int DIM = 6;

// Potential kinds; values match drift_potential_kind in include/drift_hip.h.
enum : int32_t {
    POT_GALPY_KEPLER = 0,
    POT_KEPLER = 1,
    POT_PLUMMER = 2,
    POT_MW = 3,
};

// Spherical-part table of the Milky-Way model (bulge + NFW halo):
//   a_sph(r) = -S(r^2) * r_vec,
// S(r2) = M_bulge(r)/r^3 + amp_h (ln(1 + r/a_h) - r/(a_h + r)) / r^3.
// S is tabulated on half-octave segments of r2 in [2^E_MIN, 2^E_MAX): segment
// index from the binary exponent of r2 and whether its mantissa is >= sqrt(2),
// TABLE_DEG+1 monomial coefficients per segment in the local variable
// s = (m - centre) * inv_half in [-1, 1], m = r2 / 2^e in [1, 2).
constexpr int TABLE_E_MIN = -64;
constexpr int TABLE_E_MAX = 32;
constexpr int TABLE_DEG = 15;
constexpr int TABLE_NCOEF = TABLE_DEG + 1;
constexpr int TABLE_SEGMENTS = 2 * (TABLE_E_MAX - TABLE_E_MIN);
constexpr int TABLE_DOUBLES = TABLE_SEGMENTS * TABLE_NCOEF;

// Potential parameters in the form the kernels use (derived on the host
// from drift_potential).
struct PotParams {
    int32_t kind;
    int32_t has_table;
    double amp;       // Kepler / Plummer amplitude
    double b2;        // Plummer b^2
    double disk_amp, disk_a, disk_b2;
    // spherical part outside the table range (slow path)
    double bulge_mass_total;   // 2 pi amp rc^(3-alpha) Gamma(3/2 - alpha/2)
    double bulge_small_c0;     // S_b(r2) ~ r2^(-alpha/2) * (c0 + c1 * r2) for small r2
    double bulge_small_c1;
    double bulge_exponent;     // -alpha/2
    double halo_amp, halo_a;
};

DRIFT_HD DRIFT_FI double rsqrt_d(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return rsqrt(x);
#else
    return 1.0 / std::sqrt(x);
#endif
}

DRIFT_HD DRIFT_FI float rsqrt_f(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / std::sqrt(x);
#endif
}

// Segment index and local coordinate of r2 in the spherical table.  Returns
// false when r2 is outside [2^E_MIN, 2^E_MAX) (or not finite).
DRIFT_HD DRIFT_FI bool table_locate(double r2, int& seg, double& s) {
    int e;
    double m = std::frexp(r2, &e);  // r2 = m 2^e, m in [0.5, 1)
    m *= 2.0;                        // m in [1, 2)
    e -= 1;
    if (!(e >= TABLE_E_MIN && e < TABLE_E_MAX)) return false;
    const double kSqrt2 = 1.4142135623730951;
    int half = m >= kSqrt2 ? 1 : 0;
    // half 0: [1, sqrt2), half 1: [sqrt2, 2)
    const double lo = half ? kSqrt2 : 1.0;
    const double hi = half ? 2.0 : kSqrt2;
    const double centre = 0.5 * (lo + hi);
    const double inv_half = 2.0 / (hi - lo);
    s = (m - centre) * inv_half;
    seg = 2 * (e - TABLE_E_MIN) + half;
    return true;
}

template <class T>
DRIFT_HD DRIFT_FI T horner_table(const T* c, T s) {
    T acc = c[TABLE_DEG];
#pragma unroll
    for (int k = TABLE_DEG - 1; k >= 0; --k) acc = acc * s + c[k];
    return acc;
}

// ln x for finite x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(z) = 2 (z + z^3/3 + ...), z = (m - 1)/(m + 1), |z| < 0.172,
// 12 odd terms (truncation < 1e-19); ln 2 split in two parts.  Small in
// registers, for the slow path below (the hot path uses the table).
DRIFT_HD DRIFT_FI double light_log(double x) {
    int e;
    double m = std::frexp(x, &e);  // [0.5, 1)
    if (m < 0.70710678118654752) {
        m *= 2.0;
        e -= 1;
    }
    const double z = (m - 1.0) / (m + 1.0);
    const double z2 = z * z;
    double p = 1.0 / 23.0;
#pragma unroll
    for (int k = 10; k >= 0; --k) p = p * z2 + 1.0 / (2 * k + 1);
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    return (double)e * ln2_hi + ((double)e * ln2_lo + 2.0 * z * p);
}

// 2^y: 2^n e^(f ln 2), n = round(y), |f| <= 1/2, Taylor series of e^g to
// degree 16 (|g| <= 0.35, truncation < 1e-19).
DRIFT_HD DRIFT_FI double light_exp2(double y) {
    const double n = std::rint(y);
    const double g = (y - n) * 0.69314718055994530942;
    double p = 1.0 / 20922789888000.0;  // 1/16!
    double fact = 20922789888000.0;
#pragma unroll
    for (int k = 15; k >= 0; --k) {
        fact /= (double)(k + 1);
        p = p * g + 1.0 / fact;
    }
    return std::ldexp(p, (int)n);
}

// Direct (slow-path) evaluation of S(r2) outside the tabulated range
// [2^E_MIN, 2^E_MAX), i.e. r < 2.3e-10 or r >= 65536.
DRIFT_HD DRIFT_FI double spherical_direct(const PotParams& p, double r2) {
    if (!(r2 > 0.0)) return 0.0;
    double ir = rsqrt_d(r2);
    double r = r2 * ir;
    double ir3 = ir * ir * ir;
    double sb, sh;
    if (r2 < 1.0) {
        // leading terms of the small-radius expansions of both components
        sb = light_exp2(p.bulge_exponent * light_log(r2) * 1.44269504088896340736) *
             (p.bulge_small_c0 + p.bulge_small_c1 * r2);
        double ia = 1.0 / p.halo_a;
        sh = p.halo_amp * (0.5 * ia * ia * ir - (2.0 / 3.0) * ia * ia * ia + 0.75 * r * ia * ia * ia * ia);
    } else {
        // the bulge has its total mass; u = r/a_h is large, no cancellation
        sb = p.bulge_mass_total * ir3;
        double u = r / p.halo_a;
        sh = p.halo_amp * (light_log(1.0 + u) - u / (1.0 + u)) * ir3;
    }
    return sb + sh;
}

// ---------------------------------------------------------------------------
// Accelerations.  T is the arithmetic type of the force evaluation (double for
// FP64, float for the mixed-precision mode).  All forms are the Cartesian
// equivalents of the reference formulas (shared/src/potential.rs).
// ---------------------------------------------------------------------------

// galpy Kepler (src/integrators/galpy/dopr54.rs:75-88): a = -q / |q|^3, amp 1.
// The reference divides by the cylindrical radius; at R = 0 its x/y
// accelerations are NaN, which is reproduced.
template <class T>
DRIFT_HD DRIFT_FI void galpy_kepler_acc(T x, T y, T z, T& ax, T& ay, T& az) {
    T R2 = x * x + y * y;
    T r2 = R2 + z * z;
    T ir;
    if constexpr (sizeof(T) == 8) ir = rsqrt_d(r2); else ir = rsqrt_f(r2);
    T f = -(ir * ir * ir);
    ax = f * x;
    ay = f * y;
    az = f * z;
    if (R2 == T(0)) {
        ax = T(NAN);
        ay = T(NAN);
    }
}

// KeplerPotential (shared/src/potential.rs:229-238).
template <class T>
DRIFT_HD DRIFT_FI void kepler_acc(T amp, T x, T y, T z, T& ax, T& ay, T& az) {
    T r2 = x * x + y * y + z * z;
    T ir;
    if constexpr (sizeof(T) == 8) ir = rsqrt_d(r2); else ir = rsqrt_f(r2);
    T f = -amp * (ir * ir * ir);
    ax = f * x;
    ay = f * y;
    az = f * z;
}

// PlummerPotential (shared/src/potential.rs:254-261).
template <class T>
DRIFT_HD DRIFT_FI void plummer_acc(T amp, T b2, T x, T y, T z, T& ax, T& ay, T& az) {
    T r2 = x * x + y * y + z * z;
    T ir;
    if constexpr (sizeof(T) == 8) ir = rsqrt_d(r2 + b2); else ir = rsqrt_f(r2 + b2);
    T f = -amp * (ir * ir * ir);
    ax = f * x;
    ay = f * y;
    az = f * z;
}

// Miyamoto-Nagai disk (shared/src/potential.rs:138-151):
// a_R = -amp R / D^3, a_z = -amp z (a + zb) / (zb D^3),
// zb = sqrt(z^2 + b^2), D^2 = R^2 + (a + zb)^2.
template <class T>
DRIFT_HD DRIFT_FI void mn_acc(T amp, T a, T b2, T x, T y, T z, T& ax, T& ay, T& az) {
    T R2 = x * x + y * y;
    T zb2 = z * z + b2;
    T izb, iD;
    if constexpr (sizeof(T) == 8) izb = rsqrt_d(zb2); else izb = rsqrt_f(zb2);
    T azb = a + zb2 * izb;
    T D2 = azb * azb + R2;
    if constexpr (sizeof(T) == 8) iD = rsqrt_d(D2); else iD = rsqrt_f(D2);
    T f = -amp * (iD * iD * iD);
    ax = f * x;
    ay = f * y;
    az = f * z * azb * izb;
}

}  // namespace drift
