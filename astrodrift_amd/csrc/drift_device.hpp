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
// The LDS table covers [2^-64, 2^32) (r from 2.3e-10 to 65,536).  A second,
// "far" table of the same layout covers every positive finite double r2,
// [2^-1074, 2^1024), and stays in global memory: it is read only when an
// orbit leaves the LDS range, with the same few registers as the LDS path.
constexpr int TABLE_E_MIN = -64;
constexpr int TABLE_E_MAX = 32;
constexpr int TABLE_DEG = 15;
constexpr int TABLE_NCOEF = TABLE_DEG + 1;
constexpr int TABLE_SEGMENTS = 2 * (TABLE_E_MAX - TABLE_E_MIN);
constexpr int TABLE_DOUBLES = TABLE_SEGMENTS * TABLE_NCOEF;
constexpr int FAR_E_MIN = -1074;
constexpr int FAR_E_MAX = 1024;
constexpr int FAR_SEGMENTS = 2 * (FAR_E_MAX - FAR_E_MIN);
constexpr int FAR_DOUBLES = FAR_SEGMENTS * TABLE_NCOEF;

// Potential parameters in the form the kernels use (derived on the host
// from drift_potential).
struct PotParams {
    int32_t kind;
    int32_t has_table;
    double amp;       // Kepler / Plummer amplitude
    double b2;        // Plummer b^2
    double disk_amp, disk_a, disk_b2;
    const double* far_table;  // MW: FAR_DOUBLES coefficients in device memory
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

// Segment index and local coordinate of r2 in a table of segments
// [2^E0, 2^E1).  Returns false when r2 is outside it.
template <int E0, int E1>
DRIFT_HD DRIFT_FI bool table_locate_in(double r2, int& seg, double& s) {
    int e;
    double m = std::frexp(r2, &e);  // r2 = m 2^e, m in [0.5, 1)
    m *= 2.0;                        // m in [1, 2)
    e -= 1;
    if (!(e >= E0 && e < E1)) return false;
    const double kSqrt2 = 1.4142135623730951;
    int half = m >= kSqrt2 ? 1 : 0;
    // half 0: [1, sqrt2), half 1: [sqrt2, 2)
    const double lo = half ? kSqrt2 : 1.0;
    const double hi = half ? 2.0 : kSqrt2;
    const double centre = 0.5 * (lo + hi);
    const double inv_half = 2.0 / (hi - lo);
    s = (m - centre) * inv_half;
    seg = 2 * (e - E0) + half;
    return true;
}

// The LDS table's range (false outside [2^E_MIN, 2^E_MAX) or not finite).
DRIFT_HD DRIFT_FI bool table_locate(double r2, int& seg, double& s) {
    return table_locate_in<TABLE_E_MIN, TABLE_E_MAX>(r2, seg, s);
}

template <class T>
DRIFT_HD DRIFT_FI T horner_table(const T* c, T s) {
    T acc = c[TABLE_DEG];
#pragma unroll
    for (int k = TABLE_DEG - 1; k >= 0; --k) acc = acc * s + c[k];
    return acc;
}

// S(r2) from the far table, for r2 outside the LDS table: any positive
// finite r2.  At r2 = 0 the spherical force S r_vec vanishes with r_vec, so
// S(0) is taken as 0; a non-finite r2 gives NaN.
DRIFT_HD DRIFT_FI double far_spherical(const PotParams& p, double r2) {
    if (!(r2 > 0.0)) return r2 == 0.0 ? 0.0 : r2;
    int seg;
    double s;
    if (!table_locate_in<FAR_E_MIN, FAR_E_MAX>(r2, seg, s)) return NAN;  // r2 = inf
    return horner_table(p.far_table + seg * TABLE_NCOEF, s);
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
