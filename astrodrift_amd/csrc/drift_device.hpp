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

// Potential kinds; values 0..4 match drift_potential_kind in include/drift_hip.h.
// POT_MOVING is the kernels' own flag: the static kind plus a set of moving
// origins (point masses or Plummer spheres on tabulated tracks).
enum : int32_t {
    POT_GALPY_KEPLER = 0,
    POT_KEPLER = 1,
    POT_PLUMMER = 2,
    POT_MW = 3,
    POT_NONE = 4,
    POT_MOVING = 8,
};

// Static kind of a kernel kind (without the moving-origin flag).
constexpr int32_t static_kind(int32_t kind) { return kind & ~POT_MOVING; }
// Whether a kernel kind stages the Milky-Way table in LDS.
constexpr bool kind_has_table(int32_t kind) { return static_kind(kind) == POT_MW; }

// Moving-origin tracks: for origin i and time segment s, ORIGIN_COEFS doubles
// at ((i * n_segments) + s) * ORIGIN_COEFS hold the x, y and z quintics in
// the segment's local time tau in [0, 1), each as (tau^5, tau^4, ..., tau^0)
// coefficients.
constexpr int ORIGIN_COEFS = 18;

// Spherical-part table of the Milky-Way model (bulge + NFW halo):
//   a_sph(r) = -S(r^2) * r_vec,
// S(r2) = M_bulge(r)/r^3 + amp_h (ln(1 + r/a_h) - r/(a_h + r)) / r^3.
// S is tabulated on 2^TABLE_SUBBITS segments per octave of r2, found from the
// bits of r2 alone: the biased exponent and the top TABLE_SUBBITS mantissa
// bits form the segment key, and the remaining mantissa bits are the position
// t in [0, 1) inside the segment.  Each segment holds TABLE_NCOEF monomial
// coefficients of a Chebyshev interpolant in the local variable
// s = 2t - 1 in [-1, 1].  The LDS table covers r2 in [2^TABLE_E_MIN,
// 2^TABLE_E_MAX) (r from 2^-6 = 0.016 to 64 model units, 0.12 to 512 kpc at
// R0 = 8 kpc; 30 KB).  A second, "far" table of the same layout covers every
// positive finite double r2, [2^-1074, 2^1024) (subnormal r2 are scaled by
// 2^64 to find their key); it stays in global memory and is read only when an
// orbit leaves the LDS range.  The two are built segment by segment by the
// same routine, so on the octaves they share their coefficients, and the
// force, are the same bits: the LDS range only decides which lanes read
// global memory.  Half of the former 48 octaves kept the same results and
// left room for larger dense-output stashes and a third wave per SIMD.
constexpr int TABLE_SUBBITS = 4;
constexpr int TABLE_E_MIN = -12;
constexpr int TABLE_E_MAX = 12;
constexpr int TABLE_DEG = 9;
constexpr int TABLE_NCOEF = TABLE_DEG + 1;
constexpr int TABLE_SEGMENTS = (TABLE_E_MAX - TABLE_E_MIN) << TABLE_SUBBITS;
constexpr int TABLE_DOUBLES = TABLE_SEGMENTS * TABLE_NCOEF;
// The mixed-precision mode's float copy: the same segments, each with its own
// cubic interpolant evaluated in float (worst relative error 3.3e-7, about
// five float ulps, beside the force's other float rounding), stored in 4-float
// rows so a row is one 16-byte load and three FMAs.
constexpr int TABLE32_NCOEF = 4;
constexpr int TABLE32_STRIDE = 4;
constexpr int TABLE32_FLOATS = TABLE_SEGMENTS * TABLE32_STRIDE;
constexpr int FAR_E_MIN = -1074;
constexpr int FAR_E_MAX = 1024;
constexpr int FAR_SEGMENTS = (FAR_E_MAX - FAR_E_MIN) << TABLE_SUBBITS;
constexpr int FAR_DOUBLES = FAR_SEGMENTS * TABLE_NCOEF;
static_assert((TABLE_NCOEF * 8) % 16 == 0, "segments must stay 16-byte aligned for 128-bit loads");

// Cost ordering's buckets: 1/16 octave of an orbit's step rate, the highest
// rate in bucket 0.  A zero, negative or non-finite rate (no force, a singular
// or non-finite start: such orbits end at once or are free drifts) takes the
// last, cheapest bucket.
constexpr int kOrderBins = 1024;
constexpr int kOrderBinsPerOctave = 16;

DRIFT_HD DRIFT_FI uint32_t order_bin_of_rate(double rate) {
    if (!(rate > 0.0) || !(rate < INFINITY)) return kOrderBins - 1;
    const int level = (int)std::floor(kOrderBinsPerOctave * std::log2(rate));
    const int bin = kOrderBins / 2 - 1 - level;
    return (uint32_t)(bin < 0 ? 0 : (bin > kOrderBins - 1 ? kOrderBins - 1 : bin));
}

// Potential parameters in the form the kernels use (derived on the host
// from drift_potential).
struct PotParams {
    int32_t kind;
    int32_t has_table;
    double amp;       // Kepler / Plummer amplitude
    double b2;        // Plummer b^2
    double disk_amp, disk_a, disk_b2;
    const double* far_table;  // MW: FAR_DOUBLES coefficients in device memory
    // moving origins (kind | POT_MOVING), device memory
    const double* tracks;         // n_origins x n_segments x ORIGIN_COEFS
    const double* origin_params;  // n_origins x (amp, b^2)
    int32_t n_origins, n_segments;
    double t_start;  // start of segment 0
    double inv_dt;   // segments per unit time
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

// Reciprocal square root of a positive finite x: the v_rsq_f64 estimate
// (about 23 correct bits) refined by one second-order Newton-Raphson step,
// e = 1 - x y^2, 1/sqrt(x) = y (1 - e)^(-1/2) ~ y (1 + e/2 + 3 e^2 / 8), which
// leaves an error near one ulp.  Zero and infinite x give NaN (the forces only
// take it of sums with a positive term, or of a non-finite position that ends
// the orbit anyway).
DRIFT_HD DRIFT_FI double rsqrt_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
#else
    return 1.0 / std::sqrt(x);
#endif
}

// Reciprocal of a finite, nonzero, normal x: the v_rcp_f64 estimate refined
// by two Newton-Raphson steps y <- y + y (1 - x y) (error near one ulp).
DRIFT_HD DRIFT_FI double rcp_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    return fma(y, fma(-x, y, 1.0), y);
#else
    return 1.0 / x;
#endif
}

// Segment of r2 in a table whose key range starts at biased exponent
// key0_exp (= e_min + 1023) and holds nseg segments; s is the local
// coordinate in [-1, 1).  The unsigned difference puts keys below the range,
// zero, subnormal, negative, infinite and NaN r2 all out of range.
DRIFT_HD DRIFT_FI bool table_locate_bits(double r2, int32_t key0_exp, uint32_t nseg, int& seg,
                                         double& s) {
    const uint64_t b = __builtin_bit_cast(uint64_t, r2);
    const uint32_t hi = (uint32_t)(b >> 32), lo = (uint32_t)b;
    const uint32_t rel = (hi >> (20 - TABLE_SUBBITS)) - ((uint32_t)key0_exp << TABLE_SUBBITS);
    // the mantissa bits below the segment bits, as 1 + t in [1, 2)
    const uint32_t mhi = (((hi << TABLE_SUBBITS) | (lo >> (32 - TABLE_SUBBITS))) & 0x000FFFFFu) |
                         0x3FF00000u;
    const uint64_t mb = ((uint64_t)mhi << 32) | (uint64_t)(lo << TABLE_SUBBITS);
    s = fma(2.0, __builtin_bit_cast(double, mb), -3.0);  // 2 (1 + t) - 3 = 2t - 1, exact
    seg = (int)rel;
    return rel < nseg;
}

// The LDS table's range (false outside [2^E_MIN, 2^E_MAX) or not finite).
DRIFT_HD DRIFT_FI bool table_locate(double r2, int& seg, double& s) {
    return table_locate_bits(r2, TABLE_E_MIN + 1023, (uint32_t)TABLE_SEGMENTS, seg, s);
}

template <class T, int NCOEF = TABLE_NCOEF>
DRIFT_HD DRIFT_FI T horner_table(const T* c, T s) {
    T acc = c[NCOEF - 1];
#pragma unroll
    for (int k = NCOEF - 2; k >= 0; --k) acc = acc * s + c[k];
    return acc;
}

// S(r2) from the far table, for r2 outside the LDS table: any positive
// finite r2.  At r2 = 0 the spherical force S r_vec vanishes with r_vec, so
// S(0) is taken as 0; a non-finite r2 gives NaN.
DRIFT_HD DRIFT_FI double far_spherical(const PotParams& p, double r2) {
    if (!(r2 > 0.0)) return r2 == 0.0 ? 0.0 : r2;
    int32_t key0 = FAR_E_MIN + 1023;
    if (r2 < 0x1p-1022) {  // subnormal: scale into the normal range, shift the key
        r2 *= 0x1p64;
        key0 += 64;
    }
    int seg;
    double s;
    if (!table_locate_bits(r2, key0, (uint32_t)FAR_SEGMENTS, seg, s)) return NAN;  // r2 = inf
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
    if constexpr (sizeof(T) == 8) ir = rsqrt_pos(r2); else ir = rsqrt_f(r2);
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
    if constexpr (sizeof(T) == 8) ir = rsqrt_pos(r2); else ir = rsqrt_f(r2);
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
    if constexpr (sizeof(T) == 8) ir = rsqrt_pos(r2 + b2); else ir = rsqrt_f(r2 + b2);
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
    if constexpr (sizeof(T) == 8) izb = rsqrt_pos(zb2); else izb = rsqrt_f(zb2);
    T azb = a + zb2 * izb;
    T D2 = azb * azb + R2;
    if constexpr (sizeof(T) == 8) iD = rsqrt_pos(D2); else iD = rsqrt_f(D2);
    T f = -amp * (iD * iD * iD);
    ax = f * x;
    ay = f * y;
    az = f * z * azb * izb;
}

}  // namespace drift
