// spherical_table.cpp -- host-side construction of the Milky-Way spherical
// force table used by the gfx950 kernels.
//
// The bulge (cut-off power law) and NFW halo are both spherical, so their
// summed acceleration is -S(r^2) r_vec with
//   S(r2) = M_b(r) / r^3 + amp_h (ln(1 + r/a_h) - r/(a_h + r)) / r^3,
//   M_b(r) = 2 pi amp_b rc^(3-alpha) gamma(3/2 - alpha/2, r^2/rc^2).
// S has its only singularity at r2 = 0, so on every segment of r2 (2^K equal
// mantissa intervals per octave, K = TABLE_SUBBITS) it is approximated to
// ~1e-16 by a Chebyshev interpolant, evaluated in long double here and stored
// as monomial coefficients in the segment's local variable.
// The kernels then need no log, exp or pow for these two components.
//
// gamma(a, x) is evaluated by its power series for x < a + 1 and by the Lentz
// continued fraction of Gamma(a, x) otherwise (Abramowitz & Stegun 6.5.29 /
// 6.5.31; Press et al., Numerical Recipes 2nd ed. section 6.2), in long double.
#include <cmath>
#include <cstdint>
#include <vector>

#include "drift_device.hpp"
#include "drift_internal.hpp"

namespace drift {

namespace {

using ld = long double;

ld lower_gamma_ld(ld a, ld x) {
    if (x <= 0) return 0;
    if (x < a + 1) {
        ld ap = a, del = 1 / a, sum = del;
        for (int n = 0; n < 2000; ++n) {
            ap += 1;
            del *= x / ap;
            sum += del;
            if (fabsl(del) < fabsl(sum) * 1e-21L) break;
        }
        return sum * expl(-x + a * logl(x));
    }
    const ld fpmin = 1e-4000L;
    ld bb = x + 1 - a, c = 1 / fpmin, d = 1 / bb, h = d;
    for (int i = 1; i < 2000; ++i) {
        ld an = -i * (i - a);
        bb += 2;
        d = an * d + bb;
        if (fabsl(d) < fpmin) d = fpmin;
        c = bb + an / c;
        if (fabsl(c) < fpmin) c = fpmin;
        d = 1 / d;
        ld del = d * c;
        h *= del;
        if (fabsl(del - 1) < 1e-21L) break;
    }
    return tgammal(a) - expl(-x + a * logl(x)) * h;
}

// ln(1+u) - u/(1+u), with the alternating series for small u where the
// direct difference cancels.
ld nfw_mass_shape(ld u) {
    if (u < 0.125L) {
        // sum_{k>=2} (-1)^k (k-1)/k u^k
        ld pw = u, sum = 0;
        for (int k = 2; k < 80; ++k) {
            pw *= u;
            ld add = (k % 2 == 0 ? 1 : -1) * (ld)(k - 1) / (ld)k * pw;
            sum += add;
            if (fabsl(add) < 1e-24L * fabsl(sum)) break;
        }
        return sum;
    }
    return log1pl(u) - u / (1 + u);
}

struct MwSpherical {
    ld mass_scale, a_bulge, rc2, halo_amp, halo_a;
    ld operator()(ld r2) const {
        ld r = sqrtl(r2);
        ld r3 = r2 * r;
        ld sb = mass_scale * lower_gamma_ld(a_bulge, r2 / rc2) / r3;
        ld sh = halo_amp * nfw_mass_shape(r / halo_a) / r3;
        return sb + sh;
    }
};

MwSpherical make_spherical(const drift_potential& p) {
    MwSpherical f;
    f.a_bulge = 1.5L - 0.5L * (ld)p.bulge_alpha;
    f.mass_scale = 2 * 3.14159265358979323846264338327950288L * (ld)p.bulge_amp *
                   powl((ld)p.bulge_rc, 3 - (ld)p.bulge_alpha);
    f.rc2 = (ld)p.bulge_rc * (ld)p.bulge_rc;
    f.halo_amp = p.halo_amp;
    f.halo_a = p.halo_a;
    return f;
}

// Octave e and the interval [lo, hi) of the mantissa m = r2 / 2^e covered by
// segment `seg` of a table that starts at octave e_min.
void segment_bounds(int seg, int e_min, ld& lo, ld& hi, int& e) {
    const int sub = seg & ((1 << TABLE_SUBBITS) - 1);
    e = (seg >> TABLE_SUBBITS) + e_min;
    const ld w = ldexpl(1.0L, -TABLE_SUBBITS);
    lo = 1.0L + sub * w;
    hi = lo + w;
}

// Chebyshev interpolants of S with N coefficients on the segments of
// [2^e_min, 2^e_max), as monomial coefficients in the segment's local variable
// s in [-1, 1], written through store(segment, k, value).
template <int N, class Store>
void build_table_n(const drift_potential& p, int e_min, int e_max, Store store) {
    const MwSpherical f = make_spherical(p);
    const ld pi = 3.14159265358979323846264338327950288L;
    ld node[N], basis[N][N];
    for (int j = 0; j < N; ++j) node[j] = cosl(pi * (j + 0.5L) / N);
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j) basis[k][j] = cosl(k * pi * (j + 0.5L) / N);
    const int segments = (e_max - e_min) << TABLE_SUBBITS;
    for (int seg = 0; seg < segments; ++seg) {
        ld lo, hi;
        int e;
        segment_bounds(seg, e_min, lo, hi, e);
        const ld scale = ldexpl(1.0L, e);
        const ld centre = 0.5L * (lo + hi), half = 0.5L * (hi - lo);
        // Chebyshev interpolation at the N first-kind nodes.
        ld fv[N], cheb[N];
        for (int j = 0; j < N; ++j) fv[j] = f((centre + half * node[j]) * scale);
        for (int k = 0; k < N; ++k) {
            ld acc = 0;
            for (int j = 0; j < N; ++j) acc += fv[j] * basis[k][j];
            cheb[k] = 2 * acc / N;
        }
        cheb[0] *= 0.5L;
        // Convert sum_k cheb[k] T_k(s) to monomials in s.
        ld mono[N] = {};
        ld tkm1[N] = {}, tk[N] = {}, tkp1[N];
        tkm1[0] = 1;  // T_0
        tk[1] = 1;    // T_1
        mono[0] += cheb[0];
        if (N > 1) mono[1] += cheb[1];
        for (int k = 2; k < N; ++k) {
            for (int i = 0; i < N; ++i) tkp1[i] = -tkm1[i] + (i > 0 ? 2 * tk[i - 1] : 0);
            for (int i = 0; i < N; ++i) mono[i] += cheb[k] * tkp1[i];
            for (int i = 0; i < N; ++i) {
                tkm1[i] = tk[i];
                tk[i] = tkp1[i];
            }
        }
        for (int i = 0; i < N; ++i) store(seg, i, mono[i]);
    }
}

void build_table(const drift_potential& p, int e_min, int e_max, double* coeffs) {
    build_table_n<TABLE_NCOEF>(p, e_min, e_max, [coeffs](int seg, int k, ld v) {
        coeffs[seg * TABLE_NCOEF + k] = (double)v;
    });
}

}  // namespace

void build_mw_params(const drift_potential& p, PotParams& out) {
    out = PotParams{};
    out.kind = POT_MW;
    out.has_table = 1;
    out.disk_amp = p.disk_amp;
    out.disk_a = p.disk_a;
    out.disk_b2 = p.disk_b * p.disk_b;
    out.far_table = nullptr;  // set when the far table is on the device
}

void build_mw_table(const drift_potential& p, double* coeffs) {
    build_table(p, TABLE_E_MIN, TABLE_E_MAX, coeffs);
}

void build_mw_table_f32(const drift_potential& p, float* coeffs) {
    for (int k = 0; k < TABLE32_FLOATS; ++k) coeffs[k] = 0.0f;
    build_table_n<TABLE32_NCOEF>(p, TABLE_E_MIN, TABLE_E_MAX, [coeffs](int seg, int k, ld v) {
        coeffs[seg * TABLE32_STRIDE + k] = (float)v;
    });
}

void build_mw_far_table(const drift_potential& p, double* coeffs) {
    build_table(p, FAR_E_MIN, FAR_E_MAX, coeffs);
}

// Worst relative error of a table of [2^e_min, 2^e_max) against the long-double
// S at samples_per_seg points per segment, over the segments where S and its
// coefficients are normal doubles (|S| in [2^-960, 2^1000]).  Each point is
// located the way the kernels locate it (table_locate_bits, and for the far
// table the subnormal rescaling of far_spherical).
double table_max_rel_error(const drift_potential& p, const double* coeffs, int e_min, int e_max,
                           int samples_per_seg) {
    const MwSpherical f = make_spherical(p);
    double worst = 0.0;
    uint64_t state = 0x9E3779B97F4A7C15ull;
    const int segments = (e_max - e_min) << TABLE_SUBBITS;
    for (int seg = 0; seg < segments; ++seg) {
        ld lo, hi;
        int e;
        segment_bounds(seg, e_min, lo, hi, e);
        for (int k = 0; k < samples_per_seg; ++k) {
            state = state * 6364136223846793005ull + 1442695040888963407ull;
            double u = (double)(state >> 11) * (1.0 / 9007199254740992.0);
            double m = (double)(lo + (hi - lo) * (ld)u);
            if (m >= (double)hi) continue;
            double r2 = std::ldexp(m, e);
            ld exact = f((ld)r2);
            if (!(fabsl(exact) >= ldexpl(1.0L, -960) && fabsl(exact) <= ldexpl(1.0L, 1000))) continue;
            int32_t key0 = e_min + 1023;
            double r2k = r2;
            if (r2k < 0x1p-1022) {
                r2k *= 0x1p64;
                key0 += 64;
            }
            int sg;
            double s;
            // (subnormal r2 may round into a neighbouring segment: the point is
            // then checked where the kernels would look it up)
            if (!table_locate_bits(r2k, key0, (uint32_t)segments, sg, s)) return INFINITY;
            double approx = horner_table(coeffs + sg * TABLE_NCOEF, s);
            double rel = (double)fabsl(((ld)approx - exact) / exact);
            if (rel > worst) worst = rel;
        }
    }
    return worst;
}

double mw_table_max_rel_error(const drift_potential& p, const double* coeffs, int samples_per_seg) {
    return table_max_rel_error(p, coeffs, TABLE_E_MIN, TABLE_E_MAX, samples_per_seg);
}

double mw_far_table_max_rel_error(const drift_potential& p, const double* coeffs,
                                  int samples_per_seg) {
    return table_max_rel_error(p, coeffs, FAR_E_MIN, FAR_E_MAX, samples_per_seg);
}

double mw_spherical_exact(const drift_potential& p, double r2) {
    return (double)make_spherical(p)((long double)r2);
}

}  // namespace drift
