// integrators.hpp -- per-lane DOPR54 and DOP853 step machines.
//
// One orbit per GPU lane.  Each integrator is split into init() and attempt():
// attempt() performs exactly one attempted step (accepted or rejected) and
// writes every output sample that step completes, so the kernels can run a
// flat loop "while (!done) attempt()" in lock-step across a wavefront, and the
// persistent kernel can refill a lane whose orbit has finished.
//
// The arithmetic follows the reference CPU integrators:
//   DOP853: src/integrators/galpy/dop853.rs:316-783 (galpy leung_dop853.c)
//   DOPR54: src/integrators/galpy/dopr54.rs:135-778 (galpy bovy_rk.c)
// with these GPU-specific, value-preserving restatements:
//   * DOP853 runs in its second-order (Runge-Kutta-Nystrom) form: the
//     right-hand side is (v, F(x)), so the velocity half of each stage
//     derivative is the stage velocity, a known combination of earlier forces;
//     every combination of stage derivatives is formed from the forces alone
//     with exact products of the tableau (dop853_rkn.hpp), equal to the
//     reference's up to rounding, and the stage velocities are never formed;
//   * DOP853 evaluates its three dense-output stages only on accepted steps
//     that emit at least one sample (k10/k2/k3 are recomputed before reuse, so
//     the trajectory is unchanged);
//   * DOP853's err^(1/8) is sqrt(sqrt(sqrt(err))); pow(facold, beta) with
//     beta = 0 is 1 and is dropped; divisions by a value used more than once
//     (the error scale sk, 1/safe, the step h in the dense output) and the
//     step-size updates become multiplications by a reciprocal (Newton-refined,
//     within about an ulp of the quotient); the dense coefficients rcont5..8
//     keep their common factor h out, and each sample applies it once;
//   * DOPR54's log-space error norm s = ln(e^(atol-c) + e^(rtol+max-c)) + c is
//     the logarithm of sc = e^atol + e^rtol max|y|, so the norm is computed as
//     mean((yerr/sc)^2), and round(log2(0.85 err^-0.2)) is decided by comparing
//     that mean square against the six exact boundaries of the rounding.
// Termination guards the reference lacks: a non-finite DOP853 error estimate or
// max_attempts attempts end the orbit, flagging its status and writing NaN to
// the samples not reached.
#pragma once

#include <type_traits>

#include "drift_device.hpp"
#include "dop853_rkn.hpp"
#include "dopr54_rkn.hpp"

namespace drift {

constexpr int32_t STATUS_NONFINITE = 0x100;
constexpr int32_t STATUS_STEP_LIMIT = 0x200;

// ---------------------------------------------------------------------------
// DOP853 tableau (Hairer & Wanner; the coefficients galpy and the reference
// use, src/integrators/galpy/dop853.rs:105-264).
// ---------------------------------------------------------------------------
namespace hw853 {
constexpr double uround = 2.3e-16;
constexpr double c2 = 0.526001519587677318785587544488e-1, c3 = 0.789002279381515978178381316732e-1,
                 c4 = 0.118350341907227396726757197510, c5 = 0.281649658092772603273242802490,
                 c6 = 0.333333333333333333333333333333, c7 = 0.25,
                 c8 = 0.307692307692307692307692307692, c9 = 0.651282051282051282051282051282,
                 c10 = 0.6, c11 = 0.857142857142857142857142857142, c14 = 0.1, c15 = 0.2,
                 c16 = 0.777777777777777777777777777778;
constexpr double a21 = 5.26001519587677318785587544488e-2;
constexpr double a31 = 1.97250569845378994544595329183e-2, a32 = 5.91751709536136983633785987549e-2;
constexpr double a41 = 2.95875854768068491816892993775e-2, a43 = 8.87627564304205475450678981324e-2;
constexpr double a51 = 2.41365134159266685502369798665e-1, a53 = -8.84549479328286085344864962717e-1,
                 a54 = 9.24834003261792003115737966543e-1;
constexpr double a61 = 3.7037037037037037037037037037e-2, a64 = 1.70828608729473871279604482173e-1,
                 a65 = 1.25467687566822425016691814123e-1;
constexpr double a71 = 3.7109375e-2, a74 = 1.70252211019544039314978060272e-1,
                 a75 = 6.02165389804559606850219397283e-2, a76 = -1.7578125e-2;
constexpr double a81 = 3.70920001185047927108779319836e-2, a84 = 1.70383925712239993810214054705e-1,
                 a85 = 1.07262030446373284651809199168e-1, a86 = -1.53194377486244017527936158236e-2,
                 a87 = 8.27378916381402288758473766002e-3;
constexpr double a91 = 6.24110958716075717114429577812e-1, a94 = -3.36089262944694129406857109825e0,
                 a95 = -8.68219346841726006818189891453e-1, a96 = 2.75920996994467083049415600797e1,
                 a97 = 2.01540675504778934086186788979e1, a98 = -4.34898841810699588477366255144e1;
constexpr double a101 = 4.77662536438264365890433908527e-1, a104 = -2.48811461997166764192642586468e0,
                 a105 = -5.90290826836842996371446475743e-1, a106 = 2.12300514481811942347288949897e1,
                 a107 = 1.52792336328824235832596922938e1, a108 = -3.32882109689848629194453265587e1,
                 a109 = -2.03312017085086261358222928593e-2;
constexpr double a111 = -9.3714243008598732571704021658e-1, a114 = 5.18637242884406370830023853209e0,
                 a115 = 1.09143734899672957818500254654e0, a116 = -8.14978701074692612513997267357e0,
                 a117 = -1.85200656599969598641566180701e1, a118 = 2.27394870993505042818970056734e1,
                 a119 = 2.49360555267965238987089396762e0, a1110 = -3.0467644718982195003823669022e0;
constexpr double a121 = 2.27331014751653820792359768449e0, a124 = -1.05344954667372501984066689879e1,
                 a125 = -2.00087205822486249909675718444e0, a126 = -1.79589318631187989172765950534e1,
                 a127 = 2.79488845294199600508499808837e1, a128 = -2.85899827713502369474065508674e0,
                 a129 = -8.87285693353062954433549289258e0, a1210 = 1.23605671757943030647266201528e1,
                 a1211 = 6.43392746015763530355970484046e-1;
constexpr double a141 = 5.61675022830479523392909219681e-2, a147 = 2.53500210216624811088794765333e-1,
                 a148 = -2.46239037470802489917441475441e-1, a149 = -1.24191423263816360469010140626e-1,
                 a1410 = 1.5329179827876569731206322685e-1, a1411 = 8.20105229563468988491666602057e-3,
                 a1412 = 7.56789766054569976138603589584e-3, a1413 = -8.298e-3;
constexpr double a151 = 3.18346481635021405060768473261e-2, a156 = 2.83009096723667755288322961402e-2,
                 a157 = 5.35419883074385676223797384372e-2, a158 = -5.49237485713909884646569340306e-2,
                 a1511 = -1.08347328697249322858509316994e-4, a1512 = 3.82571090835658412954920192323e-4,
                 a1513 = -3.40465008687404560802977114492e-4, a1514 = 1.41312443674632500278074618366e-1;
constexpr double a161 = -4.28896301583791923408573538692e-1, a166 = -4.69762141536116384314449447206e0,
                 a167 = 7.68342119606259904184240953878e0, a168 = 4.06898981839711007970213554331e0,
                 a169 = 3.56727187455281109270669543021e-1, a1613 = -1.39902416515901462129418009734e-3,
                 a1614 = 2.9475147891527723389556272149e0, a1615 = -9.15095847217987001081870187138e0;
constexpr double b1 = 5.42937341165687622380535766363e-2, b6 = 4.45031289275240888144113950566,
                 b7 = 1.89151789931450038304281599044, b8 = -5.8012039600105847814672114227,
                 b9 = 3.1116436695781989440891606237e-1, b10 = -1.52160949662516078556178806805e-1,
                 b11 = 2.01365400804030348374776537501e-1, b12 = 4.47106157277725905176885569043e-2;
constexpr double bhh1 = 0.244094488188976377952755905512, bhh2 = 0.733846688281611857341361741547,
                 bhh3 = 0.220588235294117647058823529412e-1;
constexpr double er1 = 0.1312004499419488073250102996e-1, er6 = -0.1225156446376204440720569753e1,
                 er7 = -0.4957589496572501915214079952, er8 = 0.1664377182454986536961530415e1,
                 er9 = -0.3503288487499736816886487290, er10 = 0.3341791187130174790297318841,
                 er11 = 0.8192320648511571246570742613e-1, er12 = -0.2235530786388629525884427845e-1;
constexpr double d41 = -0.84289382761090128651353491142e1, d46 = 0.56671495351937776962531783590,
                 d47 = -0.30689499459498916912797304727e1, d48 = 0.23846676565120698287728149680e1,
                 d49 = 0.21170345824450282767155149946e1, d410 = -0.87139158377797299206789907490,
                 d411 = 0.22404374302607882758541771650e1, d412 = 0.63157877876946881815570249290,
                 d413 = -0.88990336451333310820698117400e-1, d414 = 0.18148505520854727256656404962e2,
                 d415 = -0.91946323924783554000451984436e1, d416 = -0.44360363875948939664310572000e1;
constexpr double d51 = 0.10427508642579134603413151009e2, d56 = 0.24228349177525818288430175319e3,
                 d57 = 0.16520045171727028198505394887e3, d58 = -0.37454675472269020279518312152e3,
                 d59 = -0.22113666853125306036270938578e2, d510 = 0.77334326684722638389603898808e1,
                 d511 = -0.30674084731089398182061213626e2, d512 = -0.93321305264302278729567221706e1,
                 d513 = 0.15697238121770843886131091075e2, d514 = -0.31139403219565177677282850411e2,
                 d515 = -0.93529243588444783865713862664e1, d516 = 0.35816841486394083752465898540e2;
constexpr double d61 = 0.19985053242002433820987653617e2, d66 = -0.38703730874935176555105901742e3,
                 d67 = -0.18917813819516756882830838328e3, d68 = 0.52780815920542364900561016686e3,
                 d69 = -0.11573902539959630126141871134e2, d610 = 0.68812326946963000169666922661e1,
                 d611 = -0.10006050966910838403183860980e1, d612 = 0.77771377980534432092869265740,
                 d613 = -0.27782057523535084065932004339e1, d614 = -0.60196695231264120758267380846e2,
                 d615 = 0.84320405506677161018159903784e2, d616 = 0.11992291136182789328035130030e2;
constexpr double d71 = -0.25693933462703749003312586129e2, d76 = -0.15418974869023643374053993627e3,
                 d77 = -0.23152937917604549567536039109e3, d78 = 0.35763911791061412378285349910e3,
                 d79 = 0.93405324183624310003907691704e2, d710 = -0.37458323136451633156875139351e2,
                 d711 = 0.10409964950896230045147246184e3, d712 = 0.29840293426660503123344363579e2,
                 d713 = -0.43533456590011143754432175058e2, d714 = 0.96324553959188282948394950600e2,
                 d715 = -0.39177261675615439165231486172e2, d716 = -0.14972683625798562581422125276e3;
// step-size controller (dop853.rs:342-347); beta = 0 so facold drops out
constexpr double safe = 0.9, facc1 = 1.0 / 0.333, facc2 = 1.0 / 6.0;
}  // namespace hw853

// The reference's ternary min/max (dop853.rs:286-299): NaN handling differs
// from fmin/fmax and is kept.
DRIFT_HD DRIFT_FI double tmin(double a, double b) { return a < b ? a : b; }
DRIFT_HD DRIFT_FI double tmax(double a, double b) { return a > b ? a : b; }
DRIFT_HD DRIFT_FI double sign_of(double x, double y) { return y > 0.0 ? std::fabs(x) : -std::fabs(x); }

DRIFT_HD DRIFT_FI bool finite6(const double* y) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < DIM; i++) ok = ok && std::isfinite(y[i]);
    return ok;
}

// The DOP853 coefficient table (dop853_rkn.hpp).  On the device it is read
// from constant memory through a pointer that each use site re-derives behind
// an empty asm statement that also takes a value the site depends on (`after`,
// e.g. the previous stage's force), so the compiler fetches a site's
// coefficients with 64-byte scalar loads next to their use instead of hoisting
// all ~200 of them out of the step (there are not enough scalar registers to
// keep them) or materialising every 64-bit literal with two scalar moves.
#if defined(__HIP_DEVICE_COMPILE__)
using Dop853Coef = const __attribute__((address_space(4))) double*;
__device__ __forceinline__ Dop853Coef dop853_coef_base() {
    return (Dop853Coef)rkn853::kCoefDevice;
}
__device__ __forceinline__ Dop853Coef dop853_coef(Dop853Coef base, double after) {
    asm volatile("" : "+s"(base) : "v"(after));
    return base;
}
// the same, after all six components of a group that came before
__device__ __forceinline__ Dop853Coef dop853_coef(Dop853Coef base, const double* after) {
    asm volatile("" : "+s"(base) : "v"(after[0]), "v"(after[1]), "v"(after[2]), "v"(after[3]),
                 "v"(after[4]), "v"(after[5]));
    return base;
}
#else
using Dop853Coef = const double*;
inline Dop853Coef dop853_coef_base() { return rkn853::kCoef; }
inline Dop853Coef dop853_coef(Dop853Coef base, double) { return base; }
inline Dop853Coef dop853_coef(Dop853Coef base, const double*) { return base; }
#endif

// Uniform (per-job) DOP853 data.
struct Dop853Job {
    Dop853Coef coef;  // the coefficient table (dop853_rkn.hpp)
    const double* times;
    int32_t nt;
    uint32_t max_attempts;
    double rtol, atol;  // linear tolerances (exp of the ABI's logarithms)
    double hmax, pos_neg;
};

struct Dop853Lane {
    double y[DIM];   // (x, v)
    double F1[3];    // force at y: the acceleration half of the first stage derivative
    // the reference's t_old (the last accepted step's start) only feeds its
    // dense output, which here takes the step's start directly, so the lane
    // does not keep it
    double h, t_cur;
    double t_next;  // times[ii + 1] while ii < nt - 1 (the next sample to reach)
    int32_t ii;  // index of the last emitted sample
    int32_t reject;
    uint32_t accepted, rejected, dense;
    int32_t status;
};

DRIFT_HD DRIFT_FI Dop853Job make_dop853_job(const double* times, int32_t nt, double rtol_lin,
                                             double atol_lin, uint32_t max_attempts, double t0,
                                             double tend) {
    Dop853Job j;
    j.coef = dop853_coef_base();
    j.times = times;
    j.nt = nt;
    j.max_attempts = max_attempts;
    j.rtol = rtol_lin;
    j.atol = atol_lin;
    j.hmax = tend - t0;                 // dop853.rs:348
    j.pos_neg = sign_of(1.0, j.hmax);   // dop853.rs:349
    return j;
}

// Initial step selection (Hairer), dop853.rs:372-423.  Emits sample 0.
template <class F, class Sink>
DRIFT_HD DRIFT_FI void dop853_init(Dop853Lane& L, const double* y0, double t0, const Dop853Job& J,
                                   const F& f, Sink& sink) {
#pragma unroll
    for (int i = 0; i < DIM; i++) L.y[i] = y0[i];
    sink.emit(0, L.y);
    double k1[DIM];
    f(t0, L.y, k1);
#pragma unroll
    for (int c = 0; c < 3; c++) L.F1[c] = k1[3 + c];
    double dnf = 0.0, dny = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        double sk = J.atol + J.rtol * std::fabs(L.y[i]);
        double sqr = k1[i] / sk;
        dnf += sqr * sqr;
        sqr = L.y[i] / sk;
        dny += sqr * sqr;
    }
    double h = sign_of(tmin(std::sqrt(dny / dnf) * 0.01, std::fabs(J.hmax)), J.pos_neg);
    double k3[DIM], k2[DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) k3[i] = L.y[i] + h * k1[i];
    f(t0 + h, k3, k2);
    double der2 = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        double sk = J.atol + J.rtol * std::fabs(L.y[i]);
        double sqr = (k2[i] - k1[i]) / sk;
        der2 += sqr * sqr;
    }
    der2 = std::sqrt(der2) / h;
    double der12 = tmax(std::fabs(der2), std::sqrt(dnf));
    double h1 = std::sqrt(std::sqrt(std::sqrt(0.01 / der12)));  // (0.01/der12)^(1/8)
    L.h = sign_of(tmin(100.0 * std::fabs(h), tmin(std::fabs(h1), std::fabs(J.hmax))), J.pos_neg);
    L.t_cur = t0;
    L.ii = 0;
    L.t_next = J.nt > 1 ? J.times[1] : 0.0;
    L.reject = 0;
    L.accepted = L.rejected = L.dense = 0;
    L.status = 0;
}

// Reloads the cached next sample time after the lane's sample index was
// restored from saved state.
DRIFT_HD DRIFT_FI void dop853_refresh_next(Dop853Lane& L, const Dop853Job& J) {
    L.t_next = L.ii < J.nt - 1 ? J.times[L.ii + 1] : 0.0;
}

// Writes NaN to the samples not reached and marks the lane finished.
template <class Sink>
DRIFT_HD DRIFT_FI void abandon_orbit(int32_t& ii, int32_t nt, int32_t& status, int32_t flag,
                                     Sink& sink) {
    double nan6[DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++) nan6[i] = NAN;
    while (ii < nt - 1) {
        ii++;
        sink.emit(ii, nan6);
    }
    status |= flag;
}

// Copies stage I's coefficients (c_I and its (A^2)_Im terms) out of the table.
template <int I>
DRIFT_HD DRIFT_FI void rkn853_load(Dop853Coef C, double* L) {
#pragma unroll
    for (int q = 0; q < rkn853::kStageN[I]; q++) L[q] = C[rkn853::kStageC[I] + q];
}

// Evaluates stage I of the Runge-Kutta-Nystrom form: the stage position
// X_I = x0 + h c_I v0 + h^2 sum_m (A^2)_Im F_m (the stage velocity, which the
// force does not read, is never formed) and its force F_I.  L holds stage I's
// coefficients; stage NEXT's are fetched into Lnext before the force call,
// so their scalar loads overlap the force evaluation.  The force is taken at
// t_step + c_I h (dop853.rs:432-578 and :689-727, t_step being the step's
// start).
template <int I, int NEXT, class F>
DRIFT_HD DRIFT_FI void rkn853_stage(const Dop853Job& J, double t_step, const double* y0, double h,
                                    double hh, double (&Fs)[17][3], const double* L,
                                    double* Lnext, const F& f) {
    const double hc = h * L[0];
    double q[DIM], a[DIM];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if constexpr (I == 2)
            q[c] = fma(hc, y0[3 + c], y0[c]);
        else
            q[c] = fma(hh, rkn853::stage_sum_local<I>(Fs, c, L), fma(hc, y0[3 + c], y0[c]));
        q[3 + c] = 0.0;
    }
    if constexpr (NEXT > 0) rkn853_load<NEXT>(dop853_coef(J.coef, q[2]), Lnext);
    f(t_step + hc, q, a);
#pragma unroll
    for (int c = 0; c < 3; c++) Fs[I][c] = a[3 + c];
}

// Dense output of one accepted step (dop853.rs:636-763): the three extra
// stages, rcont1..8 and every sample in (t_step, t_step + h].  Fs holds the
// step's stage forces F1 and F6..F13 (the ones the dense stages and rcont1..8
// read; F14..F16 are computed here), y0 the step's start and k5 its end.
// Lc0 holds stage 14's coefficients on entry.  ii is the last sample emitted
// before the step and t_s the time of sample ii + 1; on return ii is the last
// sample emitted and t_s the time of the sample after it.  Non-finite samples
// set STATUS_NONFINITE in status.  rcont5..8 are kept without their factor h
// (applied per sample); their position halves are h dense_k(F), the velocity
// halves sum_j d_kj F_j.
template <class F, class Sink>
DRIFT_HD DRIFT_FI void dop853_dense_emit(const Dop853Job& J, const F& f, Sink& sink,
                                         const double* y0, const double* k5, double (&Fs)[17][3],
                                         double h, double t_step, double (&Lc0)[12],
                                         double (&Lc1)[12], int32_t& ii, double& t_s,
                                         int32_t& status) {
    const double hh = h * h;
    const double t_new = t_step + h;
    double r1[DIM], r2[DIM], r3[DIM], r4[DIM], r5[DIM], r6[DIM], r7[DIM], r8[DIM];
    // the three extra stages (dop853.rs:703-742)
    rkn853_stage<14, 15>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<15, 16>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<16, 0>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    // rcont5..8 after the stages, one component at a time, so that each
    // component's forces die as its eight sums consume them; every group's
    // coefficients are fetched again per component, behind the previous
    // group's result (scalar loads, which would not fit the scalar registers
    // all at once)
#pragma unroll
    for (int c = 0; c < 3; c++) {
        {
            const Dop853Coef C = dop853_coef(J.coef, c == 0 ? Fs[16][2] : r8[2 + c]);
            r5[c] = rkn853::dense4pos_pre(Fs, c, C) + rkn853::dense4pos_post(Fs, c, C);
            r5[3 + c] = rkn853::dense4vel_pre(Fs, c, C) + rkn853::dense4vel_post(Fs, c, C);
        }
        {
            const Dop853Coef C = dop853_coef(J.coef, r5[3 + c]);
            r6[c] = rkn853::dense5pos_pre(Fs, c, C) + rkn853::dense5pos_post(Fs, c, C);
            r6[3 + c] = rkn853::dense5vel_pre(Fs, c, C) + rkn853::dense5vel_post(Fs, c, C);
        }
        {
            const Dop853Coef C = dop853_coef(J.coef, r6[3 + c]);
            r7[c] = rkn853::dense6pos_pre(Fs, c, C) + rkn853::dense6pos_post(Fs, c, C);
            r7[3 + c] = rkn853::dense6vel_pre(Fs, c, C) + rkn853::dense6vel_post(Fs, c, C);
        }
        {
            const Dop853Coef C = dop853_coef(J.coef, r7[3 + c]);
            r8[c] = rkn853::dense7pos_pre(Fs, c, C) + rkn853::dense7pos_post(Fs, c, C);
            r8[3 + c] = rkn853::dense7vel_pre(Fs, c, C) + rkn853::dense7vel_post(Fs, c, C);
        }
    }
    // rcont1..4 last: their inputs stay live anyway, the results only
    // from here on
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        // k1 = (v0, F1), k4 (the new derivative) = (v_new, F13)
        const double k1i = i < 3 ? y0[3 + i] : Fs[1][i - 3];
        const double k4i = i < 3 ? k5[3 + i] : Fs[13][i - 3];
        r1[i] = y0[i];
        double ydiff = k5[i] - y0[i];
        r2[i] = ydiff;
        double bspl = h * k1i - ydiff;
        r3[i] = bspl;
        r4[i] = ydiff - h * k4i - bspl;
    }
    const double* times = J.times;
    const int32_t last = J.nt - 1;
    const double inv_h = rcp_fast(h);
    do {
        // fetch the following sample time now, so its latency hides
        // behind this sample's polynomial and stores
        const int32_t after = ii + 2;
        const double t_after = after <= last ? times[after] : 0.0;
        const double sv = (t_s - t_step) * inv_h;
        const double s1 = 1.0 - sv;
        const double s1h = s1 * h;
        const double s1hh = s1h * h;  // positions carry one more h
        double yy[DIM];
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            const double w = i < 3 ? s1hh : s1h;
            yy[i] = r1[i] +
                    sv * (r2[i] +
                          s1 * (r3[i] +
                                sv * (r4[i] + w * (r5[i] + sv * (r6[i] + s1 * (r7[i] + sv * r8[i]))))));
        }
        ii++;
        if (!finite6(yy)) status |= STATUS_NONFINITE;
        sink.emit(ii, yy);
        t_s = t_after;
    } while (ii < last && J.pos_neg * t_s < J.pos_neg * t_new);
}

// The trial part of one attempted DOP853 step (dop853.rs:425-619): stages 2
// to 12 from the lane's state, the 8th-order solution k5 and the error
// estimate, returned.  Fs[1] holds F1 on entry and Fs[2..12] the stage
// forces on return.
//
// The right-hand side is (v, F(x)), so the velocity half of each stage
// derivative is that stage's velocity, itself v0 plus h times a combination
// of the earlier forces.  Every combination of stage derivatives the step
// forms is therefore computed from the forces alone, with the combined
// coefficients of dop853_rkn.hpp (exact products of the tableau, rounded
// once): the same method and the same values up to rounding, with half the
// stage arithmetic and half the stage storage of the first-order form.
template <class F>
DRIFT_HD DRIFT_FI double dop853_trial(const Dop853Lane& L, const Dop853Job& J, const F& f,
                                      double h, double (&Fs)[17][3], double (&k5)[DIM]) {
    const double hh = h * h;
    const double* y0 = L.y;
    const double t_step = L.t_cur;  // start of this step

    // stage coefficients, double-buffered: each stage fetches the next one's
    double Lc0[12], Lc1[12];
    rkn853_load<2>(dop853_coef(J.coef, h), Lc0);
    rkn853_stage<2, 3>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<3, 4>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<4, 5>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<5, 6>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<6, 7>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<7, 8>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<8, 9>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<9, 10>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<10, 11>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);
    rkn853_stage<11, 12>(J, t_step, y0, h, hh, Fs, Lc1, Lc0, f);
    rkn853_stage<12, 0>(J, t_step, y0, h, hh, Fs, Lc0, Lc1, f);

    // k4 = sum_j b_j k_j and the 8th-order solution k5 = y0 + h k4
    // (dop853.rs:585-595); the position half of k4 is v0 + h bpos(F)
    double k4[DIM];
    {
        const Dop853Coef C = dop853_coef(J.coef, Fs[12][2]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            k4[3 + c] = rkn853::bvel(Fs, c, C);
            k4[c] = y0[3 + c] + h * rkn853::bpos(Fs, c, C);
        }
    }
#pragma unroll
    for (int i = 0; i < DIM; i++) k5[i] = y0[i] + h * k4[i];

    // error estimates (dop853.rs:597-619)
    double err = 0.0, err2 = 0.0;
    double e1[DIM], e2[DIM];
    {
        const Dop853Coef C = dop853_coef(J.coef, k5[5]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            e2[c] = h * rkn853::bhhpos(Fs, c, C);
            e1[c] = h * rkn853::errpos(Fs, c, C);
        }
    }
    {
        const Dop853Coef C = dop853_coef(J.coef, e1[2]);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            e2[3 + c] = k4[3 + c] - rkn853::bhhvel(Fs, c, C);
            e1[3 + c] = rkn853::errvel(Fs, c, C);
        }
    }
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        const double inv_sk = rcp_fast(J.atol + J.rtol * tmax(std::fabs(y0[i]), std::fabs(k5[i])));
        double sqr = e2[i] * inv_sk;
        err2 += sqr * sqr;
        sqr = e1[i] * inv_sk;
        err += sqr * sqr;
    }
    double deno = err + 0.01 * err2;
    if (deno <= 0.0) deno = 1.0;
    // sqrt(1/(deno*6)) as a reciprocal square root
    return std::fabs(h) * err * rsqrt_d(deno * (double)DIM);
}

// Step factors (dop853.rs:621-623, :775) as multipliers of h:
//   accept: h / max(facc2, min(facc1, fac11 / safe)) = h min(1/facc2, shrink)
//   reject: h / min(facc1, fac11 / safe)            = h shrink
// with shrink = max(1/facc1, safe err^(-1/8)) and fac11 = err^(1/8).
// err^(-1/8) is three reciprocal square roots; err = 0 (or below the normal
// range) gives an infinite factor, err = inf a zero one, as the reference's
// limits of pow(err, 1/8).  Returns shrink; grow_h is the accepted step's
// candidate h min(1/facc2, shrink).
DRIFT_HD DRIFT_FI double dop853_shrink(double err, double h, double& grow_h) {
    using namespace hw853;
    const double root8 = rsqrt_pos(rsqrt_pos(rsqrt_pos(err)));
    const double grow = err > 1e-300 ? (err < INFINITY ? safe * root8 : 0.0) : INFINITY;
    constexpr double inv_facc1 = 1.0 / facc1, inv_facc2 = 1.0 / facc2;
    const double shrink = tmax(inv_facc1, grow);
    grow_h = h * tmin(inv_facc2, shrink);
    return shrink;
}

// Accepted step's bookkeeping (dop853.rs:764-773): the new state, FSAL force,
// the next step size (clamped to hmax; no growth right after a rejection)
// and the times.
DRIFT_HD DRIFT_FI void dop853_commit_accept(Dop853Lane& L, const Dop853Job& J, const double* k5,
                                            const double* F13, double h, double hnew,
                                            double t_new) {
#pragma unroll
    for (int i = 0; i < DIM; i++) L.y[i] = k5[i];
#pragma unroll
    for (int c = 0; c < 3; c++) L.F1[c] = F13[c];
    hnew = std::fabs(hnew) > std::fabs(J.hmax) ? J.pos_neg * J.hmax : hnew;
    if (L.reject != 0) hnew = J.pos_neg * tmin(std::fabs(hnew), std::fabs(h));
    L.reject = 0;
    L.t_cur = t_new;
    L.h = hnew;
}

// One attempted DOP853 step, dop853.rs:425-782, writing the samples of an
// accepted step at once.  Returns true when the orbit has emitted its last
// sample.
template <class F, class Sink>
DRIFT_HD DRIFT_FI bool dop853_attempt(Dop853Lane& L, const Dop853Job& J, const F& f, Sink& sink) {
    using namespace hw853;
    const double h = J.pos_neg * tmax(std::fabs(L.h), 1e3 * uround);
    const double* y0 = L.y;
    const double t_step = L.t_cur;  // start of this step
    const double t_new = L.t_cur + h;
    double Fs[17][3];
#pragma unroll
    for (int c = 0; c < 3; c++) Fs[1][c] = L.F1[c];
    double k5[DIM];
    const double err = dop853_trial(L, J, f, h, Fs, k5);
    if (err != err) {
        // The reference would shrink h to its floor and retry forever.
        abandon_orbit(L.ii, J.nt, L.status, STATUS_NONFINITE, sink);
        return true;
    }
    double hnew;
    const double shrink = dop853_shrink(err, h, hnew);

    if (err <= 1.0) {
        L.accepted++;
        // FSAL derivative at the new point (stage 13): (v_new, F(x_new)); the
        // first dense stage's coefficients are fetched meanwhile
        double Lc0[12], Lc1[12];
        rkn853_load<14>(dop853_coef(J.coef, k5[2]), Lc0);
        {
            double a[DIM];
            f(t_new, k5, a);
#pragma unroll
            for (int c = 0; c < 3; c++) Fs[13][c] = a[3 + c];
        }
        if (L.ii < J.nt - 1 && J.pos_neg * L.t_next < J.pos_neg * t_new) {
            // Dense output (dop853.rs:636-763), only when a sample falls in
            // (t_step, t_new).
            L.dense++;
            double t_s = L.t_next;
            dop853_dense_emit(J, f, sink, y0, k5, Fs, h, t_step, Lc0, Lc1, L.ii, t_s, L.status);
            L.t_next = t_s;
        }
        dop853_commit_accept(L, J, k5, Fs[13], h, hnew, t_new);
    } else {
        L.rejected++;
        // t_current goes back to the step's start, which it still holds
        // (dop853.rs:777-778)
        L.reject = 1;
        L.h = h * shrink;
    }
    if (L.ii >= J.nt - 1) return true;
    if (L.accepted + L.rejected >= J.max_attempts) {
        abandon_orbit(L.ii, J.nt, L.status, STATUS_STEP_LIMIT, sink);
        return true;
    }
    return false;
}

// The inputs of one accepted step's dense output, as a lane hands them to
// another lane of its wavefront: the step's start y0 and end k5, the stage
// forces F1 and F6..F13, h and the step's start time, the orbit's output
// index and the last sample emitted before the step.
struct Dop853Dense {
    double y0[DIM], k5[DIM];
    double F[9][3];  // F1, F6, F7, ..., F13
    double h, t_step;
    int64_t idx;
    int32_t ii;
};

// stage index of Dop853Dense::F[q]
DRIFT_HD DRIFT_FI constexpr int dense_force_stage(int q) { return q == 0 ? 1 : q + 5; }

// A Dop853Dense as a record of 64-bit words, the form a wavefront's stash
// holds: y0, k5, F (row-major), h, t_step, then idx and ii as integers; one
// word of padding makes the count even (the stash moves words in pairs).
constexpr int kDenseUsedWords = 2 * DIM + 27 + 4;
constexpr int kDenseRecordWords = kDenseUsedWords + (kDenseUsedWords & 1);

struct DenseWords {
    double w[kDenseRecordWords];
};

DRIFT_HD DRIFT_FI void dense_pack(const Dop853Dense& r, DenseWords& x) {
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        x.w[i] = r.y0[i];
        x.w[DIM + i] = r.k5[i];
    }
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) x.w[2 * DIM + 3 * q + c] = r.F[q][c];
    x.w[2 * DIM + 27] = r.h;
    x.w[2 * DIM + 28] = r.t_step;
    x.w[2 * DIM + 29] = __builtin_bit_cast(double, (int64_t)r.idx);
    x.w[2 * DIM + 30] = __builtin_bit_cast(double, (int64_t)r.ii);
    static_assert(2 * DIM + 30 < kDenseRecordWords, "the record holds every word");
#pragma unroll
    for (int k = kDenseUsedWords; k < kDenseRecordWords; k++) x.w[k] = 0.0;
}

DRIFT_HD DRIFT_FI void dense_unpack(const DenseWords& x, Dop853Dense& r) {
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        r.y0[i] = x.w[i];
        r.k5[i] = x.w[DIM + i];
    }
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) r.F[q][c] = x.w[2 * DIM + 3 * q + c];
    r.h = x.w[2 * DIM + 27];
    r.t_step = x.w[2 * DIM + 28];
    r.idx = __builtin_bit_cast(int64_t, x.w[2 * DIM + 29]);
    r.ii = (int32_t)__builtin_bit_cast(int64_t, x.w[2 * DIM + 30]);
}

// The same attempted step with its dense output handed over instead of
// written: an accepted step with samples fills a Dop853Dense (idx is left to
// the caller), the lane's sample index and next sample time move
// past the samples the step covers, and round(own, req, completes) runs
// before the step is committed -- at a point every lane of the caller
// reaches, with own set where this lane's step left req -- so that whoever
// runs the request (dop853_run_dense) writes the samples.  completes is set
// where this attempt ends the orbit (its last sample, a non-finite error
// estimate or the attempt limit).  Counts, samples and the returned
// completion are those of dop853_attempt.
template <class F, class Sink, class Round>
DRIFT_HD DRIFT_FI bool dop853_attempt_deferred(Dop853Lane& L, const Dop853Job& J, const F& f,
                                               Sink& sink, Round&& round) {
    using namespace hw853;
    const double h = J.pos_neg * tmax(std::fabs(L.h), 1e3 * uround);
    const double t_step = L.t_cur;
    const double t_new = L.t_cur + h;
    const int32_t last = J.nt - 1;
    double Fs[17][3];
#pragma unroll
    for (int c = 0; c < 3; c++) Fs[1][c] = L.F1[c];
    double k5[DIM];
    const double err = dop853_trial(L, J, f, h, Fs, k5);
    const bool nonfinite = err != err;
    double hnew;
    const double shrink = dop853_shrink(err, h, hnew);
    const bool accept = err <= 1.0;  // false for a NaN
    Dop853Dense req;
    bool own = false;
    if (accept) {
        L.accepted++;
        const double* times = J.times;
        own = L.ii < last && J.pos_neg * L.t_next < J.pos_neg * t_new;
        {
            double a[DIM];
            f(t_new, k5, a);
#pragma unroll
            for (int c = 0; c < 3; c++) Fs[13][c] = a[3 + c];
        }
        if (own) {
            L.dense++;
#pragma unroll
            for (int i = 0; i < DIM; i++) {
                req.y0[i] = L.y[i];
                req.k5[i] = k5[i];
            }
#pragma unroll
            for (int q = 0; q < 9; q++)
#pragma unroll
                for (int c = 0; c < 3; c++) req.F[q][c] = Fs[dense_force_stage(q)][c];
            req.h = h;
            req.t_step = t_step;
            req.ii = L.ii;
            // the samples the step covers, as dop853_dense_emit's loop walks
            // them: t_s is the time of sample ii + 1 after each
            int32_t ii = L.ii;
            double t_s;
            do {
                const int32_t after = ii + 2;
                t_s = after <= last ? times[after] : 0.0;
                ii++;
            } while (ii < last && J.pos_neg * t_s < J.pos_neg * t_new);
            L.ii = ii;
            L.t_next = t_s;
        }
    } else if (!nonfinite) {
        L.rejected++;
    }
    const bool completes = nonfinite || L.ii >= last || L.accepted + L.rejected >= J.max_attempts;
    round(own, req, completes);
    if (nonfinite) {
        // The reference would shrink h to its floor and retry forever.
        abandon_orbit(L.ii, J.nt, L.status, STATUS_NONFINITE, sink);
        return true;
    }
    if (accept) {
        dop853_commit_accept(L, J, k5, Fs[13], h, hnew, t_new);
    } else {
        // t_current and t_old keep their values (dop853.rs:777-778)
        L.reject = 1;
        L.h = h * shrink;
    }
    if (L.ii >= last) return true;
    if (L.accepted + L.rejected >= J.max_attempts) {
        abandon_orbit(L.ii, J.nt, L.status, STATUS_STEP_LIMIT, sink);
        return true;
    }
    return false;
}

// Runs a handed-over dense output on the calling lane: the request's stage
// forces go back into their stage slots, and the samples go to the request's
// orbit through sink (whose orbit index the caller has set to req.idx).
// Returns the STATUS_NONFINITE bit if a sample was not finite.
template <class F, class Sink>
DRIFT_HD DRIFT_FI int32_t dop853_run_dense(const Dop853Dense& req, const Dop853Job& J, const F& f,
                                           Sink& sink) {
    double Fs[17][3];
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
        for (int c = 0; c < 3; c++) Fs[dense_force_stage(q)][c] = req.F[q][c];
    double Lc0[12], Lc1[12];
    rkn853_load<14>(dop853_coef(J.coef, req.h), Lc0);
    int32_t ii = req.ii;
    double t_s = J.times[ii + 1];
    int32_t status = 0;
    dop853_dense_emit(J, f, sink, req.y0, req.k5, Fs, req.h, req.t_step, Lc0, Lc1, ii, t_s, status);
    return status;
}

// ---------------------------------------------------------------------------
// DOPR54 (galpy bovy_rk.c), dopr54.rs.
// ---------------------------------------------------------------------------
namespace dp54 {
constexpr double c2 = 0.2, c3 = 0.3, c4 = 0.8, c5 = 8.0 / 9.0;
constexpr double a21 = 0.2;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double a71 = 35.0 / 384.0, a73 = 500.0 / 1113.0, a74 = 125.0 / 192.0,
                 a75 = -2187.0 / 6784.0, a76 = 11.0 / 84.0;
constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0, b5 = -2187.0 / 6784.0,
                 b6 = 11.0 / 84.0;
constexpr double e1 = b1 - 5179.0 / 57600.0, e3 = b3 - 7571.0 / 16695.0, e4 = b4 - 393.0 / 640.0,
                 e5 = b5 + 92097.0 / 339200.0, e6 = b6 - 187.0 / 2100.0, e7 = -1.0 / 40.0;
constexpr double max_stepreduce = 10000.0;  // dopr54.rs:34
constexpr double max_dt_reduce = 10000.0;   // dopr54.rs:35
constexpr double initial_step_sentinel = -9999.99;
}  // namespace dp54

struct Dopr54Job {
    const double* times;
    int32_t nt;
    uint32_t max_attempts;
    double dt;             // t[1] - t[0], the only grid spacing DOPR54 reads (dopr54.rs:713)
    double rtol, atol;     // linear tolerances e^rtol, e^atol
    double dt_one_init;    // sentinel selects the RK4 estimate
    double bound2[6];      // squared error-norm boundaries of round(log2(0.85 err^-0.2)) >= j, j = -2..3
};

struct Dopr54Lane {
    double y[DIM], a1[DIM];
    double to, dt_one, init_to, init_dt_one;
    int32_t ii;
    int32_t err;  // reference error code (dopr54.rs:407-418)
    uint32_t accepted, rejected;
    int32_t status;
};

// Boundaries err <= U_j  <=>  round(log2(0.85) - 0.2 log2(err)) >= j, with
// U_j = 0.85^5 2^(5 (1/2 - j)); stored squared for comparison with mean(e^2).
inline void dopr54_bounds(double* bound2) {
    for (int j = -2; j <= 3; ++j) {
        double u = std::pow(0.85, 5.0) * std::ldexp(std::sqrt(2.0), 2 - 5 * j);  // 2^(2.5-5j)
        bound2[j + 2] = u * u;
    }
}

DRIFT_HD DRIFT_FI double max_abs6(const double* y) {
    double m = std::fabs(y[0]);
#pragma unroll
    for (int i = 1; i < DIM; i++) m = tmax(std::fabs(y[i]), m);
    return m;
}

// rk4_estimate_step, dopr54.rs:210-368, with the scale e^s = atol + rtol max|y|.
// Each step-doubling test takes three rk4_onestep calls (dopr54.rs:135-208):
// dt from yo, dt/2 from yo, dt/2 from that half step.  Their 12 stages run as
// one loop, so the force is inlined once (for point masses the estimate is
// called out of line, rk4_estimate_call, so that its registers stay out of
// the step loop's budget).  The arithmetic and its
// order are rk4_onestep's: stage k of a step of length h from `base` at tn is
// taken at tn, tn + h/2, tn + h/2, tn + h, adds h a / (6, 3, 3, 6) to the
// step's accumulator, and the next stage point is base + (h a) (1/2, 1/2, 1).
template <class F>
DRIFT_HD DRIFT_FI double rk4_estimate_step(const double* yo, double dt, double to,
                                           const Dopr54Job& J, const F& f) {
    const double init_dt = dt;
    const double sc = J.atol + J.rtol * max_abs6(yo);
    for (int iter = 0; iter < 64; ++iter) {
        double y1[DIM], y21[DIM], base[DIM], acc[DIM], ynk[DIM], a[DIM];
#pragma unroll 1
        for (int s = 0; s < 12; ++s) {
            const int o = s >> 2, k = s & 3;  // step o (full, half, second half), stage k
            if (k == 0) {
#pragma unroll
                for (int i = 0; i < DIM; i++) {
                    base[i] = o == 2 ? y21[i] : yo[i];
                    acc[i] = base[i];
                    ynk[i] = base[i];
                }
            }
            const double h = o == 0 ? dt : dt / 2.0;
            const double tn = o == 2 ? to + dt / 2.0 : to;
            f(k == 0 ? tn : (k == 3 ? tn + h : tn + h / 2.0), ynk, a);
            const double w = (k == 0 || k == 3) ? 6.0 : 3.0;
            const double m = k == 2 ? 1.0 : 0.5;
#pragma unroll
            for (int i = 0; i < DIM; i++) {
                acc[i] += h * a[i] / w;
                ynk[i] = base[i] + h * a[i] * m;
            }
            if (k == 3 && o == 0) {
#pragma unroll
                for (int i = 0; i < DIM; i++) y1[i] = acc[i];
            }
            if (k == 3 && o == 1) {
#pragma unroll
                for (int i = 0; i < DIM; i++) y21[i] = acc[i];
            }
        }
        double e = 0.0;  // acc holds y2, the two half steps
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            double q = (y1[i] - acc[i]) / sc;
            e += q * q;
        }
        double err = std::sqrt(e / (double)DIM);
        if (!(err > 1.0)) break;
        double factor = std::ceil(std::pow(err, 0.2));
        if (factor > 1.0 && init_dt / dt * factor < dp54::max_dt_reduce) {
            dt /= factor;
        } else {
            break;
        }
    }
    return dt;
}

// rk4_estimate_step as a call of its own, the state passed by value: the
// estimate runs once per orbit and needs some 45 more registers than the
// step loop, so inlined it would set the kernel's register count and with
// it the occupancy (the kernels' waves-per-SIMD target then bounds the call's
// registers as well, and it spills inside itself if it must).  Only the
// job's tolerances are read.  Used for the forces that declare
// kFewRegisters (point masses), whose launches gain a wave per SIMD by it;
// the others keep the estimate inline (a table or an origin track cache
// passed by value would cost more than the call saves).
template <class F, class = void>
struct few_registers : std::false_type {};
template <class F>
struct few_registers<F, std::void_t<decltype(F::kFewRegisters)>>
    : std::bool_constant<F::kFewRegisters> {};

template <class F>
DRIFT_HD __attribute__((noinline)) double rk4_estimate_call(double y0, double y1, double y2,
                                                            double y3, double y4, double y5,
                                                            double dt, double to, double atol,
                                                            double rtol, F f) {
    const double yo[DIM] = {y0, y1, y2, y3, y4, y5};
    Dopr54Job J{};
    J.atol = atol;
    J.rtol = rtol;
    return rk4_estimate_step(yo, dt, to, J, f);
}

// Writes samples for every interval the lane has completed (the save after
// each dopr54_onestep, dopr54.rs:728-764) and opens the next interval.
template <class Sink>
DRIFT_HD DRIFT_FI void dopr54_emit_completed(Dopr54Lane& L, const Dopr54Job& J, Sink& sink) {
    const double dt = J.dt;
    while (L.ii < J.nt - 1) {
        const double end = L.init_to + dt;
        const bool open = (dt >= 0.0 && L.to < end) || (dt < 0.0 && L.to > end);
        if (open) break;
        L.ii++;
        if (!finite6(L.y)) L.status |= STATUS_NONFINITE;
        sink.emit(L.ii, L.y);
        L.init_to = L.to;
        L.init_dt_one = L.dt_one;
    }
}

// The start of an orbit once its initial sub-step dt_one is decided (the
// job's explicit value or the RK4 estimate): sample 0, the first force, the
// bookkeeping, and the samples an already completed interval owes.
template <class F, class Sink>
DRIFT_HD DRIFT_FI void dopr54_begin(Dopr54Lane& L, const double* y0, double t0, double dt_one,
                                    const Dopr54Job& J, const F& f, Sink& sink) {
#pragma unroll
    for (int i = 0; i < DIM; i++) L.y[i] = y0[i];
    sink.emit(0, L.y);
    L.to = t0;
    f(t0, L.y, L.a1);
    L.dt_one = dt_one;
    L.init_to = t0;
    L.init_dt_one = dt_one;
    L.ii = 0;
    L.err = 0;
    L.accepted = L.rejected = 0;
    L.status = 0;
    dopr54_emit_completed(L, J, sink);
}

// dopr54_begin after the estimate the sentinel asks for (dopr54.rs:713-726).
template <class F, class Sink>
DRIFT_HD DRIFT_FI void dopr54_init(Dopr54Lane& L, const double* y0, double t0, const Dopr54Job& J,
                                   const F& f, Sink& sink) {
    double dt_one = J.dt_one_init;
    if (dt_one == dp54::initial_step_sentinel) {
        if constexpr (few_registers<F>::value)
            dt_one = rk4_estimate_call(y0[0], y0[1], y0[2], y0[3], y0[4], y0[5], J.dt, t0, J.atol,
                                       J.rtol, f);
        else
            dt_one = rk4_estimate_step(y0, J.dt, t0, J, f);
    }
    dopr54_begin(L, y0, t0, dt_one, J, f, sink);
}

// One pass of the dopr54_onestep loop body (dopr54.rs:403-459) with
// dopr54_actualstep (:462-659).  Returns true when the orbit is complete.
//
// As for DOP853, the right-hand side is (v, F(x)): the stage derivatives'
// velocity halves are the stage velocities, themselves v0 plus dt times
// combinations of earlier forces, so the stage positions, the update and the
// error estimate are formed from the forces F_1..F_7 alone with the combined
// coefficients of dopr54_rkn.hpp (the same values up to rounding; stage 7's
// point is the update itself, as a_7j = b_j).
template <class F, class Sink>
DRIFT_HD DRIFT_FI bool dopr54_attempt(Dopr54Lane& L, const Dopr54Job& J, const F& f, Sink& sink) {
    using namespace dp54;
    using namespace rkn54;
    const double dt = J.dt;
    bool accept = false;
    double dt_one = L.dt_one;
    if (L.init_dt_one / dt_one > max_stepreduce || dt_one != dt_one) {
        dt_one = L.init_dt_one / max_stepreduce;
        accept = true;
        if (L.err % 2 == 0) L.err += 1;
    }
    if (dt >= 0.0 && dt_one > (L.init_to + dt - L.to)) dt_one = L.init_to + dt - L.to;
    if (dt < 0.0 && dt_one < (L.init_to + dt - L.to)) dt_one = L.init_to + dt - L.to;

    const double* yo = L.y;
    const double to = L.to;  // stage times to + c_i dt_one (dopr54.rs:545-607)
    const double h = dt_one, hh = dt_one * dt_one;
    double F1[3], F2[3], F3[3], F4[3], F5[3], F6[3], F7[3];
    double q[DIM], a[DIM];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F1[c] = L.a1[3 + c];
        q[3 + c] = 0.0;  // the forces read positions only
    }
#pragma unroll
    for (int c = 0; c < 3; c++) q[c] = fma(h * c2, yo[3 + c], yo[c]);
    f(to + c2 * dt_one, q, a);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F2[c] = a[3 + c];
        q[c] = fma(hh, q31 * F1[c], fma(h * c3, yo[3 + c], yo[c]));
    }
    f(to + c3 * dt_one, q, a);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F3[c] = a[3 + c];
        q[c] = fma(hh, q41 * F1[c] + q42 * F2[c], fma(h * c4, yo[3 + c], yo[c]));
    }
    f(to + c4 * dt_one, q, a);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F4[c] = a[3 + c];
        q[c] = fma(hh, q51 * F1[c] + q52 * F2[c] + q53 * F3[c], fma(h * c5, yo[3 + c], yo[c]));
    }
    f(to + c5 * dt_one, q, a);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F5[c] = a[3 + c];
        q[c] = fma(hh, q61 * F1[c] + q62 * F2[c] + q63 * F3[c] + q64 * F4[c],
                   fma(h, yo[3 + c], yo[c]));
    }
    f(to + dt_one, q, a);
    // 5th-order update (x1, v1); stage 7 is evaluated at it (FSAL)
    double yn1[DIM], yerr[DIM];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F6[c] = a[3 + c];
        yn1[c] = fma(hh, bq1 * F1[c] + bq3 * F3[c] + bq4 * F4[c] + bq5 * F5[c],
                     fma(h, yo[3 + c], yo[c]));
        yn1[3 + c] = fma(h, b1 * F1[c] + b3 * F3[c] + b4 * F4[c] + b5 * F5[c] + b6 * F6[c],
                         yo[3 + c]);
    }
    f(to + dt_one, yn1, a);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        F7[c] = a[3 + c];
        yerr[c] = hh * (eq1 * F1[c] + eq3 * F3[c] + eq4 * F4[c] + eq5 * F5[c] + eq6 * F6[c]);
        yerr[3 + c] = h * (e1 * F1[c] + e3 * F3[c] + e4 * F4[c] + e5 * F5[c] + e6 * F6[c] +
                           e7 * F7[c]);
    }

    // error norm: mean((yerr / (atol + rtol max|yo|))^2)
    const double maxabs = max_abs6(yo);
    const double inv_sc = 1.0 / (J.atol + J.rtol * maxabs);
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < DIM; i++) {
        double q = yerr[i] * inv_sc;
        e += q * q;
    }
    e = e / (double)DIM;
    if (!std::isfinite(maxabs)) e = NAN;
    double powertwo;
    if (e != e) {
        powertwo = NAN;
    } else {
        int p = -3;
#pragma unroll
        for (int j = 0; j < 6; j++) p += (e <= J.bound2[j]) ? 1 : 0;
        powertwo = (double)p;
    }
    if (powertwo >= 0.0 || accept) {
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            L.a1[i] = a[i];  // (v1, F7)
            L.y[i] = yn1[i];
        }
        L.to += dt_one;
        L.accepted++;
    } else {
        L.rejected++;
    }
    L.dt_one = (powertwo != powertwo) ? NAN : std::ldexp(dt_one, (int)powertwo);
    dopr54_emit_completed(L, J, sink);
    if (L.ii >= J.nt - 1) return true;
    if (L.accepted + L.rejected >= J.max_attempts) {
        abandon_orbit(L.ii, J.nt, L.status, STATUS_STEP_LIMIT, sink);
        return true;
    }
    return false;
}

}  // namespace drift
