// dopr54_rkn.hpp -- GENERATED by gen_dopr54_rkn.py; do not edit.
//
// galpy DOPR54 (src/integrators/galpy/dopr54.rs:486-521) in second-order
// form: stage positions x0 + dt c_i v0 + dt^2 sum_m q_im F_m, the update's
// position x0 + dt v0 + dt^2 sum_m bq_m F_m and the error estimate's
// position dt^2 sum_m eq_m F_m; exact rational products rounded once.
#pragma once

namespace drift {
namespace rkn54 {

constexpr double q31 = 0x1.70a3d70a3d70ap-5;  // (A^2)_31 = 9/200
constexpr double q41 = -0x1.eb851eb851eb8p-2;  // (A^2)_41 = -12/25
constexpr double q42 = 0x1.999999999999ap-1;  // (A^2)_42 = 4/5
constexpr double q51 = -0x1.dde5db8e88cc7p+0;  // (A^2)_51 = -12248/6561
constexpr double q52 = 0x1.a5de0dd001df7p+1;  // (A^2)_52 = 7208/2187
constexpr double q53 = -0x1.08b37c1e1f4afp+0;  // (A^2)_53 = -6784/6561
constexpr double q61 = -0x1.026c9b26c9b27p+1;  // (A^2)_61 = -533/264
constexpr double q62 = 0x1.08ba2e8ba2e8cp+2;  // (A^2)_62 = 91/22
constexpr double q63 = -0x1.b26c9b26c9b27p+0;  // (A^2)_63 = -56/33
constexpr double q64 = 0x1.45d1745d1745dp-4;  // (A^2)_64 = 7/88
constexpr double bq1 = 0x1.7555555555555p-4;  // sum_i b_i a_i1 = 35/384
constexpr double bq2 = 0x0.0p+0;  // sum_i b_i a_i2 = 0
constexpr double bq3 = 0x1.4203385a29dc9p-2;  // sum_i b_i a_i3 = 50/159
constexpr double bq4 = 0x1.0aaaaaaaaaaabp-3;  // sum_i b_i a_i4 = 25/192
constexpr double bq5 = -0x1.256f1826a439fp-5;  // sum_i b_i a_i5 = -243/6784
constexpr double eq1 = 0x1.5b97530eca864p-9;  // sum_i e_i a_i1 = 611/230400
constexpr double eq2 = 0x0.0p+0;  // sum_i e_i a_i2 = 0
constexpr double eq3 = -0x1.938a3f671a2ffp-8;  // sum_i e_i a_i3 = -514/83475
constexpr double eq4 = 0x1.4da740da740dap-7;  // sum_i e_i a_i4 = 391/38400
constexpr double eq5 = -0x1.be0505fd4b8b5p-9;  // sum_i e_i a_i5 = -4617/1356800
constexpr double eq6 = -0x1.ad1ad1ad1ad1bp-9;  // sum_i e_i a_i6 = -11/3360

}  // namespace rkn54
}  // namespace drift
