// dop853_rkn.hpp -- GENERATED by gen_dop853_rkn.py from the DOP853 tableau
// in integrators.hpp (namespace hw853); do not edit by hand.
//
// DOP853 for y = (x, v), y' = (v, F(x)): every combination of stage
// derivatives written with the forces F_m only (stages 1..12, 13 = FSAL
// derivative at the new point, 14..16 the dense-output stages):
//   stage positions  X_i = x0 + h c_i v0 + h^2 stage_i(F)
//   update           sum_j b_j k_j = (v0 + h bpos(F), bvel(F))
//   error estimates  sum_j er_j k_j = (h errpos(F), errvel(F)),
//                    k4 - bhh1 k1 - bhh2 k9 - bhh3 k12
//                      = (h bhhpos(F), bvel(F) - bhhvel(F))
//   dense output     sum_j d_kj k_j = (h dense_k pos(F), dense_k vel(F)),
//                    k = 4..7 (rcont5..8 without their factor h)
// F is indexed F[m][c] (m = stage, c = component); C points at kCoef (or
// a device copy of it), the coefficients in order of use.
#pragma once

#include "drift_device.hpp"

namespace drift {
namespace rkn853 {

constexpr int kNumCoef = 225;
constexpr double kCoef[kNumCoef] = {
    0.05260015195876773,  // [0] stage2: c
    0.0789002279381516,  // [1] stage3: c
    0.003112622984346139,  // [2] stage3: F1
    0.1183503419072274,  // [3] stage4: c
    0.0017508504286947032,  // [4] stage4: F1
    0.00525255128608411,  // [5] stage4: F2
    0.2816496580927726,  // [6] stage5: c
    0.009915816237971964,  // [7] stage5: F1
    -0.05234336665618131,  // [8] stage5: F2
    0.0820908153700972,  // [9] stage5: F3
    0.3333333333333333,  // [10] stage6: c
    0.035337931304886355,  // [11] stage6: F1
    -0.09581915952175495,  // [12] stage6: F3
    0.11603678377242416,  // [13] stage6: F4
    0.25,  // [14] stage7: c
    0.018920483189113914,  // [15] stage7: F1
    -0.03815245266364541,  // [16] stage7: F3
    0.05268745617004205,  // [17] stage7: F4
    -0.0022054866955105506,  // [18] stage7: F5
    0.3076923076923077,  // [19] stage8: c
    0.03067021189623757,  // [20] stage8: F1
    -0.07975482628537983,  // [21] stage8: F3
    0.09799120567724122,  // [22] stage8: F4
    -0.0014238754814449106,  // [23] stage8: F5
    -0.00014543770014516837,  // [24] stage8: F6
    0.6512820512820513,  // [25] stage9: c
    -0.15229089727622042,  // [26] stage9: F1
    0.46966087733519885,  // [27] stage9: F3
    -0.0681414047027942,  // [28] stage9: F4
    0.010711857531715373,  // [29] stage9: F5
    0.3119698547460421,  // [30] stage9: F6
    -0.3598261324728635,  // [31] stage9: F7
    0.6,  // [32] stage10: c
    -0.11020716722377502,  // [33] stage10: F1
    0.30128953154727944,  // [34] stage10: F3
    0.07865735349516378,  // [35] stage10: F4
    0.030838873981239478,  // [36] stage10: F5
    -0.319604147551303,  // [37] stage10: F6
    -0.6851760518136165,  // [38] stage10: F7
    0.8842016075650119,  // [39] stage10: F8
    0.8571428571428571,  // [40] stage11: c
    0.3721894995071434,  // [41] stage11: F1
    -0.5050736261155677,  // [42] stage11: F3
    -0.46149874648558276,  // [43] stage11: F4
    -0.06518509348541916,  // [44] stage11: F5
    4.09803840386656,  // [45] stage11: F6
    3.8922102844072395,  // [46] stage11: F7
    -7.025278165955343,  // [47] stage11: F8
    0.06194438303648046,  // [48] stage11: F9
    1.0,  // [49] stage12: c
    -0.7650750098382328,  // [50] stage12: F1
    0.8347994820739503,  // [51] stage12: F3
    1.755944144193115,  // [52] stage12: F4
    0.23253698859550603,  // [53] stage12: F5
    11.90371935788187,  // [54] stage12: F6
    -1.9034949405460306,  // [55] stage12: F7
    -10.951226401858388,  // [56] stage12: F8
    1.3530625395360725,  // [57] stage12: F9
    -1.960266160037863,  // [58] stage12: F10
    0.1,  // [59] stage14: c
    0.005054351963776225,  // [60] stage14: F1
    -0.41267006753617186,  // [61] stage14: F6
    -0.11888400830261149,  // [62] stage14: F7
    0.511250400071805,  // [63] stage14: F8
    -0.05239733207549386,  // [64] stage14: F9
    0.06981946420649213,  // [65] stage14: F10
    0.0031982003615129264,  // [66] stage14: F11
    -0.000371008689309057,  // [67] stage14: F12
    0.2,  // [68] stage15: c
    0.009887781381457055,  // [69] stage15: F1
    -0.007602503134799441,  // [70] stage15: F6
    0.04742334588757416,  // [71] stage15: F7
    -0.03637906592918709,  // [72] stage15: F8
    -0.021320408140087014,  // [73] stage15: F9
    0.026772748548590856,  // [74] stage15: F10
    0.0013364963323445063,  // [75] stage15: F11
    0.001054215711719072,  // [76] stage15: F12
    -0.0011726106576121006,  // [77] stage15: F13
    0.7777777777777778,  // [78] stage16: c
    0.3588663399722558,  // [79] stage16: F1
    9.38025080291835,  // [80] stage16: F6
    7.477758624830405,  // [81] stage16: F7
    -15.729096304440624,  // [82] stage16: F8
    -0.3664913832247408,  // [83] stage16: F9
    0.4520427193280568,  // [84] stage16: F10
    0.024882489771726926,  // [85] stage16: F11
    0.018743046880517162,  // [86] stage16: F12
    -0.02134289656466091,  // [87] stage16: F13
    -1.293144303668819,  // [88] stage16: F14
    0.054293734116568765,  // [89] bvel: F1
    4.450312892752409,  // [90] bvel: F6
    1.8915178993145003,  // [91] bvel: F7
    -5.801203960010585,  // [92] bvel: F8
    0.3111643669578199,  // [93] bvel: F9
    -0.1521609496625161,  // [94] bvel: F10
    0.20136540080403034,  // [95] bvel: F11
    0.04471061572777259,  // [96] bvel: F12
    0.054293734116568765,  // [97] bpos: F1
    2.9668752618349394,  // [98] bpos: F6
    1.4186384244858752,  // [99] bpos: F7
    -4.016218126161174,  // [100] bpos: F8
    0.10850859975965002,  // [101] bpos: F9
    -0.06086437986500643,  // [102] bpos: F10
    0.028766485829147193,  // [103] bpos: F11
    0.01312004499419488,  // [104] errvel: F1
    -1.2251564463762044,  // [105] errvel: F6
    -0.4957589496572502,  // [106] errvel: F7
    1.6643771824549864,  // [107] errvel: F8
    -0.35032884874997366,  // [108] errvel: F9
    0.3341791187130175,  // [109] errvel: F10
    0.08192320648511571,  // [110] errvel: F11
    -0.022355307863886294,  // [111] errvel: F12
    -0.18865188001798797,  // [112] errpos: F1
    0.9962151331241325,  // [113] errpos: F4
    0.23599761577747436,  // [114] errpos: F5
    -2.854630682350911,  // [115] errpos: F6
    -4.0828088279337065,  // [116] errpos: F7
    6.038341535948958,  // [117] errpos: F8
    0.39584534789657555,  // [118] errpos: F9
    -0.5259249995299615,  // [119] errpos: F10
    -0.014383242914573597,  // [120] errpos: F11
    0.2440944881889764,  // [121] bhhvel: F1
    0.7338466882816118,  // [122] bhhvel: F9
    0.022058823529411766,  // [123] bhhvel: F12
    -0.4538545734291735,  // [124] bhhpos: F1
    2.6987585022618616,  // [125] bhhpos: F4
    0.6812767760191379,  // [126] bhhpos: F5
    -16.885342816594818,  // [127] bhhpos: F6
    -13.987876814514605,  // [128] bhhpos: F7
    27.96175549233409,  // [129] bhhpos: F8
    0.3042333850581198,  // [130] bhhpos: F9
    -0.3335239499192925,  // [131] bhhpos: F10
    0.01457399878468182,  // [132] bhhpos: F11
    -8.428938276109013,  // [133] dense4vel_pre: F1
    0.5667149535193777,  // [134] dense4vel_pre: F6
    -3.0689499459498917,  // [135] dense4vel_pre: F7
    2.38466765651207,  // [136] dense4vel_pre: F8
    2.117034582445028,  // [137] dense4vel_pre: F9
    -0.871391583777973,  // [138] dense4vel_pre: F10
    2.2404374302607883,  // [139] dense4vel_pre: F11
    0.6315787787694688,  // [140] dense4vel_pre: F12
    2.8611698796830125,  // [141] dense4pos_pre: F1
    30.51215583852884,  // [142] dense4pos_pre: F6
    -24.61264508865549,  // [143] dense4pos_pre: F7
    -35.41960917471137,  // [144] dense4pos_pre: F8
    -3.863458628138401,  // [145] dense4pos_pre: F9
    3.7761446592257837,  // [146] dense4pos_pre: F10
    0.5382666867527854,  // [147] dense4pos_pre: F11
    0.12984961919289312,  // [148] dense4pos_pre: F12
    10.427508642579134,  // [149] dense5vel_pre: F1
    242.28349177525817,  // [150] dense5vel_pre: F6
    165.20045171727028,  // [151] dense5vel_pre: F7
    -374.5467547226902,  // [152] dense5vel_pre: F8
    -22.113666853125306,  // [153] dense5vel_pre: F9
    7.733432668472264,  // [154] dense5vel_pre: F10
    -30.674084731089398,  // [155] dense5vel_pre: F11
    -9.332130526430229,  // [156] dense5vel_pre: F12
    -17.921210306702967,  // [157] dense5pos_pre: F1
    -124.2260787072285,  // [158] dense5pos_pre: F6
    273.13795938281214,  // [159] dense5pos_pre: F7
    96.31318087284055,  // [160] dense5pos_pre: F8
    27.684870153300384,  // [161] dense5pos_pre: F9
    -29.05563647508469,  // [162] dense5pos_pre: F10
    -3.0977069495595075,  // [163] dense5pos_pre: F11
    0.4625952063992359,  // [164] dense5pos_pre: F12
    19.985053242002433,  // [165] dense6vel_pre: F1
    -387.0373087493518,  // [166] dense6vel_pre: F6
    -189.17813819516758,  // [167] dense6vel_pre: F7
    527.8081592054236,  // [168] dense6vel_pre: F8
    -11.57390253995963,  // [169] dense6vel_pre: F9
    6.8812326946963,  // [170] dense6vel_pre: F10
    -1.0006050966910838,  // [171] dense6vel_pre: F11
    0.7777137798053443,  // [172] dense6vel_pre: F12
    -8.999411590110814,  // [173] dense6pos_pre: F1
    -250.14465065648488,  // [174] dense6pos_pre: F6
    -7.345168174338206,  // [175] dense6pos_pre: F7
    324.4120585596325,  // [176] dense6pos_pre: F8
    1.3538496519789445,  // [177] dense6pos_pre: F9
    3.8566662410202324,  // [178] dense6pos_pre: F10
    -0.5618712467427652,  // [179] dense6pos_pre: F11
    -0.5475191693050367,  // [180] dense6pos_pre: F12
    -25.69393346270375,  // [181] dense7vel_pre: F1
    -154.18974869023643,  // [182] dense7vel_pre: F6
    -231.5293791760455,  // [183] dense7vel_pre: F7
    357.6391179106141,  // [184] dense7vel_pre: F8
    93.40532418362432,  // [185] dense7vel_pre: F9
    -37.45832313645163,  // [186] dense7vel_pre: F10
    104.0996495089623,  // [187] dense7vel_pre: F11
    29.8402934266605,  // [188] dense7vel_pre: F12
    75.66260800528795,  // [189] dense7pos_pre: F1
    904.821881520238,  // [190] dense7pos_pre: F6
    -991.2456491432964,  // [191] dense7pos_pre: F7
    -911.6644928639352,  // [192] dense7pos_pre: F8
    -83.34399475607842,  // [193] dense7pos_pre: F9
    73.06569397631904,  // [194] dense7pos_pre: F10
    11.227103851084003,  // [195] dense7pos_pre: F11
    -1.2324213700720028,  // [196] dense7pos_pre: F12
    -0.08899033645133331,  // [197] dense4vel_post: F13
    18.148505520854727,  // [198] dense4vel_post: F14
    -9.194632392478356,  // [199] dense4vel_post: F15
    -4.436036387594894,  // [200] dense4vel_post: F16
    -0.14125972611089993,  // [201] dense4pos_post: F13
    -14.374598829726839,  // [202] dense4pos_post: F14
    40.59398476395968,  // [203] dense4pos_post: F15
    15.697238121770845,  // [204] dense5vel_post: F13
    -31.139403219565178,  // [205] dense5vel_post: F14
    -9.35292435884448,  // [206] dense5vel_post: F15
    35.81684148639408,  // [207] dense5vel_post: F16
    0.21147048462990323,  // [208] dense5pos_post: F13
    104.24898538523482,  // [209] dense5pos_post: F14
    -327.75842904664137,  // [210] dense5pos_post: F15
    -2.778205752353508,  // [211] dense6vel_post: F13
    -60.19669523126412,  // [212] dense6vel_post: F14
    84.32040550667716,  // [213] dense6vel_post: F15
    11.99229113618279,  // [214] dense6vel_post: F16
    0.45402652434053137,  // [215] dense6pos_post: F13
    47.262978033508965,  // [216] dense6pos_post: F14
    -109.74095817349945,  // [217] dense6pos_post: F15
    -43.53345659001114,  // [218] dense7vel_post: F13
    96.32455395918828,  // [219] dense7vel_post: F14
    -39.17726167561544,  // [220] dense7vel_post: F15
    -149.72683625798564,  // [221] dense7vel_post: F16
    -0.5764911999188784,  // [222] dense7pos_post: F13
    -446.85829878732994,  // [223] dense7pos_post: F14
    1370.1440607677016,  // [224] dense7pos_post: F15
};

#if defined(__HIPCC__)
// the device copy, read with scalar loads (see dop853_coef)
static __constant__ double kCoefDevice[kNumCoef] = {
    0.05260015195876773,
    0.0789002279381516,
    0.003112622984346139,
    0.1183503419072274,
    0.0017508504286947032,
    0.00525255128608411,
    0.2816496580927726,
    0.009915816237971964,
    -0.05234336665618131,
    0.0820908153700972,
    0.3333333333333333,
    0.035337931304886355,
    -0.09581915952175495,
    0.11603678377242416,
    0.25,
    0.018920483189113914,
    -0.03815245266364541,
    0.05268745617004205,
    -0.0022054866955105506,
    0.3076923076923077,
    0.03067021189623757,
    -0.07975482628537983,
    0.09799120567724122,
    -0.0014238754814449106,
    -0.00014543770014516837,
    0.6512820512820513,
    -0.15229089727622042,
    0.46966087733519885,
    -0.0681414047027942,
    0.010711857531715373,
    0.3119698547460421,
    -0.3598261324728635,
    0.6,
    -0.11020716722377502,
    0.30128953154727944,
    0.07865735349516378,
    0.030838873981239478,
    -0.319604147551303,
    -0.6851760518136165,
    0.8842016075650119,
    0.8571428571428571,
    0.3721894995071434,
    -0.5050736261155677,
    -0.46149874648558276,
    -0.06518509348541916,
    4.09803840386656,
    3.8922102844072395,
    -7.025278165955343,
    0.06194438303648046,
    1.0,
    -0.7650750098382328,
    0.8347994820739503,
    1.755944144193115,
    0.23253698859550603,
    11.90371935788187,
    -1.9034949405460306,
    -10.951226401858388,
    1.3530625395360725,
    -1.960266160037863,
    0.1,
    0.005054351963776225,
    -0.41267006753617186,
    -0.11888400830261149,
    0.511250400071805,
    -0.05239733207549386,
    0.06981946420649213,
    0.0031982003615129264,
    -0.000371008689309057,
    0.2,
    0.009887781381457055,
    -0.007602503134799441,
    0.04742334588757416,
    -0.03637906592918709,
    -0.021320408140087014,
    0.026772748548590856,
    0.0013364963323445063,
    0.001054215711719072,
    -0.0011726106576121006,
    0.7777777777777778,
    0.3588663399722558,
    9.38025080291835,
    7.477758624830405,
    -15.729096304440624,
    -0.3664913832247408,
    0.4520427193280568,
    0.024882489771726926,
    0.018743046880517162,
    -0.02134289656466091,
    -1.293144303668819,
    0.054293734116568765,
    4.450312892752409,
    1.8915178993145003,
    -5.801203960010585,
    0.3111643669578199,
    -0.1521609496625161,
    0.20136540080403034,
    0.04471061572777259,
    0.054293734116568765,
    2.9668752618349394,
    1.4186384244858752,
    -4.016218126161174,
    0.10850859975965002,
    -0.06086437986500643,
    0.028766485829147193,
    0.01312004499419488,
    -1.2251564463762044,
    -0.4957589496572502,
    1.6643771824549864,
    -0.35032884874997366,
    0.3341791187130175,
    0.08192320648511571,
    -0.022355307863886294,
    -0.18865188001798797,
    0.9962151331241325,
    0.23599761577747436,
    -2.854630682350911,
    -4.0828088279337065,
    6.038341535948958,
    0.39584534789657555,
    -0.5259249995299615,
    -0.014383242914573597,
    0.2440944881889764,
    0.7338466882816118,
    0.022058823529411766,
    -0.4538545734291735,
    2.6987585022618616,
    0.6812767760191379,
    -16.885342816594818,
    -13.987876814514605,
    27.96175549233409,
    0.3042333850581198,
    -0.3335239499192925,
    0.01457399878468182,
    -8.428938276109013,
    0.5667149535193777,
    -3.0689499459498917,
    2.38466765651207,
    2.117034582445028,
    -0.871391583777973,
    2.2404374302607883,
    0.6315787787694688,
    2.8611698796830125,
    30.51215583852884,
    -24.61264508865549,
    -35.41960917471137,
    -3.863458628138401,
    3.7761446592257837,
    0.5382666867527854,
    0.12984961919289312,
    10.427508642579134,
    242.28349177525817,
    165.20045171727028,
    -374.5467547226902,
    -22.113666853125306,
    7.733432668472264,
    -30.674084731089398,
    -9.332130526430229,
    -17.921210306702967,
    -124.2260787072285,
    273.13795938281214,
    96.31318087284055,
    27.684870153300384,
    -29.05563647508469,
    -3.0977069495595075,
    0.4625952063992359,
    19.985053242002433,
    -387.0373087493518,
    -189.17813819516758,
    527.8081592054236,
    -11.57390253995963,
    6.8812326946963,
    -1.0006050966910838,
    0.7777137798053443,
    -8.999411590110814,
    -250.14465065648488,
    -7.345168174338206,
    324.4120585596325,
    1.3538496519789445,
    3.8566662410202324,
    -0.5618712467427652,
    -0.5475191693050367,
    -25.69393346270375,
    -154.18974869023643,
    -231.5293791760455,
    357.6391179106141,
    93.40532418362432,
    -37.45832313645163,
    104.0996495089623,
    29.8402934266605,
    75.66260800528795,
    904.821881520238,
    -991.2456491432964,
    -911.6644928639352,
    -83.34399475607842,
    73.06569397631904,
    11.227103851084003,
    -1.2324213700720028,
    -0.08899033645133331,
    18.148505520854727,
    -9.194632392478356,
    -4.436036387594894,
    -0.14125972611089993,
    -14.374598829726839,
    40.59398476395968,
    15.697238121770845,
    -31.139403219565178,
    -9.35292435884448,
    35.81684148639408,
    0.21147048462990323,
    104.24898538523482,
    -327.75842904664137,
    -2.778205752353508,
    -60.19669523126412,
    84.32040550667716,
    11.99229113618279,
    0.45402652434053137,
    47.262978033508965,
    -109.74095817349945,
    -43.53345659001114,
    96.32455395918828,
    -39.17726167561544,
    -149.72683625798564,
    -0.5764911999188784,
    -446.85829878732994,
    1370.1440607677016,
};
#endif

constexpr int k_c2 = 0;
constexpr int k_c3 = 1;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage3(const FA& F, int c, CP C) {
    return C[2] * F[1][c];
}
constexpr int k_c4 = 3;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage4(const FA& F, int c, CP C) {
    return C[4] * F[1][c] + C[5] * F[2][c];
}
constexpr int k_c5 = 6;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage5(const FA& F, int c, CP C) {
    return C[7] * F[1][c] + C[8] * F[2][c] + C[9] * F[3][c];
}
constexpr int k_c6 = 10;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage6(const FA& F, int c, CP C) {
    return C[11] * F[1][c] + C[12] * F[3][c] + C[13] * F[4][c];
}
constexpr int k_c7 = 14;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage7(const FA& F, int c, CP C) {
    return C[15] * F[1][c] + C[16] * F[3][c] + C[17] * F[4][c] + C[18] * F[5][c];
}
constexpr int k_c8 = 19;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage8(const FA& F, int c, CP C) {
    return C[20] * F[1][c] + C[21] * F[3][c] + C[22] * F[4][c] + C[23] * F[5][c] + C[24] * F[6][c];
}
constexpr int k_c9 = 25;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage9(const FA& F, int c, CP C) {
    return C[26] * F[1][c] + C[27] * F[3][c] + C[28] * F[4][c] + C[29] * F[5][c] + C[30] * F[6][c] + C[31] * F[7][c];
}
constexpr int k_c10 = 32;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage10(const FA& F, int c, CP C) {
    return C[33] * F[1][c] + C[34] * F[3][c] + C[35] * F[4][c] + C[36] * F[5][c] + C[37] * F[6][c] + C[38] * F[7][c] + C[39] * F[8][c];
}
constexpr int k_c11 = 40;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage11(const FA& F, int c, CP C) {
    return C[41] * F[1][c] + C[42] * F[3][c] + C[43] * F[4][c] + C[44] * F[5][c] + C[45] * F[6][c] + C[46] * F[7][c] + C[47] * F[8][c] + C[48] * F[9][c];
}
constexpr int k_c12 = 49;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage12(const FA& F, int c, CP C) {
    return C[50] * F[1][c] + C[51] * F[3][c] + C[52] * F[4][c] + C[53] * F[5][c] + C[54] * F[6][c] + C[55] * F[7][c] + C[56] * F[8][c] + C[57] * F[9][c] + C[58] * F[10][c];
}
constexpr int k_c14 = 59;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage14(const FA& F, int c, CP C) {
    return C[60] * F[1][c] + C[61] * F[6][c] + C[62] * F[7][c] + C[63] * F[8][c] + C[64] * F[9][c] + C[65] * F[10][c] + C[66] * F[11][c] + C[67] * F[12][c];
}
constexpr int k_c15 = 68;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage15(const FA& F, int c, CP C) {
    return C[69] * F[1][c] + C[70] * F[6][c] + C[71] * F[7][c] + C[72] * F[8][c] + C[73] * F[9][c] + C[74] * F[10][c] + C[75] * F[11][c] + C[76] * F[12][c] + C[77] * F[13][c];
}
constexpr int k_c16 = 78;
template <class FA, class CP>
DRIFT_HD DRIFT_FI double stage16(const FA& F, int c, CP C) {
    return C[79] * F[1][c] + C[80] * F[6][c] + C[81] * F[7][c] + C[82] * F[8][c] + C[83] * F[9][c] + C[84] * F[10][c] + C[85] * F[11][c] + C[86] * F[12][c] + C[87] * F[13][c] + C[88] * F[14][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double bvel(const FA& F, int c, CP C) {
    return C[89] * F[1][c] + C[90] * F[6][c] + C[91] * F[7][c] + C[92] * F[8][c] + C[93] * F[9][c] + C[94] * F[10][c] + C[95] * F[11][c] + C[96] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double bpos(const FA& F, int c, CP C) {
    return C[97] * F[1][c] + C[98] * F[6][c] + C[99] * F[7][c] + C[100] * F[8][c] + C[101] * F[9][c] + C[102] * F[10][c] + C[103] * F[11][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double errvel(const FA& F, int c, CP C) {
    return C[104] * F[1][c] + C[105] * F[6][c] + C[106] * F[7][c] + C[107] * F[8][c] + C[108] * F[9][c] + C[109] * F[10][c] + C[110] * F[11][c] + C[111] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double errpos(const FA& F, int c, CP C) {
    return C[112] * F[1][c] + C[113] * F[4][c] + C[114] * F[5][c] + C[115] * F[6][c] + C[116] * F[7][c] + C[117] * F[8][c] + C[118] * F[9][c] + C[119] * F[10][c] + C[120] * F[11][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double bhhvel(const FA& F, int c, CP C) {
    return C[121] * F[1][c] + C[122] * F[9][c] + C[123] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double bhhpos(const FA& F, int c, CP C) {
    return C[124] * F[1][c] + C[125] * F[4][c] + C[126] * F[5][c] + C[127] * F[6][c] + C[128] * F[7][c] + C[129] * F[8][c] + C[130] * F[9][c] + C[131] * F[10][c] + C[132] * F[11][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense4vel_pre(const FA& F, int c, CP C) {
    return C[133] * F[1][c] + C[134] * F[6][c] + C[135] * F[7][c] + C[136] * F[8][c] + C[137] * F[9][c] + C[138] * F[10][c] + C[139] * F[11][c] + C[140] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense4pos_pre(const FA& F, int c, CP C) {
    return C[141] * F[1][c] + C[142] * F[6][c] + C[143] * F[7][c] + C[144] * F[8][c] + C[145] * F[9][c] + C[146] * F[10][c] + C[147] * F[11][c] + C[148] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense5vel_pre(const FA& F, int c, CP C) {
    return C[149] * F[1][c] + C[150] * F[6][c] + C[151] * F[7][c] + C[152] * F[8][c] + C[153] * F[9][c] + C[154] * F[10][c] + C[155] * F[11][c] + C[156] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense5pos_pre(const FA& F, int c, CP C) {
    return C[157] * F[1][c] + C[158] * F[6][c] + C[159] * F[7][c] + C[160] * F[8][c] + C[161] * F[9][c] + C[162] * F[10][c] + C[163] * F[11][c] + C[164] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense6vel_pre(const FA& F, int c, CP C) {
    return C[165] * F[1][c] + C[166] * F[6][c] + C[167] * F[7][c] + C[168] * F[8][c] + C[169] * F[9][c] + C[170] * F[10][c] + C[171] * F[11][c] + C[172] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense6pos_pre(const FA& F, int c, CP C) {
    return C[173] * F[1][c] + C[174] * F[6][c] + C[175] * F[7][c] + C[176] * F[8][c] + C[177] * F[9][c] + C[178] * F[10][c] + C[179] * F[11][c] + C[180] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense7vel_pre(const FA& F, int c, CP C) {
    return C[181] * F[1][c] + C[182] * F[6][c] + C[183] * F[7][c] + C[184] * F[8][c] + C[185] * F[9][c] + C[186] * F[10][c] + C[187] * F[11][c] + C[188] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense7pos_pre(const FA& F, int c, CP C) {
    return C[189] * F[1][c] + C[190] * F[6][c] + C[191] * F[7][c] + C[192] * F[8][c] + C[193] * F[9][c] + C[194] * F[10][c] + C[195] * F[11][c] + C[196] * F[12][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense4vel_post(const FA& F, int c, CP C) {
    return C[197] * F[13][c] + C[198] * F[14][c] + C[199] * F[15][c] + C[200] * F[16][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense4pos_post(const FA& F, int c, CP C) {
    return C[201] * F[13][c] + C[202] * F[14][c] + C[203] * F[15][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense5vel_post(const FA& F, int c, CP C) {
    return C[204] * F[13][c] + C[205] * F[14][c] + C[206] * F[15][c] + C[207] * F[16][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense5pos_post(const FA& F, int c, CP C) {
    return C[208] * F[13][c] + C[209] * F[14][c] + C[210] * F[15][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense6vel_post(const FA& F, int c, CP C) {
    return C[211] * F[13][c] + C[212] * F[14][c] + C[213] * F[15][c] + C[214] * F[16][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense6pos_post(const FA& F, int c, CP C) {
    return C[215] * F[13][c] + C[216] * F[14][c] + C[217] * F[15][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense7vel_post(const FA& F, int c, CP C) {
    return C[218] * F[13][c] + C[219] * F[14][c] + C[220] * F[15][c] + C[221] * F[16][c];
}
template <class FA, class CP>
DRIFT_HD DRIFT_FI double dense7pos_post(const FA& F, int c, CP C) {
    return C[222] * F[13][c] + C[223] * F[14][c] + C[224] * F[15][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage3_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage4_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[2][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage5_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[2][c] + L[3] * F[3][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage6_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage7_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage8_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c] + L[5] * F[6][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage9_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c] + L[5] * F[6][c] + L[6] * F[7][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage10_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c] + L[5] * F[6][c] + L[6] * F[7][c] + L[7] * F[8][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage11_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c] + L[5] * F[6][c] + L[6] * F[7][c] + L[7] * F[8][c] + L[8] * F[9][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage12_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[3][c] + L[3] * F[4][c] + L[4] * F[5][c] + L[5] * F[6][c] + L[6] * F[7][c] + L[7] * F[8][c] + L[8] * F[9][c] + L[9] * F[10][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage14_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[6][c] + L[3] * F[7][c] + L[4] * F[8][c] + L[5] * F[9][c] + L[6] * F[10][c] + L[7] * F[11][c] + L[8] * F[12][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage15_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[6][c] + L[3] * F[7][c] + L[4] * F[8][c] + L[5] * F[9][c] + L[6] * F[10][c] + L[7] * F[11][c] + L[8] * F[12][c] + L[9] * F[13][c];
}
template <class FA>
DRIFT_HD DRIFT_FI double stage16_local(const FA& F, int c, const double* L) {
    return L[1] * F[1][c] + L[2] * F[6][c] + L[3] * F[7][c] + L[4] * F[8][c] + L[5] * F[9][c] + L[6] * F[10][c] + L[7] * F[11][c] + L[8] * F[12][c] + L[9] * F[13][c] + L[10] * F[14][c];
}

// number of coefficients of stage i (c_i and its terms), from kStageC[i]
constexpr int kStageN[17] = {0, 0, 1, 2, 3, 4, 4, 5, 6, 7, 8, 9, 10, 0, 9, 10, 11};

// index of c_i in kCoef for stage i (-1: no such stage)
constexpr int kStageC[17] = {-1, -1, k_c2, k_c3, k_c4, k_c5, k_c6, k_c7, k_c8, k_c9, k_c10, k_c11, k_c12, -1, k_c14, k_c15, k_c16};

// sum_m (A^2)_im F_m[c] for stage I (0 for stage 2)
template <int I, class FA, class CP>
DRIFT_HD DRIFT_FI double stage_sum(const FA& F, int c, CP C) {
    if constexpr (I == 2) return 0.0;
    if constexpr (I == 3) return stage3(F, c, C);
    if constexpr (I == 4) return stage4(F, c, C);
    if constexpr (I == 5) return stage5(F, c, C);
    if constexpr (I == 6) return stage6(F, c, C);
    if constexpr (I == 7) return stage7(F, c, C);
    if constexpr (I == 8) return stage8(F, c, C);
    if constexpr (I == 9) return stage9(F, c, C);
    if constexpr (I == 10) return stage10(F, c, C);
    if constexpr (I == 11) return stage11(F, c, C);
    if constexpr (I == 12) return stage12(F, c, C);
    if constexpr (I == 14) return stage14(F, c, C);
    if constexpr (I == 15) return stage15(F, c, C);
    if constexpr (I == 16) return stage16(F, c, C);
}

// the same from a local copy L of the stage's coefficients (L[0] = c_I)
template <int I, class FA>
DRIFT_HD DRIFT_FI double stage_sum_local(const FA& F, int c, const double* L) {
    if constexpr (I == 2) return 0.0;
    if constexpr (I == 3) return stage3_local(F, c, L);
    if constexpr (I == 4) return stage4_local(F, c, L);
    if constexpr (I == 5) return stage5_local(F, c, L);
    if constexpr (I == 6) return stage6_local(F, c, L);
    if constexpr (I == 7) return stage7_local(F, c, L);
    if constexpr (I == 8) return stage8_local(F, c, L);
    if constexpr (I == 9) return stage9_local(F, c, L);
    if constexpr (I == 10) return stage10_local(F, c, L);
    if constexpr (I == 11) return stage11_local(F, c, L);
    if constexpr (I == 12) return stage12_local(F, c, L);
    if constexpr (I == 14) return stage14_local(F, c, L);
    if constexpr (I == 15) return stage15_local(F, c, L);
    if constexpr (I == 16) return stage16_local(F, c, L);
}

}  // namespace rkn853
}  // namespace drift
