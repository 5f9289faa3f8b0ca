// Deterministic FP64 exp, sin and cos for the weight quadrature.
//
// Built from + - * /, truncation and exact power-of-two scaling only, so every IEEE-conforming build
// (device or host, with contraction off) returns the same bits. The near-zero double-layer
// weights of coplanar triangle pairs are rounding residues whose FP16 codes depend on the
// last bit of every exp/sin/cos term; with these functions the device weights equal the CPU
// checker's bit for bit.
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions", 1980): reduce the
// argument by ln 2 (exp) or pi/2 (sin, cos) with the constant split into parts whose
// products with the quotient are exact, then evaluate the Taylor polynomial of the reduced
// argument in Horner form (degree 13 for exp on |r| <= ln2/2, 17/16 for sin/cos on
// |r| <= pi/4: truncation below 2^-53 relative). Valid for |x| < 2^20 (sin, cos) and for
// the full exp range; accuracy about 1 ulp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pppm_amd
{

namespace detm
{
constexpr double LN2_HI = 0.6931471806019545;       // ln 2 rounded to 32 significant bits
constexpr double LN2_LO = -4.2009150726810846e-11;  // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double PIO2_1 = 1.5707963267341256;       // pi/2 rounded to 33 significant bits
constexpr double PIO2_2 = 6.077100506303966e-11;    // next 33 bits
constexpr double PIO2_3 = 2.0222662487959506e-21;   // remainder
constexpr double TWO_OVER_PI = 0.6366197723675814;
// 1/n!
constexpr double F2 = 0.5, F3 = 0.16666666666666666, F4 = 0.041666666666666664, F5 = 0.008333333333333333,
                 F6 = 0.001388888888888889, F7 = 0.0001984126984126984, F8 = 2.48015873015873e-05,
                 F9 = 2.7557319223985893e-06, F10 = 2.755731922398589e-07, F11 = 2.505210838544172e-08,
                 F12 = 2.08767569878681e-09, F13 = 1.6059043836821613e-10, F14 = 1.1470745597729725e-11,
                 F15 = 7.647163731819816e-13, F16 = 4.779477332387385e-14, F17 = 2.8114572543455206e-15;

__host__ __device__ inline double pow2i(int k)  // 2^k for -1022 <= k <= 1023, exact
{
    const uint64_t bits = (uint64_t)(k + 1023) << 52;
    double d;
    __builtin_memcpy(&d, &bits, sizeof d);
    return d;
}
// nearest integer, halves away from zero: trunc(|x| + 0.5) with x's sign, the value of a round
// trip through a 64-bit integer (one instruction per truncation instead of an integer conversion)
__host__ __device__ inline double round_half_away(double x) { return x >= 0 ? __builtin_trunc(x + 0.5) : -__builtin_trunc(-x + 0.5); }
}  // namespace detm

__host__ __device__ inline double det_exp(double x)
{
    using namespace detm;
    if (x != x)
        return x;
    if (x > 709.0)
        return __builtin_huge_val();
    if (x < -745.0)
        return 0.0;
    const double kd = round_half_away(x * INV_LN2);
    const int k = (int)kd;
    const double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = F13;
    p = F12 + r * p;
    p = F11 + r * p;
    p = F10 + r * p;
    p = F9 + r * p;
    p = F8 + r * p;
    p = F7 + r * p;
    p = F6 + r * p;
    p = F5 + r * p;
    p = F4 + r * p;
    p = F3 + r * p;
    p = F2 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    // split the scaling so that results down to the subnormal range stay exact
    if (k < -1000)
        return (p * pow2i(k + 1000)) * pow2i(-1000);
    if (k > 1000)
        return (p * pow2i(k - 1000)) * pow2i(1000);
    return p * pow2i(k);
}

// sin and cos of x together
__host__ __device__ inline void det_sincos(double x, double *s, double *c)
{
    using namespace detm;
    const double qd = round_half_away(x * TWO_OVER_PI);
    const int q = (int)qd;  // |qd| < 2^20 in the valid range
    const double r = ((x - qd * PIO2_1) - qd * PIO2_2) - qd * PIO2_3;
    const double z = r * r;
    // sin r = r - r z (1/3! - z (1/5! - z (1/7! - ... - z/17!)))
    double ps = F17;
    ps = F15 - z * ps;
    ps = F13 - z * ps;
    ps = F11 - z * ps;
    ps = F9 - z * ps;
    ps = F7 - z * ps;
    ps = F5 - z * ps;
    ps = F3 - z * ps;
    const double sr = r - (r * z) * ps;
    // cos r = 1 - z (1/2! - z (1/4! - ... - z/16!))
    double pc = F16;
    pc = F14 - z * pc;
    pc = F12 - z * pc;
    pc = F10 - z * pc;
    pc = F8 - z * pc;
    pc = F6 - z * pc;
    pc = F4 - z * pc;
    pc = F2 - z * pc;
    const double cr = 1.0 - z * pc;
    switch ((int)(q & 3))
    {
    case 0:
        *s = sr;
        *c = cr;
        break;
    case 1:
        *s = cr;
        *c = -sr;
        break;
    case 2:
        *s = -sr;
        *c = -cr;
        break;
    default:
        *s = -cr;
        *c = sr;
        break;
    }
}

}  // namespace pppm_amd
