// trig.hpp -- sine and cosine of a heading angle, shared by host and device code.
//
// The stage kernels turn every updated heading into (cos, sin) once per particle and stage.  The
// angles are already wrapped into [-pi, pi), so the general-purpose library routine (with its
// large-argument reduction path) does more work than needed; this one is exact to about one ulp
// for |x| <= 2^19:
//   * Cody-Waite reduction x = k pi/2 + r, k = rint(2x/pi), with pi/2 split into two 33-bit parts
//     P1, P2 (k P1 and k P2 are exact) and a tail P3, so |r| <= pi/4 carries no reduction error
//     above 1e-30;
//   * the Taylor series of sin to r^17 and of cos to r^18 on |r| <= pi/4 (truncation below 1e-19,
//     coefficients the doubles nearest 1/n!), evaluated by Estrin's scheme with explicit fma, and
//     the leading 1 - r^2/2 of cos carried with its rounding error;
//   * quadrant k mod 4 selects and negates.
// Host and device evaluate the same operations (explicit fma, no contraction), so their results
// are bitwise equal.  Measured against extended-precision sinl/cosl: at most 1.4 ulp on [-4, 4].
#pragma once

#include <cmath>
#include <cstring>

#if defined(__HIPCC__)
#define SPP_HD __host__ __device__
#else
#define SPP_HD
#endif

namespace spp {

// fma(a, k, c) for a constant k, as one VOP3 instruction on the device.  Written as fma(k, z, c)
// with two literal constants, the compiler keeps one of them in a VGPR and copies it into the
// accumulator of a v_fmac before every term (an extra VALU operation per term); here k lives in an
// SGPR and c in a VGPR the compiler hoists.  The product and the single rounding are those of fma,
// so the value is bitwise the same on host and device.
SPP_HD inline double fma_k(double a, double k, double c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(k), "v"(c));
    return r;
#else
    return std::fma(a, k, c);
#endif
}

// (-1)^k x: the parity of k moved into x's sign bit (the same bits as k odd ? -x : x, without the
// compare and select).
SPP_HD inline double negate_if_odd(double x, int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __hiloint2double(__double2hiint(x) ^ (int)((unsigned)k << 31), __double2loint(x));
#else
    unsigned long long b;
    std::memcpy(&b, &x, sizeof b);
    b ^= (unsigned long long)((unsigned)k & 1u) << 63;
    std::memcpy(&x, &b, sizeof x);
    return x;
#endif
}

SPP_HD inline void sincos_heading(double x, double* sn, double* cs)
{
    constexpr double kTwoOverPi = 0.6366197723675814;
    constexpr double P1 = 1.5707963267341256;      // leading 33 bits of pi/2
    constexpr double P2 = 6.077100506303966e-11;   // next 33 bits
    constexpr double P3 = 2.0222662487959506e-21;  // pi/2 - P1 - P2, rounded
    const double kf = rint(x * kTwoOverPi);
    double r = fma(-kf, P1, x);
    r = fma(-kf, P2, r);
    r = fma(-kf, P3, r);
    const double z = r * r;
    // sin r = r + r z P(z), P = S1 + S2 z + ... + S8 z^7, S_m = (-1)^m / (2m+1)!, and
    // cos r = (1 - z/2) + z^2 Q(z), Q = C2 + C3 z + ... + C9 z^7, C_m = (-1)^m / (2m)!, both by
    // Estrin's scheme (pairs, then z^2, then z^4) for a short dependency chain
    const double z2 = z * z;
    const double z4 = z2 * z2;
    const double p01 = fma_k(z, 0.008333333333333333, -0.16666666666666666);
    const double p23 = fma_k(z, 2.7557319223985893e-06, -0.0001984126984126984);
    const double p45 = fma_k(z, 1.6059043836821613e-10, -2.505210838544172e-08);
    const double p67 = fma_k(z, 2.8114572543455206e-15, -7.647163731819816e-13);
    const double p = fma(fma(p67, z2, p45), z4, fma(p23, z2, p01));
    const double s = fma(r * z, p, r);
    const double q01 = fma_k(z, -0.001388888888888889, 0.041666666666666664);
    const double q23 = fma_k(z, -2.755731922398589e-07, 2.48015873015873e-05);
    const double q45 = fma_k(z, -1.1470745597729725e-11, 2.08767569878681e-09);
    const double q67 = fma_k(z, -1.5619206968586225e-16, 4.779477332387385e-14);
    const double q = fma(fma(q67, z2, q45), z4, fma(q23, z2, q01));
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    const double c = w + (((1.0 - w) - hz) + z2 * q);
    const int k = (int)kf;
    const double a = (k & 1) ? c : s;  // sin x up to sign
    const double b = (k & 1) ? s : c;  // cos x up to sign
    *sn = (k & 2) ? -a : a;
    *cs = ((k + 1) & 2) ? -b : b;
}

// sin(a - b) for two headings a, b in [-pi, pi], as the alignment torque of the reference
// evaluates it (sin(p.theta - theta), shared/particle.h:750 and :305-306): the difference is
// rounded once, and the sine of that double is taken to about one ulp *relative*, also where
// a - b is close to 0, +-pi or +-2 pi (aligned and anti-aligned neighbours).  The product form
// sin a cos b - cos a sin b of stored unit vectors cannot do this: its terms are ~1 and their
// roundings (~1e-16 absolute) swamp a small difference.
//   * reduction d = k pi + r, k = rint(d / pi) in {-2, ..., 2}, pi = PA + PB + PC with PA, PB of
//     33 significant bits (k PA, k PB exact); r is exact when it is small (Sterbenz), so the
//     result keeps the relative accuracy of fl(a - b) itself;
//   * sin r on |r| <= pi/2 by its Taylor series to r^21 (truncation below 1.3e-18), coefficients
//     the doubles nearest (-1)^m/(2m+1)!, Estrin's scheme with explicit fma;
//   * sin(k pi + r) = (-1)^k sin r: XOR_SIGN moves k's parity into the sign bit (one fewer compare
//     and select per pair), otherwise the sign is selected.  The two are bitwise equal; they differ
//     in the registers the compiler allocates around them, which decides the occupancy of some
//     stage kernels (the XOR form costs the continuous model's batched kernels a wave per SIMD).
template <bool XOR_SIGN = true>
SPP_HD inline double sin_diff(double a, double b)
{
    constexpr double kInvPi = 0.3183098861837907;
    constexpr double PA = 3.1415926534682512;      // leading 33 bits of pi
    constexpr double PB = 1.2154201012607932e-10;  // next 33 bits
    constexpr double PC = 4.044532497591901e-21;   // pi - PA - PB, rounded
    const double d = a - b;
    const double kf = rint(d * kInvPi);
    double r = fma(-kf, PA, d);
    r = fma(-kf, PB, r);
    r = fma(-kf, PC, r);
    const double z = r * r;
    const double z2 = z * z;
    const double z4 = z2 * z2;
    const double p01 = fma_k(z, 0.008333333333333333, -0.16666666666666666);
    const double p23 = fma_k(z, 2.7557319223985893e-06, -0.0001984126984126984);
    const double p45 = fma_k(z, 1.6059043836821613e-10, -2.505210838544172e-08);
    const double p67 = fma_k(z, 2.8114572543455206e-15, -7.647163731819816e-13);
    const double p89 = fma_k(z, 1.9572941063391263e-20, -8.22063524662433e-18);
    const double q0 = fma(p23, z2, p01);
    const double q1 = fma(p67, z2, p45);
    const double p = fma(fma(p89, z4, q1), z4, q0);
    const double s = fma(r * z, p, r);
    if (XOR_SIGN) return negate_if_odd(s, (int)kf);
    return ((int)kf & 1) ? -s : s;
}

// Natural logarithm of a positive normal double x, to about one ulp (the noise draw's log(1 - u),
// u a 53-bit uniform in [0, 1)).  The general-purpose library routine carries a double-double
// result through its whole evaluation (about 85 instructions on this GPU); this one is about 40:
//   * x = 2^k m with m in [sqrt(1/2), sqrt(2)) (frexp, then one doubling), f = m - 1 exact;
//   * log(1 + f) = 2 atanh(s) with s = f / (2 + f): since 2s = f - s f, log(1 + f) = f - s (f - R)
//     with R = 2 atanh(s)/s - 2 = z P(z), z = s^2 <= 0.0295, P the Taylor series of
//     sum_k 2 z^k / (2k + 3) to z^8 (truncation below 3e-17 of the result), by Estrin's scheme; the
//     leading f is exact, so the rounding of s only reaches the correction s (f - R);
//   * log x = k ln2_hi + (f - (s (f - R) - k ln2_lo)), ln2_hi of 40 significant bits (k ln2_hi
//     exact for |k| < 2^13).
// Host and device evaluate the same operations (IEEE division, explicit fma), so their results are
// bitwise equal; measured against extended-precision logl: at most one ulp on [2^-53, 1].
SPP_HD inline double log_pos(double x)
{
    constexpr double kSqrtHalf = 0.7071067811865476;
    constexpr double kLn2Hi = 0.6931471805601177;     // ln 2 to 40 significant bits
    constexpr double kLn2Lo = -1.7239444525614835e-13;  // ln 2 - kLn2Hi, rounded
#if defined(__HIP_DEVICE_COMPILE__)
    double m = __builtin_amdgcn_frexp_mant(x);
    int k = __builtin_amdgcn_frexp_exp(x);
#else
    int k;
    double m = std::frexp(x, &k);
#endif
    if (m < kSqrtHalf) {
        m = m + m;
        k -= 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    const double z2 = z * z;
    const double z4 = z2 * z2;
    const double z8 = z4 * z4;
    const double p01 = fma_k(z, 0.4, 0.6666666666666666);
    const double p23 = fma_k(z, 0.2222222222222222, 0.2857142857142857);
    const double p45 = fma_k(z, 0.15384615384615385, 0.18181818181818182);
    const double p67 = fma_k(z, 0.11764705882352941, 0.13333333333333333);
    const double p = fma(0.10526315789473684, z8, fma(fma(p67, z2, p45), z4, fma(p23, z2, p01)));
    const double R = z * p;
    const double kd = (double)k;
    const double lo = fma(s, f - R, -kd * kLn2Lo);
    return fma(kd, kLn2Hi, f - lo);
}

}  // namespace spp
