// spp_device.hpp -- device-side math shared by the engine's kernels (gfx950).
//
// Compiled with -ffp-contract=off: products and sums round exactly as the reference's x86-64 -O3
// build does (no FMA), so squared distances are bit-identical on both sides; fma() appears only
// where the result is provably exact or where bitwise agreement is not needed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spp_engine.h"
#include "trig.hpp"

namespace spp {

constexpr double kPi = 3.14159265358979323846;  // the reference's PI = M_PI
constexpr double kTwoPi = 6.283185307179586;    // 2*M_PI as the reference evaluates it

// One particle as its neighbours read it: wrapped position (reference r) and heading angle
// (reference theta, wrapped into [-pi, pi)), 32 B so a gather is one aligned quarter of a 128-B
// line.  Neighbours need theta itself: the alignment torque is sin(theta_j - theta_i) evaluated
// from the angles (sin_diff), which keeps its relative accuracy for nearly aligned headings.  The
// fourth field carries theta_old (the step's starting heading) in the outputs of the middle RK
// stages, so the later stages of the step read it with the record; it is 0 in step-final records.
// The same 32-B type also holds generic 4-vectors (forces and torque, RK4 accumulators).
struct alignas(32) P4 {
    double x, y, z, w;
};

// Streaming (nontemporal) stores for a stage's per-particle outputs.  Nothing in the same launch
// reads them back; the next launch reads them after the kernel boundary.  Marked as streaming,
// they leave the L2 during the launch instead of waiting for the write-back at its end.  At
// 262,144 particles this makes the RK2 stage kernels about 10 % faster; large engines keep
// ordinary stores (the engine's kStreamMaxN).  A stage wave writes
// whole lines (64 consecutive slots); the rebuild's scattered and partial-line writes (the
// permutation, the slot-major list) became 1.4-1.8x slower with the same hint, so they keep
// ordinary stores.
typedef double stream_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_stream(P4* p, double x, double y, double z, double w)
{
    stream_d2* q = reinterpret_cast<stream_d2*>(p);
    __builtin_nontemporal_store(stream_d2{x, y}, q);
    __builtin_nontemporal_store(stream_d2{z, w}, q + 1);
}
__device__ __forceinline__ void store_stream(double2* p, const double2& v)
{
    __builtin_nontemporal_store(stream_d2{v.x, v.y}, reinterpret_cast<stream_d2*>(p));
}
__device__ __forceinline__ void store_stream(double* p, double v) { __builtin_nontemporal_store(v, p); }
// a streaming or an ordinary store, chosen per launch (a uniform branch)
__device__ __forceinline__ void store_out(bool stream, P4* p, double x, double y, double z, double w)
{
    if (stream)
        store_stream(p, x, y, z, w);
    else
        *p = P4{x, y, z, w};
}
template <class T>
__device__ __forceinline__ void store_out(bool stream, T* p, const T& v)
{
    if (stream)
        store_stream(p, v);
    else
        *p = v;
}

// Per-launch physics constants (kernel arguments, i.e. SGPRs).
struct Phys {
    double Lx, Ly, Lx2, Ly2, invLx2, invLy2;
    double dt, half_dt, dt6;
    double v0;
    double c_align;   // REPULSIVE: g / (pi R_a^2); CONTINUOUS: (1 - alpha) / pi
    double c_self;    // CONTINUOUS: alpha / pi
    double g;         // CONTINUOUS: drift = g * summed torque
    double Ra2;       // CONTINUOUS predicate d2 < R_a^2 (particle.h:305)
    // Exact squared-distance thresholds: sqrt(d2) < R  <=>  d2 < T(R) for the correctly rounded
    // sqrt the reference uses (cell.h:84-85, particle.h:738,747), so no sqrt is needed to decide.
    double T_align, T_rep, T_max;
    double rc, sigma, A;
    double rc13;      // (sigma/rc)^13 for R12 truncated
    double yuk_c;     // exp(-rc/sigma)(1/rc^2 + 1/(sigma rc)) for Yukawa truncated
    double inv_sigma;
    double noise_amp; // sqrt(2 D dt)
    uint64_t seed;
    // A particle at most L - img_m from the centre along both axes has every listed neighbour on the
    // same side of the periodic edges (displacement-triggered lists: neighbours are closer than
    // r_list + the skin), so its pair differences need no minimum image.  Infinite: never assumed.
    double img_m;
};

// ------------------------------------------------------------------------------------------------
// Periodic_Transform, shared/c2dvector.h:67-73, one component: x -= L2*(int)floor(x/L2 + 0.5).
// Positions are wrapped ones plus one stage's displacement, so k = floor(x/L2 + 0.5) is -1, 0 or
// 1 and is decided by comparing x with +-L; only within 1e-9 L of the edges (where rounding of
// x/L2 + 0.5 could decide) or far outside is the reference's division evaluated.  Either way the
// result is bit-identical to the reference formula.
__device__ __forceinline__ double periodic_wrap(double x, double L, double L2)
{
    double ax = fabs(x), k;
    if (fabs(ax - L) < 1e-9 * L || ax >= 2.9 * L)
        k = (double)(int)floor(x / L2 + 0.5);
    else
        k = ax < L ? 0.0 : (x > 0 ? 1.0 : -1.0);
    return x - L2 * k;
}

// Minimum image of a difference of two wrapped coordinates (|dx| <= 2L).  k = rint(dx / 2L) is
// in {-1, 0, 1}, so fma(-k, 2L, dx) rounds once, exactly like the reference's dx - 2L*k; k agrees
// with floor(dx/2L + 0.5) except within an ulp of |dx| = L, i.e. for pairs farther apart than L,
// beyond every cutoff and list radius (the engine requires 2L/3 >= r_list).
__device__ __forceinline__ double min_image(double dx, double L2, double invL2)
{
    double k = rint(dx * invL2);
    return fma(-k, L2, dx);
}

// The same difference when the caller knows that k = 0 (both coordinates away from the periodic
// edges, closer than L to each other): fma(-0, 2L, dx) is dx, so skipping the image is exact.
template <bool IMG>
__device__ __forceinline__ double image_diff(double dx, double L2, double invL2)
{
    return IMG ? min_image(dx, L2, invL2) : dx;
}

// 1/sqrt(x) for a positive finite x: the hardware estimate and the correction the device library
// applies to it (e = 1 - x y^2, y + y e (1/2 + 3/8 e)), without the library's guards for zero,
// infinity and NaN, which pair distances never are; the same value as the library's rsqrt there.
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(y * -x, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Theta wrap of the RK stages, shared/particle.h:599: th - floor(th/(2 pi) + 0.5)*2*pi, with the
// same comparison shortcut as periodic_wrap (bit-identical result).
__device__ __forceinline__ double wrap_angle(double th)
{
    double a = fabs(th), k;
    if (fabs(a - kPi) < 1e-9 || a >= 2.9 * kPi)
        k = floor(th / kTwoPi + 0.5);
    else
        k = a < kPi ? 0.0 : (th > 0 ? 1.0 : -1.0);
    return th - k * 2 * kPi;
}

// Heading unit vector v = (cos theta, sin theta) of a record's angle (the reference's v, set from
// theta at every Move).  Angles a step wrote are in [-pi, pi); angles uploaded by a caller may be
// anything, and those beyond the heading routine's range take the library's reduction.
__device__ __forceinline__ void heading_unit(double th, double& c, double& s)
{
    if (fabs(th) <= 524288.0)
        sincos_heading(th, &s, &c);
    else
        sincos(th, &s, &c);
}

// ------------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and the angular-noise stream: counter
// (id, step_lo, step_hi, 'SPPn'), key (seed_lo, seed_hi), Box-Muller on two 53-bit uniforms.
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
    }
}

__host__ __device__ __forceinline__ double noise_normal(uint64_t seed, uint32_t id, int64_t step)
{
    uint32_t c[4] = {id, (uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), 0x5350506Eu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double ua = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * (1.0 / 9007199254740992.0);
    double ub = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) * (1.0 / 9007199254740992.0);
    double sn, cs;
    sincos_heading(kTwoPi * ub, &sn, &cs);  // cos(2 pi u_b), the heading routine (|arg| < 2 pi)
    return sqrt(-2.0 * log_pos(1.0 - ua)) * cs;  // 1 - ua in [2^-53, 1]: positive, normal
}

// ------------------------------------------------------------------------------------------------
// Pair accumulation of particle i against neighbour j (full list: each particle sums its own side
// of every pair, which equals what the reference's Newton-3 half-list update adds to it).
//   REPULSIVE  (particle.h:728-757): F = law(dr, d) for d < rc; s1 += sin(th_j - th_i) for d < Ra
//   CONTINUOUS (particle.h:296-338): for d2 < Ra^2, s1 += sin(th_j - th_i), s2 += (dr x v_i)/d
// me holds i's record (x, y, theta); (mc, ms) = (cos, sin) of theta_i, needed by the CONTINUOUS
// self-term only.  sin(th_j - th_i) is taken from the angles (sin_diff), like particle.h:750.
// IMG = false: the caller guarantees that no pair crosses a periodic edge (image_diff).
// ONE = true (REPULSIVE with a repulsion law): the caller guarantees T_rep == T_align == T_max, so a
// pair inside T_max feels both laws and the two inner tests (each a divergent branch) are skipped;
// the arithmetic and its order are those of the general path, so the sums are bitwise equal.

// The repulsion law's force on i, added to (fx, fy), for a pair at squared distance d2 > 0.
template <int REP>
__device__ __forceinline__ void repulsion_add(const Phys& P, double d2, double dx, double dy, double& fx, double& fy)
{
    double ri = rsqrt_pos(d2);  // 1/d
    double s;
    if (REP == SPP_REP_HARMONIC) {  // Spring, force-fields.h:87-97: k (rc - d)/d
        s = P.A * (P.rc * ri - 1.0);
    } else if (REP == SPP_REP_R12_TRUNCATED || REP == SPP_REP_R12) {  // :14-58
        double q = P.sigma * ri, q2 = q * q, q4 = q2 * q2, q8 = q4 * q4;
        double q13 = q8 * q4 * q;
        s = (REP == SPP_REP_R12 ? P.A * q13 : P.A * (q13 - P.rc13)) * ri;
    } else {  // Yukawa (truncated), :60-85
        double d = d2 * ri;
        double e = exp(-d * P.inv_sigma) * (ri * ri + ri * P.inv_sigma);
        s = (REP == SPP_REP_YUKAWA ? P.A * e : P.A * (e - P.yuk_c)) * ri;
    }
    fx += s * dx;
    fy += s * dy;
}

template <int MODEL, int REP, bool IMG = true, bool ONE = false>
__device__ __forceinline__ void pair_accumulate(const Phys& P, const P4& me, double mc, double ms, const P4& pj,
                                                double& fx, double& fy, double& s1, double& s2)
{
    double dx = image_diff<IMG>(me.x - pj.x, P.Lx2, P.invLx2);
    double dy = image_diff<IMG>(me.y - pj.y, P.Ly2, P.invLy2);
    double d2 = dx * dx + dy * dy;
    if (MODEL == SPP_MODEL_CONTINUOUS) {
        if (d2 < P.Ra2) {
            s1 += sin_diff<false>(pj.z, me.z);
            if (P.c_self != 0.0) s2 += (dx * ms - dy * mc) * rsqrt(d2);
        }
        return;
    } else {
        if (d2 >= P.T_max) return;
        if (ONE && REP != SPP_REP_NONE) {
            repulsion_add<REP>(P, d2, dx, dy, fx, fy);
            s1 += sin_diff(pj.z, me.z);
            return;
        }
        if (REP != SPP_REP_NONE && d2 < P.T_rep) repulsion_add<REP>(P, d2, dx, dy, fx, fy);
        if (d2 < P.T_align) s1 += sin_diff(pj.z, me.z);
    }
}

// Drift torque of particle i from its accumulated sums.
template <int MODEL>
__device__ __forceinline__ double drift_torque(const Phys& P, double s1, double s2)
{
    if (MODEL == SPP_MODEL_CONTINUOUS) return P.g * (P.c_align * s1 - P.c_self * s2);
    return P.c_align * s1;
}

}  // namespace spp
