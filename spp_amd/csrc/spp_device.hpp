// spp_device.hpp -- device-side math shared by the engine's kernels (gfx950).
//
// Everything here is compiled with -ffp-contract=off so that products and sums round exactly as
// the reference's x86-64 -O3 build does (it has no FMA); the Verlet predicate sqrt(d2) < rv then
// sees bit-identical d2 on both sides and neighbour sets match exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spp_engine.h"

namespace spp {

constexpr double kPi = 3.14159265358979323846;  // the reference's PI = M_PI
constexpr double kTwoPi = 6.283185307179586;    // 2*M_PI as the reference evaluates it

// One particle's stage state, 32 B so a neighbour gather is one 128-B line access.
struct alignas(32) P4 {
    double x, y, th, w;
};

// Per-launch constants of the physics (kept in kernel arguments, i.e. SGPRs).
struct Phys {
    double Lx, Ly, Lx2, Ly2;
    double dt, half_dt, dt6;
    double v0;
    double c_align;    // REPULSIVE: g / (pi R_a^2); CONTINUOUS: (1 - alpha) / pi
    double c_self;     // CONTINUOUS: alpha / pi
    double g;          // CONTINUOUS: drift = g * summed torque
    double Ra, Ra2;    // alignment radius and its square (continuous predicate d2 < Ra^2)
    double rc, sigma, A;
    double rc13;       // (sigma/rc)^13 for R12 truncated
    double yuk_c;      // exp(-rc/sigma)(1/rc^2 + 1/(sigma rc)) for Yukawa truncated
    double guard2;     // d2 >= guard2 implies sqrt(d2) >= every cutoff (skip the sqrt)
    double noise_amp;  // sqrt(2 D dt)
    uint64_t seed;
};

// ------------------------------------------------------------------------------------------------
// Periodic_Transform, shared/c2dvector.h:67-73, one component: x -= L2*(int)floor(x/L2 + 0.5).
// Used for particle positions (once per particle per stage), where it must match bit for bit.
__device__ __forceinline__ double periodic_wrap(double x, double L2)
{
    return x - L2 * (double)(int)floor(x / L2 + 0.5);
}

// Minimum image of a coordinate difference of two wrapped positions (|dx| <= 2L).  Bit-identical
// to periodic_wrap(dx, 2L) except when dx lies within an ulp of +-L, i.e. for pairs farther apart
// than L, beyond every cutoff and list radius (the engine requires 2L/3 >= r_list).
__device__ __forceinline__ double min_image(double dx, double L, double L2)
{
    if (dx >= L)
        dx -= L2;
    else if (dx < -L)
        dx += L2;
    return dx;
}

// Theta wrap of the RK stages, shared/particle.h:599.
__device__ __forceinline__ double wrap_angle(double th)
{
    return th - floor(th / kTwoPi + 0.5) * 2 * kPi;
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
    return sqrt(-2.0 * log(1.0 - ua)) * cos(kTwoPi * ub);
}

// ------------------------------------------------------------------------------------------------
// Pair accumulation for particle i against neighbour j (full list: each particle sums its own
// side of every pair, which equals the reference's Newton-3 half-list update for that particle).
//   REPULSIVE  (particle.h:728-757): F = law(dr, d) for d < rc, tau += g sin(th_j - th_i)/(pi Ra^2)
//   CONTINUOUS (particle.h:296-338): for d2 < Ra^2, s1 += sin(th_j - th_i), s2 += (dr x v_i)/d
template <int MODEL, int REP>
__device__ __forceinline__ void pair_accumulate(const Phys& P, double xi, double yi, double thi,
                                                double vix, double viy, const P4& pj, double& fx,
                                                double& fy, double& s1, double& s2)
{
    double dx = min_image(xi - pj.x, P.Lx, P.Lx2);
    double dy = min_image(yi - pj.y, P.Ly, P.Ly2);
    double d2 = dx * dx + dy * dy;
    if (MODEL == SPP_MODEL_CONTINUOUS) {
        if (d2 < P.Ra2) {
            s1 += sin(pj.th - thi);
            if (P.c_self != 0.0) {
                double d = sqrt(d2);
                s2 += (dx * viy - dy * vix) / d;
            }
        }
        return;
    } else {
        if (d2 >= P.guard2) return;
        double d = sqrt(d2);
        if (REP != SPP_REP_NONE && d < P.rc) {
            double s;
            if (REP == SPP_REP_HARMONIC) {  // Spring, force-fields.h:87-97
                s = P.A * (P.rc - d) / d;
            } else if (REP == SPP_REP_R12_TRUNCATED || REP == SPP_REP_R12) {  // :14-58
                double sd = P.sigma / d, sd2 = sd * sd, sd4 = sd2 * sd2, sd8 = sd4 * sd4;
                double sd13 = sd8 * sd4 * sd;
                s = (REP == SPP_REP_R12 ? P.A * sd13 : P.A * (sd13 - P.rc13)) / d;
            } else {  // Yukawa (truncated), :60-85
                double e = exp(-d / P.sigma) * (1. / d2 + 1. / (P.sigma * d));
                s = (REP == SPP_REP_YUKAWA ? P.A * e : P.A * (e - P.yuk_c)) / d;
            }
            fx += s * dx;
            fy += s * dy;
        }
        if (d < P.Ra) s1 += sin(pj.th - thi);
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
