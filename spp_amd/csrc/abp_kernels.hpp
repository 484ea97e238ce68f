// abp_kernels.hpp -- single-bead active Brownian particles and a closed ring membrane (gfx950).
//
// Included by engine.hip after the Vicsek stage kernels.  The model restates ActiveBrownianChain
// with nb = 1 (shared/particle.h:792-1057) and the bead box's ring (parallel/beadbox.h:185-196,
// serial/beadbox.h:286-297):
//   * particles 0 .. n_membrane-1 (caller order) are ring beads with F0 = 0, the rest swimmers
//     with F0 = v0 (Set_F0; the swimming speed is mob_par * F0);
//   * Reset: torque = torque0, f = F0 * vhat (particle.h:878-883);
//   * every pair within rc feels one repulsion law (the nb = 1 branch applies a single law to
//     every pair, particle.h:968-981);
//   * ring bonds i <-> i+1 (mod n_membrane): Spring(dr, d, l_bond, k_bond) = k (l - d) dr / d
//     (force-fields.h:87-97);
//   * ring bending (no reference counterpart): E = kappa * sum_i (1 - cos phi_i), phi_i the angle
//     between the bonds that meet at bead i;
//   * Move: theta += mob_rot h torque + noise; r += mob_perp h f + (mob_par - mob_perp) h (f.v) v
//     (Move / Move_Runge_Kutta2_1/2, particle.h:898-962; stage 1 projects on the heading before
//     the update, stage 2 on the midpoint heading, Euler on the new heading).
#pragma once

// d(cos phi)/da and d(cos phi)/db for the angle between bond vectors a (into the bead) and b
// (out of it): cos phi = a.b / (|a||b|).
__device__ __forceinline__ void bend_grad(double ax, double ay, double bx, double by, double& gax,
                                          double& gay, double& gbx, double& gby)
{
    double aa = ax * ax + ay * ay, bb = bx * bx + by * by, ab = ax * bx + ay * by;
    double inv = 1.0 / sqrt(aa * bb);
    double c = ab * inv;
    gax = bx * inv - c * ax / aa;
    gay = by * inv - c * ay / aa;
    gbx = ax * inv - c * bx / bb;
    gby = ay * inv - c * by / bb;
}

// Bond and bending forces on ring bead `pid`, read from the stage input through the inverse
// permutation (ring neighbours are caller indices pid +- 1, pid +- 2).
__device__ __forceinline__ void ring_forces(const StageArgs& a, const Phys& P, int pid, const P4& me,
                                            double& fx, double& fy)
{
    const AbpConst& A = a.abp;
    const int Nm = A.n_membrane;
    const P4 qm1 = stage_record(a, a.inv[(pid + Nm - 1) % Nm]);
    const P4 qp1 = stage_record(a, a.inv[(pid + 1) % Nm]);
    // bond vectors along the ring: t1 = r_k - r_{k-1}, t2 = r_{k+1} - r_k
    double t1x = spp::min_image(me.x - qm1.x, P.Lx2, P.invLx2);
    double t1y = spp::min_image(me.y - qm1.y, P.Ly2, P.invLy2);
    double t2x = spp::min_image(qp1.x - me.x, P.Lx2, P.invLx2);
    double t2y = spp::min_image(qp1.y - me.y, P.Ly2, P.invLy2);
    // Spring on bond (k, k+1) with dr = r_k - r_{k+1} = -t2, and on bond (k-1, k) with
    // dr = r_{k-1} - r_k = -t1, which acts on bead k with the opposite sign
    double d2 = sqrt(t2x * t2x + t2y * t2y), d1 = sqrt(t1x * t1x + t1y * t1y);
    fx += A.k_bond * (A.l_bond - d2) * (-t2x) / d2;
    fy += A.k_bond * (A.l_bond - d2) * (-t2y) / d2;
    fx -= A.k_bond * (A.l_bond - d1) * (-t1x) / d1;
    fy -= A.k_bond * (A.l_bond - d1) * (-t1y) / d1;
    if (A.kappa != 0.0) {
        const P4 qm2 = stage_record(a, a.inv[(pid + Nm - 2) % Nm]);
        const P4 qp2 = stage_record(a, a.inv[(pid + 2) % Nm]);
        double t0x = spp::min_image(qm1.x - qm2.x, P.Lx2, P.invLx2);
        double t0y = spp::min_image(qm1.y - qm2.y, P.Ly2, P.invLy2);
        double t3x = spp::min_image(qp2.x - qp1.x, P.Lx2, P.invLx2);
        double t3y = spp::min_image(qp2.y - qp1.y, P.Ly2, P.invLy2);
        double ax, ay, bx, by;
        // F_k = kappa [dc_k/da - dc_k/db - dc_{k+1}/da + dc_{k-1}/db]
        double gx = 0, gy = 0;
        bend_grad(t0x, t0y, t1x, t1y, ax, ay, bx, by);  // angle at k-1
        gx += bx;
        gy += by;
        bend_grad(t1x, t1y, t2x, t2y, ax, ay, bx, by);  // angle at k
        gx += ax - bx;
        gy += ay - by;
        bend_grad(t2x, t2y, t3x, t3y, ax, ay, bx, by);  // angle at k+1
        gx -= ax;
        gy -= ay;
        fx += A.kappa * gx;
        fy += A.kappa * gy;
    }
}

// Bead pairs of two chains (ActiveBrownianChain::Interact, particle.h:983-1011), this particle's
// side: beads at s = v sep ((1 - nb)/2 + k), dr = (dr0 + s_this) - s_that, R12_Repulsive_Truncated
// within rc_b (force-fields.h:14-39); adds F to f and s_this x F to the torque.
__device__ __forceinline__ void bead_pairs(const AbpConst& A, double dx0, double dy0, double mc, double ms, int na,
                                           double cj, double sj, int nbj, double& fx, double& fy, double& tq)
{
    for (int k = 0; k < na; ++k) {
        const double ok = A.sep * ((1 - na) / 2.0 + k);
        const double stx = mc * ok, sty = ms * ok;
        for (int l = 0; l < nbj; ++l) {
            const double ol = A.sep * ((1 - nbj) / 2.0 + l);
            const double dx = dx0 + stx - cj * ol, dy = dy0 + sty - sj * ol;
            const double d2 = dx * dx + dy * dy;
            if (d2 < A.T_b) {
                const double d = sqrt(d2);
                const double q = A.sigma_b / d, q2 = q * q, q4 = q2 * q2, q8 = q4 * q4;
                const double st = A.A_b * (q8 * q4 * q - A.rc13_b);
                const double Fx = st * dx / d, Fy = st * dy / d;
                fx += Fx;
                fy += Fy;
                tq += stx * Fy - sty * Fx;
            }
        }
    }
}

// Pairs of a particle with nb beads (1 for ring beads) when the swimmers are chains: two single
// beads feel the configured law (the reference's nb = 1 branch), every other pair the bead law.
// The neighbour's class comes from its caller index, its heading from its record.
template <int REP>
__device__ __forceinline__ void chain_gather(const StageArgs& a, const Phys& P, int i, const P4& me, double mc,
                                             double ms, int na, double& fx, double& fy, double& tq)
{
    const AbpConst& A = a.abp;
    const int cnt = a.ncount[i];
    double s1 = 0, s2 = 0;
    for (int k = 0; k < cnt; ++k) {
        const int j = a.nbr32[(size_t)k * a.N + i];
        const P4 q = stage_record(a, j);
        const int nbj = (unsigned)a.id[j] < (unsigned)A.n_membrane ? 1 : A.nb;
        if (na == 1 && nbj == 1) {
            spp::pair_accumulate<SPP_MODEL_REPULSIVE, REP>(P, me, mc, ms, q, fx, fy, s1, s2);
        } else {
            const double dx0 = spp::min_image(me.x - q.x, P.Lx2, P.invLx2);
            const double dy0 = spp::min_image(me.y - q.y, P.Ly2, P.invLy2);
            double cj, sj;
            spp::heading_unit(q.z, cj, sj);
            bead_pairs(A, dx0, dy0, mc, ms, na, cj, sj, nbj, fx, fy, tq);
        }
    }
}

// Forces and torque on one particle: Reset (f = F0 v, torque = torque0) + pair repulsion + ring
// forces.
template <int REP>
__device__ __forceinline__ void abp_forces(const StageArgs& a, const Phys& P0, int i, int pid, const P4& me,
                                           double mc, double ms, double& fx, double& fy, double& tq)
{
    Phys P = P0;
    const bool bead = (unsigned)pid < (unsigned)a.abp.n_membrane;
    P.v0 = bead ? 0.0 : a.abp.F0;
    tq = a.abp.torque0;
    if (a.abp.nb <= 1) {
        double t_unused;
        gather_forces_flat<SPP_MODEL_REPULSIVE, REP>(a, P, i, me, mc, ms, fx, fy, t_unused);
    } else {
        fx = mc * P.v0;
        fy = ms * P.v0;
        chain_gather<REP>(a, P, i, me, mc, ms, bead ? 1 : a.abp.nb, fx, fy, tq);
    }
    if (bead) ring_forces(a, P, pid, me, fx, fy);
}

// One pair's contribution to particle si's force and torque (caller's Particle::Interact): the
// class rules of chain_gather for the single neighbour sj.
template <int REP>
__device__ __forceinline__ void abp_pair_side(const StageArgs& a, const Phys& P, int si, int sj, double& fx,
                                              double& fy, double& tq)
{
    const AbpConst& A = a.abp;
    const P4 me = stage_record(a, si), q = stage_record(a, sj);
    double mc, ms;
    spp::heading_unit(me.z, mc, ms);
    const int na = A.nb > 1 && !((unsigned)a.id[si] < (unsigned)A.n_membrane) ? A.nb : 1;
    const int nbj = A.nb > 1 && !((unsigned)a.id[sj] < (unsigned)A.n_membrane) ? A.nb : 1;
    if (na == 1 && nbj == 1) {
        double s1 = 0, s2 = 0;
        spp::pair_accumulate<SPP_MODEL_REPULSIVE, REP>(P, me, mc, ms, q, fx, fy, s1, s2);
    } else {
        const double dx0 = spp::min_image(me.x - q.x, P.Lx2, P.invLx2);
        const double dy0 = spp::min_image(me.y - q.y, P.Ly2, P.invLy2);
        double cj, sj_;
        spp::heading_unit(q.z, cj, sj_);
        bead_pairs(A, dx0, dy0, mc, ms, na, cj, sj_, nbj, fx, fy, tq);
    }
}

// Mobilities of a particle: (parallel, perpendicular, rotational).
__device__ __forceinline__ void abp_mobility(const AbpConst& A, int pid, double& mpar, double& mperp, double& krot)
{
    const bool chain = A.nb > 1 && !((unsigned)pid < (unsigned)A.n_membrane);
    mpar = chain ? A.c_par : A.mob_par;
    mperp = chain ? A.c_perp : A.mob_perp;
    krot = chain ? A.c_rot : A.mob_rot;
}

template <int REP, int INTEG, int STAGE>
__global__ __launch_bounds__(kBlock) void k_stage_abp(StageArgs a)
{
    int b = xcd_block();
    int i = b * kBlock + threadIdx.x;
    bool live = i < a.N;
    const Phys& P = a.P;
    const AbpConst& A = a.abp;
    constexpr bool final_stage = INTEG == SPP_EULER || STAGE == 2;
    double x = 0, y = 0, dmax = 0;
    if (live) {
        P4 me = stage_record(a, i);
        double mc, ms;  // heading of the stage input (stage 1: the old one; stage 2: the midpoint)
        spp::heading_unit(me.z, mc, ms);
        int pid = a.id[i];
        double dn, th_in = 0, thold = 0;
        double2 u = make_double2(0, 0);
        if (STAGE <= 1) {
            if (a.id_out) a.id_out[i] = pid;
            th_in = me.z;
            if (INTEG == SPP_EULER) u = a.U_in[i];
            dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)pid, noise_step(a));
        } else {
            thold = a.th0[i];
            dn = a.dth[i];
            u = a.U_in[i];
        }
        double fx, fy, tq;
        abp_forces<REP>(a, P, i, pid, me, mc, ms, fx, fy, tq);
        double mpar, mperp, krot;
        abp_mobility(A, pid, mpar, mperp, krot);
        const double dm = mpar - mperp;
        double th, sn, cs;
        if (INTEG == SPP_EULER) {
            th = th_in + krot * P.dt * (tq);
            th += dn;
            spp::sincos_heading(th, &sn, &cs);
            u.x += fx * (P.dt * mperp);
            u.y += fy * (P.dt * mperp);
            double fv = fx * cs + fy * sn;
            u.x += cs * (fv * (P.dt * dm));
            u.y += sn * (fv * (P.dt * dm));
            a.U_out[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
        } else if (STAGE == 1) {
            a.dth[i] = dn;
            th = th_in + krot * P.half_dt * (tq);
            th += dn / 2;
            th = spp::wrap_angle(th);
            double xx = me.x + fx * (P.half_dt * mperp);
            double yy = me.y + fy * (P.half_dt * mperp);
            double fv = fx * mc + fy * ms;
            xx += mc * (fv * (P.half_dt * dm));
            yy += ms * (fv * (P.half_dt * dm));
            x = spp::periodic_wrap(xx, P.Lx, P.Lx2);
            y = spp::periodic_wrap(yy, P.Ly, P.Ly2);
        } else {
            th = thold + krot * P.dt * (tq);
            th += dn;
            th = spp::wrap_angle(th);
            u.x = u.x + fx * (P.dt * mperp);
            u.y = u.y + fy * (P.dt * mperp);
            double fv = fx * mc + fy * ms;
            u.x += mc * (fv * (P.dt * dm));
            u.y += ms * (fv * (P.dt * dm));
            a.U_out[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
        }
        a.oxy[i] = make_double2(x, y);
        a.oth[i] = th;
    }
    check_disp(a, P, b, live, i, x, y, final_stage ? a.kstep + 1 : a.kstep, dmax);
    finish_disp(a, b, dmax);
}

template <int REP>
__global__ __launch_bounds__(kBlock) void k_interact_abp(StageArgs a, P4* force)
{
    int i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.N) return;
    P4 me = stage_record(a, i);
    double mc, ms, fx, fy, tq;
    spp::heading_unit(me.z, mc, ms);
    abp_forces<REP>(a, a.P, i, a.id[i], me, mc, ms, fx, fy, tq);
    force[i] = P4{fx, fy, tq, 0.0};
}

// Split-call move for the single-bead model (forces from k_interact_abp).
__global__ __launch_bounds__(kBlock) void k_move_abp(MoveArgs a)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const Phys& P = a.P;
    const AbpConst& A = a.abp;
    const double2 me = a.cur.xy[i];
    double th_in = a.cur.th[i];
    double mc, ms;
    spp::heading_unit(th_in, mc, ms);
    P4 f = a.force[i];
    double2 u = a.cur.U[i];
    double mpar, mperp, krot;
    abp_mobility(A, a.cur.id[i], mpar, mperp, krot);
    const double dm = mpar - mperp;
    double th, x, y, sn, cs;
    if (a.stage <= 1) {
        double dn = a.inj ? a.inj[i] : P.noise_amp * spp::noise_normal(P.seed, (uint32_t)a.cur.id[i], a.step);
        a.dth[i] = dn;
        if (a.integ == SPP_EULER) {
            th = th_in + krot * P.dt * (f.z);
            th += dn;
            spp::sincos_heading(th, &sn, &cs);
            u.x += f.x * (P.dt * mperp);
            u.y += f.y * (P.dt * mperp);
            double fv = f.x * cs + f.y * sn;
            u.x += cs * (fv * (P.dt * dm));
            u.y += sn * (fv * (P.dt * dm));
            a.cur.U[i] = u;
            x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
            y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
        } else {
            a.old[i] = P4{u.x, u.y, th_in, 0.0};
            th = th_in + krot * P.half_dt * (f.z);
            th += dn / 2;
            th = spp::wrap_angle(th);
            double xx = me.x + f.x * (P.half_dt * mperp);
            double yy = me.y + f.y * (P.half_dt * mperp);
            double fv = f.x * mc + f.y * ms;
            xx += mc * (fv * (P.half_dt * dm));
            yy += ms * (fv * (P.half_dt * dm));
            x = spp::periodic_wrap(xx, P.Lx, P.Lx2);
            y = spp::periodic_wrap(yy, P.Ly, P.Ly2);
        }
    } else {
        P4 o = a.old[i];
        th = o.z + krot * P.dt * (f.z);
        th += a.dth[i];
        th = spp::wrap_angle(th);
        u.x = o.x + f.x * (P.dt * mperp);
        u.y = o.y + f.y * (P.dt * mperp);
        double fv = f.x * mc + f.y * ms;
        u.x += mc * (fv * (P.dt * dm));
        u.y += ms * (fv * (P.dt * dm));
        a.cur.U[i] = u;
        x = spp::periodic_wrap(u.x, P.Lx, P.Lx2);
        y = spp::periodic_wrap(u.y, P.Ly, P.Ly2);
    }
    a.cur.xy[i] = make_double2(x, y);
    a.cur.th[i] = th;
}
