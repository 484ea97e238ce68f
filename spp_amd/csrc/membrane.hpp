// membrane.hpp -- shape and rotation observables of a ring membrane and the swimmers it encloses
// (host C++, no GPU).
//
// Follows Box::Compute_All_Variables (serial/beadbox.h:433-496) and Save_All_Variables (:498-506).
// Positions are the unwrapped ones (r_original); velocities are finite differences against the
// positions saved by the previous call, v = (r - r_old) / (t - t_old).  With <.> the mean over
// the membrane beads (mb) or over the swimmers (sw), R and V the membrane's centre of mass and
// its velocity:
//   mb_xx, mb_xy, mb_yy  gyration tensor <(r - R)(r - R)^T> of the beads
//   mb_l                 <(r - R) x (v - V)>   (2D cross product a.x b.y - a.y b.x)
//   mb_Rg = sqrt(mb_xx + mb_yy),  mb_omega = mb_l / (mb_xx + mb_yy)
//   mb_Delta = ((l1 - l2) / (l1 + l2))^2 with l1 >= l2 the gyration tensor's eigenvalues
//   sw_omega = <(r - R) x (v - V)>_sw / <|r - R|^2>_sw   (the swimmers' rotation about the
//              membrane's centre, in the membrane's frame)
// The reference also accumulates a swimmer spin that it never resets or writes; it is left out.
#pragma once

#include <cmath>
#include <ostream>

namespace spp_gpu {

struct MembraneVariables {
    double mb_r_cm[2] = {0, 0}, mb_v_cm[2] = {0, 0};
    double mb_xx = 0, mb_xy = 0, mb_yy = 0, mb_l = 0, mb_Rg = 0, mb_Delta = 0, mb_omega = 0;
    double sw_r_cm[2] = {0, 0}, sw_v_cm[2] = {0, 0};
    double sw_xx = 0, sw_xy = 0, sw_yy = 0, sw_l = 0, sw_Rg = 0, sw_Delta = 0, sw_omega = 0;
};

// Beads are entries [0, Nm), swimmers [Nm, Nm + Ns) of x/y (current) and xo/yo (previous);
// dt_elapsed = t - t_old.
inline MembraneVariables membrane_variables(const double* x, const double* y, const double* xo, const double* yo,
                                            int Nm, int Ns, double dt_elapsed)
{
    MembraneVariables m;
    auto moments = [&](int b, int n, double* rc, double* vc, double& xx, double& xy, double& yy, double& l) {
        for (int i = b; i < b + n; ++i) {
            double vx = (x[i] - xo[i]) / dt_elapsed, vy = (y[i] - yo[i]) / dt_elapsed;
            rc[0] += x[i];
            rc[1] += y[i];
            vc[0] += vx;
            vc[1] += vy;
            l += x[i] * vy - y[i] * vx;
            xx += x[i] * x[i];
            yy += y[i] * y[i];
            xy += x[i] * y[i];
        }
        for (int k = 0; k < 2; ++k) {
            rc[k] /= n;
            vc[k] /= n;
        }
        xx = xx / n - rc[0] * rc[0];
        yy = yy / n - rc[1] * rc[1];
        xy = xy / n - rc[0] * rc[1];
        l /= n;
    };
    auto shape = [](double xx, double xy, double yy, double& Rg, double& Delta) {
        double s = std::sqrt((xx - yy) * (xx - yy) + 4 * xy * xy);
        double l1 = 0.5 * (xx + yy) + 0.5 * s, l2 = 0.5 * (xx + yy) - 0.5 * s;
        Rg = std::sqrt(xx + yy);
        Delta = (l1 - l2) / (l1 + l2);
        Delta *= Delta;
    };
    const double* R = m.mb_r_cm;
    const double* V = m.mb_v_cm;
    moments(0, Nm, m.mb_r_cm, m.mb_v_cm, m.mb_xx, m.mb_xy, m.mb_yy, m.mb_l);
    m.mb_l -= R[0] * V[1] - R[1] * V[0];
    shape(m.mb_xx, m.mb_xy, m.mb_yy, m.mb_Rg, m.mb_Delta);
    m.mb_omega = m.mb_l / (m.mb_xx + m.mb_yy);
    if (Ns > 0) {
        const double* r = m.sw_r_cm;
        const double* v = m.sw_v_cm;
        moments(Nm, Ns, m.sw_r_cm, m.sw_v_cm, m.sw_xx, m.sw_xy, m.sw_yy, m.sw_l);
        // <r x v> - r_sw x V - R x v_sw + R x V  (beadbox.h:475)
        m.sw_l += -(r[0] * V[1] - r[1] * V[0]) - (R[0] * v[1] - R[1] * v[0]) + (R[0] * V[1] - R[1] * V[0]);
        shape(m.sw_xx, m.sw_xy, m.sw_yy, m.sw_Rg, m.sw_Delta);
        double dx = R[0] - r[0], dy = R[1] - r[1];
        m.sw_omega = m.sw_l / (m.sw_xx + m.sw_yy + dx * dx + dy * dy);
    }
    return m;
}

// The quantities file's header (written once, into an empty file) and one row, column order of
// Save_All_Variables (beadbox.h:502-504).
inline void write_variables_header(std::ostream& os)
{
    os << "#\ttime\tx_cm\ty_cm\tvx_cm\tvy_cm\tangular momentum\tangular freq.\tGyration raduis\tAsphericity"
          "\tQxx\tQxy\tQyy\tsw. ang. freq.\n";
}

inline void write_variables_row(std::ostream& os, double t, const MembraneVariables& m)
{
    os << t << "\t" << m.mb_r_cm[0] << "\t" << m.mb_r_cm[1] << "\t" << m.mb_v_cm[0] << "\t" << m.mb_v_cm[1] << "\t"
       << m.mb_l << "\t" << m.mb_omega << "\t" << m.mb_Rg << "\t" << m.mb_Delta << "\t" << m.mb_xx << "\t" << m.mb_xy
       << "\t" << m.mb_yy << "\t" << m.sw_omega << std::endl;
}

}  // namespace spp_gpu
