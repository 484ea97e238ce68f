// Drives spp_gpu::membrane_variables (spp_amd/csrc/membrane.hpp) from text input for the CPU tests.
// stdin: "Nm Ns dt" then Nm + Ns lines "x y x_old y_old"; stdout: "name value" lines, then the
// quantities-file header and one row at t = 1.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <vector>

#include "../../spp_amd/csrc/membrane.hpp"

int main()
{
    int Nm = 0, Ns = 0;
    double dt = 0;
    if (!(std::cin >> Nm >> Ns >> dt)) return 2;
    int n = Nm + Ns;
    std::vector<double> x(n), y(n), xo(n), yo(n);
    for (int i = 0; i < n; ++i)
        if (!(std::cin >> x[i] >> y[i] >> xo[i] >> yo[i])) return 2;
    spp_gpu::MembraneVariables m = spp_gpu::membrane_variables(x.data(), y.data(), xo.data(), yo.data(), Nm, Ns, dt);
    auto put = [](const char* k, double v) { std::printf("%s %.17g\n", k, v); };
    put("mb_x_cm", m.mb_r_cm[0]);
    put("mb_y_cm", m.mb_r_cm[1]);
    put("mb_vx_cm", m.mb_v_cm[0]);
    put("mb_vy_cm", m.mb_v_cm[1]);
    put("mb_xx", m.mb_xx);
    put("mb_xy", m.mb_xy);
    put("mb_yy", m.mb_yy);
    put("mb_l", m.mb_l);
    put("mb_Rg", m.mb_Rg);
    put("mb_Delta", m.mb_Delta);
    put("mb_omega", m.mb_omega);
    put("sw_x_cm", m.sw_r_cm[0]);
    put("sw_y_cm", m.sw_r_cm[1]);
    put("sw_l", m.sw_l);
    put("sw_omega", m.sw_omega);
    std::ostringstream os;
    spp_gpu::write_variables_header(os);
    spp_gpu::write_variables_row(os, 1.0, m);
    std::fputs(os.str().c_str(), stdout);
    return 0;
}
