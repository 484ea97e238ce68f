// vesicle_demo.cpp -- the reference's vesicle data-gathering loop (serial/membrane.cpp:24-64) on the
// GPU engine through the C++ Box facade: a ring membrane of Nm beads around Ns active Brownian
// swimmers; Multi_Step(period) chunks, a binary frame every saving_period chunks and a row of the
// membrane observables every quantities_period chunks, with resume from an existing trajectory
// (membrane.cpp:163-185).
//
// usage: vesicle_demo OUT.bin QUANT.dat Nm Ns F0 steps period saving_period quantities_period [resume]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "box.hpp"

int main(int argc, char** argv)
{
    if (argc < 10) {
        std::fprintf(stderr,
                     "usage: %s OUT.bin QUANT.dat Nm Ns F0 steps period saving_period quantities_period [resume]\n",
                     argv[0]);
        return 2;
    }
    const char* out_name = argv[1];
    const char* quant_name = argv[2];
    int Nm = std::atoi(argv[3]);
    int Ns = std::atoi(argv[4]);
    double F0 = std::atof(argv[5]);
    int steps = std::atoi(argv[6]);
    int period = std::atoi(argv[7]);
    int saving = std::atoi(argv[8]);
    int quantities = std::atoi(argv[9]);
    bool resume = argc > 10 && std::atoi(argv[10]) != 0;
    if (Nm < 3 || Ns < 0 || period <= 0 || saving <= 0 || quantities <= 0) {
        std::fprintf(stderr, "vesicle_demo: need Nm >= 3, Ns >= 0 and positive periods\n");
        return 2;
    }

    // touching beads of diameter 1 on the ring; the box leaves room for the ring to move
    double R = 0.5 / std::sin(M_PI / Nm);
    spp_params p;
    spp_params_default(&p, SPP_MODEL_ABP);
    p.n = Nm + Ns;
    p.n_membrane = Nm;
    p.Lx = p.Ly = R + 6;
    p.integrator = SPP_RK2;
    p.dt = 1e-3;
    p.repulsion = SPP_REP_HARMONIC;
    p.A_rep = 10.0;
    p.sigma_rep = 1.0;
    p.rc_rep = 1.0;
    p.r_list = 1.3;
    p.D_rot = 0.125;  // Set_Dr(1.0/8), membrane.cpp:121
    p.v0 = F0;
    p.mob_par = 0.5;
    p.mob_perp = 0.5;
    p.mob_rot = 1.0;
    p.k_bond = 100.0;
    p.l_bond = 1.0;
    p.kappa_bend = 10.0;
    p.seed = 5;
    try {
        spp_gpu::Box box(p);
        std::ofstream out, quant;
        if (resume && box.Positioning_Particles(out_name)) {
            out.open(out_name, std::ios::out | std::ios::app | std::ios::binary);
            quant.open(quant_name, std::ios::out | std::ios::app);
        } else {
            std::vector<double> x(Nm + Ns), y(Nm + Ns), th(Nm + Ns, 0.0);
            if (spp_ring_membrane(Nm, 1.0, R, x.data(), y.data()) != SPP_OK ||
                (Ns > 0 && spp_disk_state(11, Ns, R - 1.0, x.data() + Nm, y.data() + Nm, th.data() + Nm) != SPP_OK))
                throw spp_gpu::Error(SPP_ERR_ARG, "initial state");
            box.Init(x, y, th);
            out.open(out_name, std::ios::out | std::ios::binary);
            quant.open(quant_name, std::ios::out);
            box.Save_Particles_Positions();
            out << &box;
        }
        box.Save_Particles_Positions();
        for (int s = 0; s < steps; s += period) {
            box.Multi_Step(period);
            int chunk = s / period + 1;
            if (chunk % saving == 0) out << &box;
            if (chunk % quantities == 0) box.Save_All_Variables(quant);
        }
        std::printf("vesicle_demo: t=%.6f Rg=%.6f asphericity=%.6f\n", box.t, box.variables.mb_Rg,
                    box.variables.mb_Delta);
    } catch (const spp_gpu::Error& err) {
        std::fprintf(stderr, "vesicle_demo: error %d: %s\n", err.code, err.what());
        return 1;
    }
    return 0;
}
