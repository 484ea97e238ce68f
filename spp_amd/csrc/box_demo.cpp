// box_demo.cpp -- the reference's data-gathering loop (parallel/repulsive-vicsek.cpp:52-105) on the
// GPU engine through the C++ Box facade: Multi_Step(cell_update_period) chunks, the polarization
// time series and a binary trajectory frame every saving_period chunks, with resume from an
// existing trajectory (Positioning_Particles, :155-180); at the end the trajectory is read back
// frame by frame (operator>>).
//
// usage: box_demo OUT.bin POLAR.dat N L steps period saving_period [resume]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

#include "box.hpp"

int main(int argc, char** argv)
{
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s OUT.bin POLAR.dat N L steps period saving_period [resume]\n", argv[0]);
        return 2;
    }
    const char* out_name = argv[1];
    const char* polar_name = argv[2];
    int N = std::atoi(argv[3]);
    double L = std::atof(argv[4]);
    int steps = std::atoi(argv[5]);
    int period = std::atoi(argv[6]);
    int saving = std::atoi(argv[7]);
    bool resume = argc > 8 && std::atoi(argv[8]) != 0;

    spp_params p;
    spp_params_default(&p, SPP_MODEL_REPULSIVE);
    p.n = N;
    p.Lx = p.Ly = L;
    p.dt = 0.005;
    p.v0 = 1.0;
    p.R_align = 1.0;
    p.repulsion = SPP_REP_HARMONIC;
    p.A_rep = 10.0;
    p.rc_rep = 1.0;
    p.D_rot = 0.05;
    p.r_list = 1.3;
    p.seed = 2;
    try {
        spp_gpu::Box box(p);
        std::ofstream out;
        if (resume && box.Positioning_Particles(out_name)) {
            out.open(out_name, std::ios::out | std::ios::app | std::ios::binary);
        } else {
            std::vector<double> x(N), y(N), th(N);
            spp_random_state(7, N, L, L, x.data(), y.data(), th.data());
            box.Init(x, y, th);
            out.open(out_name, std::ios::out | std::ios::binary);
            out << &box;
        }
        std::ofstream pol(polar_name, resume ? std::ios::app : std::ios::out);
        for (int s = 0; s < steps; s += period) {
            box.Multi_Step(period);
            box.Save_Polarization(pol);
            if ((s / period + 1) % saving == 0) out << &box;
        }
        std::printf("box_demo: t=%.6f polarization=%.6f\n", box.t, box.polarization);
        // read the whole trajectory back frame by frame (operator>>) into a second box; the last
        // frame must reproduce the final state to float32 rounding
        out.close();
        std::ifstream in(out_name, std::ios::binary);
        spp_gpu::Box back(p);
        int frames = 0;
        while (in >> &back) ++frames;
        std::vector<double> xw, yw, xu, yu, th, bxw, byw, bxu, byu, bth;
        box.Get_State(xw, yw, xu, yu, th);
        back.Get_State(bxw, byw, bxu, byu, bth);
        double err = 0;
        for (int i = 0; i < N; ++i) {
            err = std::max(err, std::fabs(bxu[i] - (double)(float)xu[i]));
            err = std::max(err, std::fabs(bth[i] - (double)(float)th[i]));
        }
        std::printf("box_demo: read back %d frames, last t=%.6f, max error %.3g\n", frames, back.t, err);
    } catch (const spp_gpu::Error& err) {
        std::fprintf(stderr, "box_demo: error %d: %s\n", err.code, err.what());
        return 1;
    }
    return 0;
}
