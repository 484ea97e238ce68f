// Reads every healthy frame of a trajectory (argv[1]; argv[2] = 1 if particles carry theta),
// writes them to argv[3] and prints one line per frame: t Ns Nm first_x.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "../../spp_amd/csrc/trajectory.hpp"

int main(int argc, char** argv)
{
    if (argc < 4) return 2;
    std::ifstream is(argv[1], std::ios::binary);
    bool th = std::atoi(argv[2]) != 0;
    auto frames = spp_traj::read_healthy_frames(is, th);
    std::ofstream os(argv[3], std::ios::binary);
    for (auto& f : frames) {
        spp_traj::write_frame(os, f);
        std::printf("%.17g %d %d %.9g\n", f.t, f.Ns(), f.Nm(), f.Ns() ? f.x[0] : (f.Nm() ? f.mx[0] : 0.0f));
    }
    return 0;
}
