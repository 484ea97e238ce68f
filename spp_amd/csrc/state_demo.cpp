// state_demo.cpp -- the state-vector and particle-level surface of the C++ Box facade on the GPU
// engine, the way the reference's Lyapunov code drives a Box (shared/lyapunov.h: Save the state,
// Load a perturbed copy, step, take the difference):
//   1. Save(State) at step s, run, Load(State), run again: the replay against the first run;
//   2. Load(state + eps * unit deviation), run, and the grown deviation's magnitude;
//   3. Reset_Forces and particle(i).Interact(particle(j)) over every listed pair against
//      Box::Interact's forces.
//
// usage: state_demo N L steps
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "box.hpp"

static double max_abs_diff(const std::vector<double>& a, const std::vector<double>& b)
{
    double m = 0;
    for (size_t i = 0; i < a.size(); ++i) m = std::max(m, std::fabs(a[i] - b[i]));
    return m;
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s N L steps\n", argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]);
    const double L = std::atof(argv[2]);
    const int steps = std::atoi(argv[3]);

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
    p.integrator = SPP_RK2;
    p.seed = 2;
    try {
        spp_gpu::Box box(p);
        std::vector<double> x(N), y(N), th(N);
        spp_random_state(7, N, L, L, x.data(), y.data(), th.data());
        box.Init(x, y, th);
        box.Multi_Step(steps);

        // 1. save, run, load, run again
        spp_gpu::State s0(N, L, L), s1(N, L, L), s2(N, L, L);
        box.Save(s0);
        box.Multi_Step(steps);
        box.Save(s1);
        box.Load(s0);
        box.Multi_Step(steps);
        box.Save(s2);
        std::printf("replay t=%.9f %.9f step=%lld %lld max_diff=%.3g %.3g %.3g\n", s1.clock.t, s2.clock.t,
                    (long long)s1.clock.step, (long long)s2.clock.step, max_abs_diff(s1.xu, s2.xu), max_abs_diff(s1.yu, s2.yu),
                    max_abs_diff(s1.theta, s2.theta));

        // 2. a perturbed copy: eps times a random unit deviation, stepped the same
        std::mt19937_64 g(11);
        std::uniform_real_distribution<double> ud(-1.0, 1.0);
        spp_gpu::State u(N, L, L);
        for (int i = 0; i < N; ++i) {
            u.x[i] = u.xu[i] = ud(g);
            u.y[i] = u.yu[i] = ud(g);
            u.theta[i] = ud(g);
        }
        u.Unit();
        const double eps = 1e-7;
        spp_gpu::State start = s0 + u * eps;
        box.Load(start);
        spp_gpu::State dev0 = start - s0;
        box.Multi_Step(steps);
        spp_gpu::State s3(N, L, L);
        box.Save(s3);
        spp_gpu::State dev = s3 - s1;
        std::printf("perturb initial=%.6g final=%.6g\n", dev0.Magnitude(), dev.Magnitude());

        // 3. particle-level interactions over the neighbour list against Box::Interact
        box.Load(s1);
        box.Update_Cells();
        box.Interact();
        std::vector<double> fx(N), fy(N), tq(N), gx(N), gy(N), gq(N);
        box.Get_Forces(fx, fy, tq);
        std::vector<int32_t> off(N + 1);
        int64_t nnz = 0;
        spp_get_neighbors(box.engine(), nullptr, nullptr, 0, &nnz);
        std::vector<int32_t> idx(std::max<int64_t>(nnz, 1));
        if (spp_get_neighbors(box.engine(), off.data(), idx.data(), (int64_t)idx.size(), &nnz) != SPP_OK) {
            std::fprintf(stderr, "state_demo: spp_get_neighbors failed\n");
            return 1;
        }
        box.Reset_Forces();
        int64_t pairs = 0;
        for (int i = 0; i < N; ++i)
            for (int k = off[i]; k < off[i + 1]; ++k)
                if (idx[k] > i) {
                    box.particle(i).Interact(box.particle(idx[k]));
                    ++pairs;
                }
        box.Get_Forces(gx, gy, gq);
        double scale = 0;
        for (int i = 0; i < N; ++i) scale = std::max({scale, std::fabs(fx[i]), std::fabs(fy[i]), std::fabs(tq[i])});
        std::printf("pairs %lld max_diff=%.3g %.3g %.3g scale=%.6g\n", (long long)pairs, max_abs_diff(fx, gx),
                    max_abs_diff(fy, gy), max_abs_diff(tq, gq), scale);
    } catch (const spp_gpu::Error& err) {
        std::fprintf(stderr, "state_demo: error %d: %s\n", err.code, err.what());
        return 1;
    }
    return 0;
}
