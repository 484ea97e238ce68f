// State arithmetic of the Box facade (spp_amd/csrc/box.hpp, the reference's State_Hyper_Vector):
// prints the checks' measured errors and values, one per line, for tests/test_state_vector.py.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../spp_amd/csrc/box.hpp"

int main()
{
    const int n = 200;
    const double L = 10.0;
    std::mt19937_64 g(3);
    std::uniform_real_distribution<double> U(-L, L), A(-M_PI, M_PI), E(-1e-3, 1e-3);
    spp_gpu::State s1(n, L, L), s2(n, L, L), d(n, L, L);
    for (int i = 0; i < n; ++i) {
        s1.x[i] = U(g);
        s1.y[i] = U(g);
        s1.theta[i] = A(g);
        s1.xu[i] = s1.x[i] + 2 * L * (i % 3);
        s1.yu[i] = s1.y[i];
        // s2: s1 moved by up to 1e-3, across the periodic edges for some particles
        s2.x[i] = s1.x[i] + E(g);
        s2.y[i] = s1.y[i] + E(g);
        s2.theta[i] = s1.theta[i] + E(g);
        s2.xu[i] = s1.xu[i] + (s2.x[i] - s1.x[i]);
        s2.yu[i] = s1.yu[i] + (s2.y[i] - s1.y[i]);
    }
    s2.Periodic_Transform();
    // (s2 - s1) is a small deviation even across the edges (minimum image), and adding it back
    // restores s2 (positions mod the box, theta mod 2 pi, unwrapped positions exactly)
    spp_gpu::State dev = s2 - s1;
    double mx = 0;
    for (int i = 0; i < n; ++i) mx = std::fmax(mx, std::fmax(std::fabs(dev.x[i]), std::fabs(dev.theta[i])));
    std::printf("max_dev %.6g\n", mx);
    spp_gpu::State back = s1 + dev;
    double er = 0, eu = 0;
    for (int i = 0; i < n; ++i) {
        er = std::fmax(er, std::fabs(back.x[i] - s2.x[i]) + std::fabs(back.y[i] - s2.y[i]) +
                               std::fabs(std::remainder(back.theta[i] - s2.theta[i], 2 * M_PI)));
        eu = std::fmax(eu, std::fabs(back.xu[i] - s2.xu[i]) + std::fabs(back.yu[i] - s2.yu[i]));
    }
    std::printf("roundtrip %.6g %.6g\n", er, eu);
    // an edge case: x = 9.9 and -9.9 in a box of half width 10 are 0.2 apart
    spp_gpu::State a(1, L, L), b(1, L, L);
    a.x[0] = 9.9;
    b.x[0] = -9.9;
    std::printf("edge %.17g\n", (a - b).x[0]);
    // Square wraps theta into (-pi, pi]: theta = 3 pi / 2 counts as -pi / 2
    spp_gpu::State c(1, L, L);
    c.theta[0] = 1.5 * M_PI;
    std::printf("square %.17g\n", c.Square());
    // Unit has magnitude 1; the product of a vector with itself is its Square when nothing wraps
    spp_gpu::State u = dev;
    u.Unit();
    std::printf("unit %.17g %.17g\n", u.Magnitude(), std::fabs(dev * dev - dev.Square()));
    // Max_Index: the largest |theta|
    spp_gpu::State m(5, L, L);
    m.theta = {0.1, -2.5, 0.3, 2.4, -0.2};
    std::printf("maxindex %d\n", m.Max_Index());
    return 0;
}
