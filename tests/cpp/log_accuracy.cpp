// Accuracy of spp::log_pos (spp_amd/csrc/trig.hpp), the noise draw's log(1 - u), against
// extended-precision logl.  Prints the largest error in ulps of the exact result for: the noise's
// own arguments 1 - u (u a 53-bit uniform in [0, 1), so 1 - u in [2^-53, 1]); arguments within a
// few ulps of 1 and of the reduction's boundaries (powers of two and sqrt(1/2) times them); and a
// log-uniform spread over [2^-60, 2^10].  Then a few exact values.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../spp_amd/csrc/trig.hpp"

static double ulps(double got, long double want)
{
    double w = (double)want;
    if (w == 0) return got == 0 ? 0.0 : INFINITY;
    double u = std::nextafter(std::fabs(w), INFINITY) - std::fabs(w);
    return (double)(std::fabs((long double)got - want) / u);
}

int main()
{
    std::mt19937_64 g(5);
    double worst[3] = {0, 0, 0};
    // the noise's arguments: 1 - k 2^-53, k a 53-bit integer
    for (long t = 0; t < 3000000; ++t) {
        const unsigned long long k = g() >> 11;
        const double x = 1.0 - (double)k * 0x1p-53;
        worst[0] = std::fmax(worst[0], ulps(spp::log_pos(x), logl((long double)x)));
    }
    // near 1 and near the reduction's boundaries, a few ulps either side
    const double centres[] = {1.0, 0.5, 0.25, 2.0, 0x1p-30, 0x1p-52, 0.7071067811865476, 0.35355339059327379,
                              1.4142135623730951};
    for (double c : centres) {
        double x = c;
        for (int s = 0; s < 2000; ++s) x = std::nextafter(x, 0.0);
        for (int s = 0; s < 4000; ++s, x = std::nextafter(x, INFINITY))
            worst[1] = std::fmax(worst[1], ulps(spp::log_pos(x), logl((long double)x)));
    }
    std::uniform_real_distribution<double> E(-60.0, 10.0);
    for (long t = 0; t < 2000000; ++t) {
        const double x = std::exp2(E(g));
        worst[2] = std::fmax(worst[2], ulps(spp::log_pos(x), logl((long double)x)));
    }
    std::printf("%.6f %.6f %.6f\n", worst[0], worst[1], worst[2]);
    std::printf("%.17g %.17g %.17g %.17g\n", spp::log_pos(1.0), spp::log_pos(2.0), spp::log_pos(0.5),
                spp::log_pos(0x1p-53));
    return 0;
}
