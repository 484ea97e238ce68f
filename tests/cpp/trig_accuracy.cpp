// Accuracy of spp::sincos_heading (spp_amd/csrc/trig.hpp) against extended-precision sinl/cosl.
// Prints the largest error in ulps of sin and cos over random angles in [-4, 4] and over angles
// within 2^-k of multiples of pi/2, then the values at a few exact points.
#include <cmath>
#include <cstdio>
#include <random>

#include "../../spp_amd/csrc/trig.hpp"

static double ulps(double got, long double want)
{
    double w = (double)want;
    double u = std::nextafter(std::fabs(w), INFINITY) - std::fabs(w);
    return (double)(std::fabs((long double)got - want) / (w == 0 ? 4.9e-324 : u));
}

int main()
{
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> U(-4.0, 4.0), V(-1.0, 1.0);
    double es = 0, ec = 0, ea = 0;
    for (long t = 0; t < 2000000; ++t) {
        double x = t < 400000 ? (int)(t % 5 - 2) * 1.5707963267948966 + std::ldexp(V(g), -(int)(t % 45)) : U(g);
        double s, c;
        spp::sincos_heading(x, &s, &c);
        long double S = sinl((long double)x), C = cosl((long double)x);
        es = std::fmax(es, ulps(s, S));
        ec = std::fmax(ec, ulps(c, C));
        ea = std::fmax(ea, (double)std::fmax(std::fabs(s - S), std::fabs(c - C)));
    }
    std::printf("%.6f %.6f %.6g\n", es, ec, ea);
    const double pts[] = {0.0, 1.5707963267948966, -3.141592653589793, 3.141592653589793, 0.5, -2.5};
    for (double x : pts) {
        double s, c;
        spp::sincos_heading(x, &s, &c);
        std::printf("%.17g %.17g %.17g\n", x, s, c);
    }
    return 0;
}
