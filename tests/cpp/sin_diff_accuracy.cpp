// Accuracy of spp::sin_diff (spp_amd/csrc/trig.hpp), the alignment torque's sin(theta_j - theta_i),
// against extended-precision sinl of the same rounded difference (what the reference's
// sin(p.theta - theta) computes).  Prints the largest relative error in ulps for four families of
// heading pairs in [-pi, pi): uniform, nearly aligned (|a - b| down to 1e-14), nearly anti-aligned
// (a - b near +-pi) and aligned across the wrap (a - b near +-2 pi); then a few exact values.
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

static double wrap(double t) { return t - std::floor(t / 6.283185307179586 + 0.5) * 6.283185307179586; }

int main()
{
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> U(-3.141592653589793, 3.141592653589793), V(-1.0, 1.0);
    double worst[4] = {0, 0, 0, 0};
    for (long t = 0; t < 4000000; ++t) {
        const int fam = (int)(t & 3);
        double a = U(g), b;
        const double tiny = std::ldexp(V(g), -(int)(t / 4 % 47));  // spreads from 1 down to 1e-14
        if (fam == 0)
            b = U(g);
        else if (fam == 1)
            b = wrap(a + tiny);
        else if (fam == 2)
            b = wrap(a + 3.141592653589793 + tiny);
        else {
            a = 3.141592653589793 - std::fabs(tiny) * 0.5;  // a near pi, b near -pi: same heading
            b = -3.141592653589793 + std::fabs(V(g)) * std::fabs(tiny) * 0.5;
        }
        const double d = a - b;
        const double s = spp::sin_diff(a, b);
        worst[fam] = std::fmax(worst[fam], ulps(s, sinl((long double)d)));
    }
    std::printf("%.6f %.6f %.6f %.6f\n", worst[0], worst[1], worst[2], worst[3]);
    // exact values: equal headings give 0; a tiny difference is returned as is
    std::printf("%.17g %.17g %.17g %.17g\n", spp::sin_diff(0.3, 0.3), spp::sin_diff(1e-12, 0.0),
                spp::sin_diff(0.0, -1e-300), spp::sin_diff(1.5707963267948966, 0.0));
    return 0;
}
