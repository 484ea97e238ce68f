// The engine's cell index (spp_amd/csrc/binning.hpp) against the division it stands for,
// floor((x + L) nc / 2L) clamped to [0, nc), on uniform coordinates and on coordinates within a few
// ulps of every kind of cell edge, over random box sizes and grids.  Prints "checked mismatches".
#include <cmath>
#include <cstdio>
#include <random>

#include "../../spp_amd/csrc/binning.hpp"

static int by_division(double x, double L, double L2, int nc)
{
    const int c = (int)std::floor((x + L) * nc / L2);
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

int main()
{
    std::mt19937_64 g(20261204);
    long n = 0, bad = 0;
    for (int trial = 0; trial < 1500; ++trial) {
        const double L = std::uniform_real_distribution<double>(0.6, 3000.0)(g);
        const int nc = 3 + (int)(g() % 30000);
        const double L2 = 2 * L, inv = 1.0 / L2;
        std::uniform_real_distribution<double> U(-L, L);
        for (int k = 0; k < 4000; ++k) {
            double x = U(g);
            if (k % 2 == 0) {  // a few ulps either side of a cell edge (or of the box edges)
                const int c = (int)(g() % (nc + 1));
                x = -L + L2 * c / nc;
                const int steps = (int)(g() % 6);
                const double dir = (g() & 1) ? 1e300 : -1e300;
                for (int s = 0; s < steps; ++s) x = std::nextafter(x, dir);
                if (x < -L) x = -L;
                if (x >= L) x = std::nextafter(L, -1e300);
            }
            ++n;
            if (spp_bin::cell_coord(x, L, L2, inv, nc, 0) != by_division(x, L, L2, nc)) ++bad;
        }
    }
    std::printf("%ld %ld\n", n, bad);
    return 0;
}
