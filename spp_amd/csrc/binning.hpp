// binning.hpp -- the cell index of a wrapped coordinate, shared by host and device code (the
// binning, the tile sort and the list build call it; tests/cpp/cell_index_check.cpp checks it).
#pragma once

#include <cmath>

#include "trig.hpp"  // SPP_HD

namespace spp_bin {

// Cell index floor((x + L) nc / 2L) along one axis (quirk: the reference's truncating form).  The
// quotient is formed as a product by the rounded reciprocal, which lies within 4 ulp of the rounded
// quotient, so both have the same floor unless the product sits within 8 ulp of an integer; only
// then (almost never) is the division itself evaluated.  The result is always the division's.
SPP_HD inline int cell_coord(double x, double L, double L2, double invL2, int nc, int quirk)
{
    int c;
    if (quirk) {  // serial/box.h:116-117: the cast truncates (x + L) before scaling
        c = (int)((double)((int)(x + L) * nc) / L2);
    } else {
        const double q = (x + L) * nc;
        const double a = q * invL2;
        const double k = std::floor(a);
        const double f = a - k;
        const double m = 8.0 * 2.220446049250313e-16 * std::fabs(a) + 1e-300;
        c = (f >= m && f <= 1.0 - m) ? (int)k : (int)std::floor(q / L2);
    }
    return c < 0 ? 0 : (c >= nc ? nc - 1 : c);
}

}  // namespace spp_bin
