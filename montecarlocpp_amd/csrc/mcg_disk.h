// mcg_disk.h -- area of a rectangle that lies inside a disk.
//
// Used to give tally cells that a cylindrical pore cuts their material volume
// (cell volume x (1 - area inside the pore / cell area)).  Method: with
// X = x - cx, the overlap is the integral over X of the chord length of the
// disk clipped to [y0, y1]; substituting X = r sin(theta) removes the square
// root's singular derivative at X = +-r, and the clipped chord has kinks only
// where the disk's boundary crosses y0 or y1, so the theta range is split at
// those angles and each smooth piece is integrated with 24-point
// Gauss-Legendre quadrature (nodes and weights computed once by Newton's
// method on the Legendre polynomial).  Compiled by the engine's host builder
// and by the CPU checker; tests/test_geometry.py pins it against brute force.
#pragma once

#include <algorithm>
#include <cmath>
#include <vector>

namespace mcg_disk {

struct GaussLegendre {
    static constexpr int N = 24;
    double x[N], w[N];
    GaussLegendre() {
        for (int i = 0; i < N; ++i) {
            double z = std::cos(3.141592653589793 * (i + 0.75) / (N + 0.5));
            for (int it = 0; it < 100; ++it) {
                double p1 = 1.0, p2 = 0.0;
                for (int k = 1; k <= N; ++k) {
                    const double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * k - 1.0) * z * p2 - (k - 1.0) * p3) / k;
                }
                const double dp = N * (z * p1 - p2) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) {
                    x[i] = z;
                    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
                    break;
                }
            }
        }
    }
};

inline const GaussLegendre& gl() {
    static const GaussLegendre g;
    return g;
}

// Area of [x0, x1] x [y0, y1] inside the disk of radius r centred at (cx, cy).
inline double rect_in_disk(double x0, double x1, double y0, double y1, double cx, double cy,
                           double r) {
    const double X0 = std::max(x0 - cx, -r), X1 = std::min(x1 - cx, r);
    const double Y0 = y0 - cy, Y1 = y1 - cy;
    if (!(X1 > X0) || !(Y1 > Y0) || Y0 >= r || Y1 <= -r) return 0.0;
    // wholly inside (every corner in the disk) or wholly outside (the
    // rectangle's nearest point to the centre outside it): exact answers
    const double ax = std::max(std::fabs(x0 - cx), std::fabs(x1 - cx));
    const double ay = std::max(std::fabs(y0 - cy), std::fabs(y1 - cy));
    if (ax * ax + ay * ay <= r * r) return (x1 - x0) * (y1 - y0);
    const double nx = std::max({x0 - cx, 0.0, cx - x1}), ny = std::max({y0 - cy, 0.0, cy - y1});
    if (nx * nx + ny * ny >= r * r) return 0.0;
    // integrand in theta: clipped chord at X = r sin(theta), times r cos(theta)
    auto f = [&](double th) {
        const double h = r * std::cos(th);   // half chord at X = r sin(theta)
        const double top = std::min(Y1, h), bot = std::max(Y0, -h);
        return top > bot ? (top - bot) * h : 0.0;
    };
    const double t0 = std::asin(X0 / r), t1 = std::asin(X1 / r);
    std::vector<double> cuts{t0, t1};
    for (double Y : {Y0, Y1}) {
        if (std::fabs(Y) < r) {
            const double a = std::acos(std::fabs(Y) / r);   // where the boundary meets |y| = |Y|
            for (double c : {a, -a})
                if (c > t0 && c < t1) cuts.push_back(c);
        }
    }
    std::sort(cuts.begin(), cuts.end());
    const GaussLegendre& g = gl();
    double area = 0.0;
    for (size_t k = 0; k + 1 < cuts.size(); ++k) {
        const double a = cuts[k], b = cuts[k + 1];
        if (!(b > a)) continue;
        const double m = 0.5 * (a + b), hw = 0.5 * (b - a);
        double s = 0.0;
        for (int i = 0; i < GaussLegendre::N; ++i) s += g.w[i] * f(m + hw * g.x[i]);
        area += s * hw;
    }
    return area;
}

}  // namespace mcg_disk
