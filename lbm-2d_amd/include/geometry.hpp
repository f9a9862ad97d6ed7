// geometry.hpp -- obstacle shapes for the lbm-2d solver interface (src/geometry.hpp of
// the reference): NACA 4-digit section, affine transform, point-in-polygon.
#pragma once

#include <cmath>
#include <utility>
#include <vector>

// NACA 4-digit section "mpxx" (Jacobs et al., NACA Report 460) on n cosine-spaced
// chord stations, as one closed polygon: trailing edge -> upper surface -> leading
// edge -> lower surface (src/geometry.hpp:17-90 uses the same vertex order).
inline std::vector<std::pair<double, double>> naca_coords(int series, int n = 200)
{
    const double m = (series / 1000) / 100.0;
    const double p = ((series / 100) % 10) / 10.0;
    const double t = (series % 100) / 100.0;
    std::vector<double> xs(n), yt(n), yc(n, 0.0), slope(n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double beta = static_cast<double>(k) / (n - 1) * M_PI;
        const double x = 0.5 * (1.0 - std::cos(beta));
        xs[k] = x;
        // half thickness
        yt[k] = 5.0 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                           0.1015 * x * x * x * x);
        if (m > 1e-10 && p > 1e-10) {
            if (x < p) {
                yc[k] = m / (p * p) * (2.0 * p * x - x * x);
                slope[k] = 2.0 * m / (p * p) * (p - x);
            } else {
                const double q = (1.0 - p) * (1.0 - p);
                yc[k] = m / q * ((1.0 - 2.0 * p) + 2.0 * p * x - x * x);
                slope[k] = 2.0 * m / q * (p - x);
            }
        }
    }
    std::vector<std::pair<double, double>> poly;
    poly.reserve(2 * n);
    for (int k = n - 1; k >= 0; --k) {  // upper surface, trailing edge to leading edge
        const double th = std::atan(slope[k]);
        poly.emplace_back(xs[k] - yt[k] * std::sin(th), yc[k] + yt[k] * std::cos(th));
    }
    for (int k = 1; k < n; ++k) {  // lower surface back to the trailing edge
        const double th = std::atan(slope[k]);
        poly.emplace_back(xs[k] + yt[k] * std::sin(th), yc[k] - yt[k] * std::cos(th));
    }
    return poly;
}

// Scale about the origin, rotate by angle_deg, then translate to (cx, cy)
// (src/geometry.hpp:95-114).
inline void transform_points(std::vector<std::pair<double, double>>& pts, double cx, double cy, double scale,
                             double angle_deg)
{
    const double a = angle_deg * M_PI / 180.0;
    const double c = std::cos(a), s = std::sin(a);
    for (auto& pt : pts) {
        const double xs = pt.first * scale, ys = pt.second * scale;
        pt = {xs * c - ys * s + cx, xs * s + ys * c + cy};
    }
}

// Even-odd ray crossing test (src/geometry.hpp:120-134): a horizontal ray from
// (px, py) toward +x crosses an edge when the edge straddles py and the crossing
// lies to the right of px.
inline bool point_in_polygon(double px, double py, const std::vector<std::pair<double, double>>& poly)
{
    bool inside = false;
    const size_t n = poly.size();
    for (size_t i = 0, j = n - 1; i < n; j = i++) {
        const double xi = poly[i].first, yi = poly[i].second;
        const double xj = poly[j].first, yj = poly[j].second;
        if ((yi > py) != (yj > py) && px < (xj - xi) * (py - yi) / (yj - yi) + xi) inside = !inside;
    }
    return inside;
}
