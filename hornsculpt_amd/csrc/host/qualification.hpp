// Regular-rule qualification for disjoint panel pairs, shared by the host
// planner and the device assembly kernels.
//
// Restates modules/bem/galerkin/regular_qualification.hpp:17-121: a centroid
// ball per panel (outward-rounded radius), a 20-degree minimum-angle shape
// test, and the qualified degree 4/6/8 from (k*h, ball separation in units of
// the larger edge, 0.2*tolerance budget). Zero means "use the adaptive path".
#pragma once

#include "space.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>
#include <string>

namespace hsbem {

struct PanelBounds {
    Vec3 center;
    double radius{}, max_edge{};
    unsigned qualified_shape{}; // 1 when every interior angle is at least 20 degrees
};

inline PanelBounds make_panel_bounds(const Panel &p) {
    PanelBounds b{p.centroid, 0, p.max_edge, 1};
    double edge[3]{};
    for (unsigned i = 0; i < 3; ++i) {
        const Vec3 &a = p.vertices[i], &c = p.vertices[(i + 1) % 3];
        edge[i] = std::hypot(a.x - c.x, a.y - c.y, a.z - c.z);
        b.radius = std::max(b.radius, std::hypot(a.x - p.centroid.x, a.y - p.centroid.y, a.z - p.centroid.z));
    }
    b.radius = std::nextafter(b.radius * (1 + 8 * std::numeric_limits<double>::epsilon()),
                              std::numeric_limits<double>::infinity());
    constexpr double cos20 = .939692620785908384;
    bool ok = true;
    for (unsigned i = 0; i < 3; ++i) {
        const double a = edge[i], c = edge[(i + 1) % 3], d = edge[(i + 2) % 3];
        ok = ok && a > 0 && c > 0 && (a * a + c * c - d * d) <= 2 * a * c * cos20;
    }
    ok = ok && std::isfinite(b.radius) && b.max_edge > 0;
    b.qualified_shape = ok ? 1u : 0u;
    return b;
}

// Same multiplication/comparison order as the reference so that the boundary
// cases agree bit for bit between planner, kernels and reference plans.
HS_HD inline unsigned qualified_regular_degree(const PanelBounds &t, const PanelBounds &s, double k,
                                               double tolerance, unsigned minimum_degree = 4) {
    if (!t.qualified_shape || !s.qualified_shape)
        return 0;
    const double h = t.max_edge > s.max_edge ? t.max_edge : s.max_edge, kh = k * h,
                 dx = t.center.x - s.center.x, dy = t.center.y - s.center.y, dz = t.center.z - s.center.z,
                 d2 = dx * dx + dy * dy + dz * dz, rsum = t.radius + s.radius, budget = .2 * tolerance;
    auto separated = [&](double gap) {
        const double bound = rsum + gap * h;
        return d2 >= bound * bound;
    };
    constexpr double reserve = 1e-6;
    if (minimum_degree <= 4 && ((kh <= .5 && separated(3) && reserve + 4 * 2.3e-6 <= budget) ||
                                (kh <= .25 && separated(8) && reserve + 4 * 1e-7 <= budget)))
        return 4;
    if (minimum_degree <= 6 && ((kh <= 2 && separated(1) && reserve + 4 * 6e-7 <= budget) ||
                                (kh <= 1 && separated(2) && reserve + 4 * 2.4e-8 <= budget)))
        return 6;
    if (minimum_degree <= 8 && kh <= 2 &&
        ((separated(.5) && reserve + 4 * 7.5e-7 <= budget) || (separated(1) && reserve + 4 * 3.5e-8 <= budget)))
        return 8;
    return 0;
}

// Starting degree of the observer/far-field panel rule from the largest k*h
// (regular_qualification.hpp:98-111); k*h above 3.25 is rejected.
inline unsigned minimum_endpoint_degree(double max_kh, unsigned requested = 4) {
    if (!std::isfinite(max_kh) || max_kh < 0)
        throw std::invalid_argument("panel rule selection needs a finite k h");
    if (max_kh > 3.25)
        throw std::runtime_error("largest k h = " + std::to_string(max_kh) +
                                 " is above the qualified limit 3.25: use a finer mesh");
    const unsigned electrical = max_kh <= .25 ? 4 : max_kh <= 1 ? 6 : max_kh <= 2 ? 8 : 12;
    return std::max(electrical, requested);
}

inline unsigned refined_endpoint_degree(unsigned d) {
    return d < 4 ? 4 : d < 6 ? 6 : d < 8 ? 8 : d < 12 ? 12 : d < 14 ? 14 : 0;
}

} // namespace hsbem
