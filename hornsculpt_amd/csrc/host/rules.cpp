// Rule tables. gauss_legendre01 follows pair_quadrature.cpp:685-713 (Newton on
// P_n from the cos(pi(i+3/4)/(n+1/2)) guess, weights 1/((1-x^2)P_n'^2)), so its
// nodes agree bit for bit with the planner it replaces. triangle_rule returns
// the positive-weight Dunavant rules of degree 2/4/6/8 (Dunavant 1985, Int. J.
// Numer. Meth. Eng. 21:1129) and the collapsed tensor Gauss rules of degree
// 12/14 (pair_quadrature.cpp:715-769); all weights sum to 1/2.
#include "rules.hpp"

#include <cmath>
#include <limits>
#include <numbers>
#include <stdexcept>

namespace hsbem {

std::vector<GaussPoint> gauss_legendre01(unsigned order) {
    if (order < 2 || order > 20)
        throw std::invalid_argument("Gauss-Legendre order must lie in 2..20");
    std::vector<GaussPoint> rule(order);
    for (unsigned i = 0; i < (order + 1) / 2; ++i) {
        double x = std::cos(std::numbers::pi * (static_cast<double>(i) + .75) / (static_cast<double>(order) + .5));
        double dp = 0;
        for (unsigned it = 0; it < 32; ++it) {
            double p0 = 1, p1 = x;
            for (unsigned n = 2; n <= order; ++n) {
                const double p = ((2 * n - 1) * x * p1 - (n - 1) * p0) / n;
                p0 = p1;
                p1 = p;
            }
            dp = order * (x * p1 - p0) / (x * x - 1);
            const double step = p1 / dp;
            x -= step;
            if (std::abs(step) < 4 * std::numeric_limits<double>::epsilon())
                break;
            if (it == 31)
                throw std::runtime_error("Gauss-Legendre node iteration did not converge");
        }
        const double w = 1 / ((1 - x * x) * dp * dp);
        rule[i] = {(1 - x) * .5, w};
        rule[order - 1 - i] = {(1 + x) * .5, w};
    }
    return rule;
}

std::vector<TrianglePoint> triangle_rule(unsigned degree) {
    if (degree == 12 || degree == 14) {
        // Collapsed square x=u, y=(1-u)v; the Jacobian (1-u) adds one degree in u.
        const auto g = gauss_legendre01((degree + 2) / 2);
        std::vector<TrianglePoint> rule;
        rule.reserve(g.size() * g.size());
        for (const auto &u : g)
            for (const auto &v : g)
                rule.push_back({{(1 - u.position) * (1 - v.position), u.position, (1 - u.position) * v.position},
                                u.weight * v.weight * (1 - u.position)});
        return rule;
    }
    if (degree == 2)
        return {{{2.0 / 3, 1.0 / 6, 1.0 / 6}, 1.0 / 6},
                {{1.0 / 6, 2.0 / 3, 1.0 / 6}, 1.0 / 6},
                {{1.0 / 6, 1.0 / 6, 2.0 / 3}, 1.0 / 6}};
    std::vector<TrianglePoint> rule;
    // Orbit generators: three points (a,a,b) permuted, six points (a,b,c) permuted.
    auto orbit3 = [&](double a, double b, double w) {
        for (unsigned i = 0; i < 3; ++i) {
            TrianglePoint p{{a, a, a}, w};
            p.bary[i] = b;
            rule.push_back(p);
        }
    };
    auto orbit6 = [&](double a, double b, double c, double w) {
        const double abc[3]{a, b, c};
        for (unsigned i = 0; i < 3; ++i)
            for (unsigned j = 0; j < 3; ++j)
                if (i != j)
                    rule.push_back({{abc[i], abc[j], abc[3 - i - j]}, w});
    };
    if (degree == 4) {
        orbit3(.445948490915965, .108103018168070, .1116907948390055);
        orbit3(.091576213509771, .816847572980459, .054975871827661);
    } else if (degree == 6) {
        orbit3(.249286745170910, .501426509658179, .0583931378631895);
        orbit3(.063089014491502, .873821971016996, .0254224531851035);
        orbit6(.310352451033785, .636502499121399, .053145049844816, .041425537809187);
    } else if (degree == 8) {
        rule.push_back({{1.0 / 3, 1.0 / 3, 1.0 / 3}, .0721578038388935});
        orbit3(.459292588292723, .081414823414554, .0475458171336425);
        orbit3(.170569307751760, .658861384496480, .051608685267359);
        orbit3(.050547228317031, .898905543365938, .016229248811599);
        orbit6(.263112829634638, .728492392955404, .008394777409958, .0136151570872175);
    } else {
        throw std::invalid_argument("triangle rule degree must be one of 2, 4, 6, 8, 12, 14");
    }
    return rule;
}

} // namespace hsbem
