#include "fmm_scaled.hpp"

#include "fmm_levels.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numbers>
#include <stdexcept>

namespace hsbem {
namespace {

using C = std::complex<double>;

std::size_t lm(unsigned l, int m) { return static_cast<std::size_t>(l) * l + l + m; }

std::size_t coefficient_count(unsigned order) {
    if (order > kMaximumSphericalOrder)
        throw std::invalid_argument("Scaled spherical reference order exceeds its bounded CPU contract.");
    return static_cast<std::size_t>(order + 1) * (order + 1);
}

double component(Vec3 v, unsigned i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }

void check_domain(double k, double a) {
    if (!std::isfinite(k) || k < 0 || !std::isfinite(a) || a <= 0 || k * a > kMaximumSphericalElectricalRadius)
        throw std::invalid_argument("Scaled spherical waves require k>=0, a>0 and supported ka.");
}

void check_expansion(unsigned order, double a, const std::vector<C> &c, double k) {
    check_domain(k, a);
    if (c.size() != coefficient_count(order))
        throw std::invalid_argument("Scaled spherical expansion has an inconsistent coefficient count.");
    for (const C &v : c)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("Scaled spherical expansion contains a non-finite coefficient.");
}

// product rule a*b with gradients
WaveValue product(const WaveValue &a, const WaveValue &b) {
    WaveValue r;
    r.value = a.value * b.value;
    for (unsigned i = 0; i < 3; ++i)
        r.gradient[i] = a.gradient[i] * b.value + a.value * b.gradient[i];
    return r;
}

WaveValue scaled(WaveValue a, double f) {
    a.value *= f;
    for (C &g : a.gradient)
        g *= f;
    return a;
}

WaveValue difference(WaveValue a, const WaveValue &b) {
    a.value -= b.value;
    for (unsigned i = 0; i < 3; ++i)
        a.gradient[i] -= b.gradient[i];
    return a;
}

// Normalised solid harmonics |u|^l Y_lm(u/|u|) as polynomials in (x + iy, z, r^2):
// the Y_lm recursions with mu -> z, rho e^{i phi} -> x + iy, and the two-term step
// carrying r^2; gradients by the product rule, finite at the poles and the origin.
std::vector<WaveValue> solid_harmonics(unsigned order, Vec3 u) {
    std::vector<WaveValue> out(coefficient_count(order));
    const WaveValue xy{{u.x, u.y}, {C{1, 0}, C{0, 1}, C{}}}, z{{u.z, 0}, {C{}, C{}, C{1, 0}}},
        r2{{u.x * u.x + u.y * u.y + u.z * u.z, 0}, {C{2 * u.x, 0}, C{2 * u.y, 0}, C{2 * u.z, 0}}};
    WaveValue diag;
    diag.value = 1 / std::sqrt(4 * std::numbers::pi);
    for (unsigned m = 0; m <= order; ++m) {
        const int im = static_cast<int>(m);
        if (m)
            diag = scaled(product(xy, diag), -std::sqrt((2. * m + 1) / (2. * m)));
        out[lm(m, im)] = diag;
        if (m < order)
            out[lm(m + 1, im)] = scaled(product(z, diag), std::sqrt(2. * m + 3));
        for (unsigned l = m + 2; l <= order; ++l) {
            const double den = static_cast<double>(l) * l - static_cast<double>(m) * m,
                         a = std::sqrt((4. * l * l - 1) / den),
                         b = std::sqrt((2. * l + 1) * ((l - 1.) * (l - 1.) - static_cast<double>(m) * m) /
                                       ((2. * l - 3) * den));
            out[lm(l, im)] = difference(scaled(product(z, out[lm(l - 1, im)]), a),
                                        scaled(product(r2, out[lm(l - 2, im)]), b));
        }
        if (m)
            for (unsigned l = m; l <= order; ++l) {
                WaveValue neg = out[lm(l, im)];
                const double sign = m % 2 ? -1. : 1.;
                neg.value = sign * std::conj(neg.value);
                for (C &g : neg.gradient)
                    g = sign * std::conj(g);
                out[lm(l, -im)] = neg;
            }
    }
    return out;
}

// Jhat_l(z) = (2l+1)!! j_l(z) / z^l, l = 0..order: the power series while its first
// correction z^2 / (2 (2l+3)) is at most one (alternating, decreasing terms), the
// j_l table divided by z^l / (2l+1)!! beyond.
std::vector<double> normalized_regular(unsigned order, double z) {
    if (!std::isfinite(z) || z < 0 || z > 2 * kMaximumSphericalElectricalRadius)
        throw std::invalid_argument("Scaled regular reference projection exceeds its stable LF domain.");
    std::vector<double> out(order + 1);
    const RegularRadialTable table = z > 1 ? regular_radial_table(order, z) : RegularRadialTable{};
    double lead = 1;
    for (unsigned l = 0; l <= order; ++l) {
        if (l)
            lead *= z / (2. * l + 1);
        if (z > 1 && z * z > 2 * (2. * l + 3)) {
            out[l] = table.values[l] / lead;
            continue;
        }
        double term = 1, sum = 1;
        for (unsigned n = 1; n < 128; ++n) {
            term *= -z * z / (2. * n * (2. * l + 2 * n + 1));
            sum += term;
            if (std::abs(term) < std::numeric_limits<double>::epsilon() * std::abs(sum))
                break;
        }
        out[l] = sum;
    }
    return out;
}

void accumulate(WaveValue &out, const WaveValue &basis, C c) {
    out.value += c * basis.value;
    for (unsigned i = 0; i < 3; ++i)
        out.gradient[i] += c * basis.gradient[i];
}

// (a/r)^(l+1) Hhat_l(kr), l = 0..order+1, with t = a/r:
// O_0 = t e^{-ikr}, O_1 = t (1 + ikr) O_0, O_{l+1} = t O_l - (ka)^2 / ((2l-1)(2l+1)) O_{l-1}
std::vector<C> outgoing_radial(unsigned order, double k, double a, double r) {
    const double t = a / r, ka = k * a;
    const C i{0, 1};
    std::vector<C> out(order + 2);
    out[0] = t * std::exp(-i * (k * r));
    out[1] = t * (1.0 + i * (k * r)) * out[0];
    for (unsigned l = 1; l <= order; ++l)
        out[l + 1] = t * out[l] - ka * ka / ((2. * l - 1) * (2. * l + 1)) * out[l - 1];
    return out;
}

} // namespace

std::vector<WaveValue> scaled_regular_waves(unsigned order, double k, double a, Vec3 p) {
    check_domain(k, a);
    const auto solid = solid_harmonics(order, vscale(p, 1 / a));
    const auto radial = normalized_regular(order + 1, k * std::hypot(p.x, p.y, p.z));
    std::vector<WaveValue> out(solid.size());
    for (unsigned l = 0; l <= order; ++l)
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            const std::size_t j = lm(l, m);
            out[j].value = radial[l] * solid[j].value;
            for (unsigned i = 0; i < 3; ++i)
                out[j].gradient[i] = radial[l] * solid[j].gradient[i] / a -
                                     (k * k * component(p, i) / (2. * l + 3)) * radial[l + 1] * solid[j].value;
        }
    return out;
}

std::vector<WaveValue> scaled_outgoing_waves(unsigned order, double k, double a, Vec3 p) {
    check_domain(k, a);
    const double r = std::hypot(p.x, p.y, p.z);
    if (!std::isfinite(r) || r <= 0)
        throw std::invalid_argument("Scaled outgoing basis requires a finite point outside its singular center.");
    const Vec3 u = vscale(p, 1 / r);
    const auto solid = solid_harmonics(order, u);
    const auto radial = outgoing_radial(order, k, a, r);
    std::vector<WaveValue> out(coefficient_count(order));
    for (unsigned l = 0; l <= order; ++l) {
        const C dr = (l / r) * radial[l] - (2. * l + 1) / a * radial[l + 1];
        const double norm = 1 / (a * (2. * l + 1));
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            const std::size_t j = lm(l, m);
            WaveValue w;
            w.value = norm * radial[l] * solid[j].value;
            for (unsigned i = 0; i < 3; ++i)
                w.gradient[i] = norm * (dr * component(u, i) * solid[j].value +
                                        radial[l] / r * (solid[j].gradient[i] - l * component(u, i) * solid[j].value));
            out[j] = w;
        }
    }
    return out;
}

WaveValue evaluate_scaled_multipole(const ScaledExpansion &s, double k, Vec3 p) {
    check_expansion(s.order, s.radius, s.coefficients, k);
    const auto basis = scaled_outgoing_waves(s.order, k, s.radius, p);
    WaveValue out;
    for (std::size_t j = 0; j < basis.size(); ++j)
        accumulate(out, basis[j], s.coefficients[j]);
    return out;
}

WaveValue evaluate_scaled_local(const ScaledExpansion &s, double k, Vec3 p) {
    check_expansion(s.order, s.radius, s.coefficients, k);
    const auto basis = scaled_regular_waves(s.order, k, s.radius, p);
    WaveValue out;
    for (std::size_t j = 0; j < basis.size(); ++j)
        accumulate(out, basis[j], s.coefficients[j]);
    return out;
}

std::vector<C> scaled_multipole_to_spectrum(const ScaledExpansion &s, double k, const SphereRule &rule) {
    check_expansion(s.order, s.radius, s.coefficients, k);
    auto c = s.coefficients;
    double f = 1; // (ka)^l / (2l+1)!!
    for (unsigned l = 0; l <= s.order; ++l) {
        if (l)
            f *= k * s.radius / (2. * l + 1);
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m)
            c[lm(l, m)] *= f;
    }
    return multipole_to_spectrum(c, s.order, rule);
}

ScaledExpansion spectrum_to_scaled_local(const std::vector<C> &incoming, const SphereRule &rule, unsigned order,
                                         double k, double a) {
    check_domain(k, a);
    coefficient_count(order);
    ScaledExpansion out{order, a, spectrum_to_local(incoming, rule, order)};
    double f = 1;
    for (unsigned l = 0; l <= order; ++l) {
        if (l)
            f *= k * a / (2. * l + 1);
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m)
            out.coefficients[lm(l, m)] *= f;
    }
    return out;
}

AxialPlan make_axial_plan(AxialKind kind, unsigned ps, unsigned pt, double k, double as, double at, double d,
                          unsigned polar_count, double projection_radius) {
    check_domain(k, as);
    check_domain(k, at);
    coefficient_count(ps);
    coefficient_count(pt);
    if (kind != AxialKind::multipole_to_parent && kind != AxialKind::multipole_to_local &&
        kind != AxialKind::local_to_child)
        throw std::invalid_argument("Scaled axial translation kind is invalid.");
    if (!std::isfinite(d) || d < 0 || polar_count < 2 * std::max(ps, pt) + 1 || polar_count > 1024)
        throw std::invalid_argument("Scaled axial plan needs finite separation and independently resolved polar "
                                    "quadrature.");
    if (!std::isfinite(projection_radius) || projection_radius < 0)
        throw std::invalid_argument("Scaled axial projection radius must be finite and nonnegative.");
    const double probe = projection_radius > 0                  ? projection_radius
                         : kind == AxialKind::multipole_to_parent ? as * 1.01
                         : kind == AxialKind::multipole_to_local  ? std::min(at, d * .5)
                                                                  : at;
    if (!(probe > 0) || (kind == AxialKind::multipole_to_local && probe >= d))
        throw std::invalid_argument("Scaled axial Cauchy probe intersects the outgoing singular center.");
    AxialPlan plan{kind, ps, pt, polar_count, {0}, {}, probe};
    const unsigned mmax = std::min(ps, pt);
    for (unsigned m = 0; m <= mmax; ++m)
        plan.block_offsets.push_back(plan.block_offsets.back() +
                                     static_cast<std::size_t>(pt - m + 1) * (ps - m + 1));
    plan.coefficients.assign(plan.block_offsets.back(), C{});
    const SphereRule sphere = make_sphere_rule(polar_count - 1);
    const auto target_radial = normalized_regular(pt + 1, k * probe);
    for (unsigned ring = 0; ring < polar_count; ++ring) {
        // azimuth 0 of each ring; the azimuth integral is analytic (2 pi x Gauss weight)
        const SphereSample &s = sphere.samples[static_cast<std::size_t>(ring) * 2 * polar_count];
        const Vec3 u = s.direction;
        const double w = s.weight * (2 * polar_count);
        const Vec3 point = vadd({0, 0, d}, vscale(u, probe));
        const auto angular = spherical_harmonics(std::max(ps, pt), u);
        std::vector<WaveValue> src, tgt;
        if (kind == AxialKind::multipole_to_parent) {
            src = scaled_outgoing_waves(ps, k, as, vscale(u, probe));
            tgt = scaled_regular_waves(pt, k, at, point);
        } else if (kind == AxialKind::multipole_to_local) {
            src = scaled_outgoing_waves(ps, k, as, point);
        } else {
            src = scaled_regular_waves(ps, k, as, point);
        }
        for (unsigned m = 0; m <= mmax; ++m)
            for (unsigned l = m; l <= pt; ++l) {
                const std::size_t row = lm(l, static_cast<int>(m));
                const double dj = l * target_radial[l] - k * k * probe * probe / (2. * l + 3) * target_radial[l + 1],
                             den = target_radial[l] * target_radial[l] + dj * dj;
                for (unsigned n = m; n <= ps; ++n) {
                    const std::size_t col = lm(n, static_cast<int>(m));
                    C src_normal{};
                    for (unsigned i = 0; i < 3; ++i)
                        src_normal += src[col].gradient[i] * component(u, i);
                    C v{};
                    if (kind == AxialKind::multipole_to_parent) {
                        C tgt_normal{};
                        for (unsigned i = 0; i < 3; ++i)
                            tgt_normal += tgt[row].gradient[i] * component(u, i);
                        v = probe * probe * (src[col].value * std::conj(tgt_normal) - src_normal * std::conj(tgt[row].value));
                    } else {
                        if (!(den > 0) || !std::isfinite(den))
                            throw std::runtime_error("Scaled axial Cauchy projection has an unresolved radial "
                                                     "normalization.");
                        v = std::pow(at / probe, l) * std::conj(angular[row]) *
                            (target_radial[l] * src[col].value + dj * probe * src_normal) / den;
                    }
                    plan.coefficients[plan.block_offsets[m] + static_cast<std::size_t>(l - m) * (ps - m + 1) + n - m] +=
                        w * v;
                }
            }
    }
    for (const C &v : plan.coefficients)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::runtime_error("Scaled axial plan has a non-finite translation coefficient.");
    return plan;
}

std::vector<C> apply_axial_plan(const AxialPlan &p, const std::vector<C> &source) {
    if (source.size() != coefficient_count(p.source_order) ||
        p.block_offsets.size() != std::min(p.source_order, p.target_order) + 2 || p.block_offsets.front() != 0 ||
        p.block_offsets.back() != p.coefficients.size())
        throw std::invalid_argument("Scaled axial application has inconsistent dimensions.");
    std::vector<C> out(coefficient_count(p.target_order));
    for (const C &v : source)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("Scaled axial source contains a non-finite coefficient.");
    for (unsigned m = 0; m <= std::min(p.source_order, p.target_order); ++m) {
        const std::size_t expected = static_cast<std::size_t>(p.target_order - m + 1) * (p.source_order - m + 1);
        if (p.block_offsets[m] > p.block_offsets[m + 1] || p.block_offsets[m + 1] > p.coefficients.size() ||
            p.block_offsets[m + 1] - p.block_offsets[m] != expected)
            throw std::invalid_argument("Scaled axial application has an invalid m-block span.");
        for (const int sm : {static_cast<int>(m), -static_cast<int>(m)}) {
            for (unsigned l = m; l <= p.target_order; ++l) {
                C sum{};
                for (unsigned n = m; n <= p.source_order; ++n)
                    sum += p.coefficients[p.block_offsets[m] + static_cast<std::size_t>(l - m) * (p.source_order - m + 1) +
                                          n - m] *
                           source[lm(n, sm)];
                out[lm(l, sm)] = sum;
            }
            if (m == 0)
                break;
        }
    }
    for (const C &v : out)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::runtime_error("Scaled axial application produced a non-finite coefficient.");
    return out;
}

} // namespace hsbem
