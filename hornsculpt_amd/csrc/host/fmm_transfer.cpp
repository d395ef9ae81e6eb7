#include "fmm_transfer.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <math.h>
#include <numbers>
#include <stdexcept>

namespace hsbem {
namespace {

constexpr unsigned kMaximumOrder = 1024;
constexpr double kPi = std::numbers::pi;
using C = std::complex<double>;

void check_order(unsigned order) {
    if (order > kMaximumOrder)
        throw std::invalid_argument("Wideband angular order exceeds the supported CPU contract.");
}

double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

std::size_t lm(unsigned l, int m) { return static_cast<std::size_t>(l) * l + l + m; }

void check_rule(const SphereRule &r) {
    check_order(r.order);
    const auto n = static_cast<std::size_t>(r.order + 1);
    if (r.samples.size() != 2 * n * n)
        throw std::invalid_argument("Wideband sphere rule has an inconsistent sample count.");
    for (const auto &s : r.samples)
        if (!std::isfinite(s.weight) || s.weight <= 0 || !std::isfinite(dot(s.direction, s.direction)) ||
            std::abs(dot(s.direction, s.direction) - 1) > 1e-12)
            throw std::invalid_argument("Wideband sphere rule has an invalid direction or weight.");
}

void check_values(const std::vector<C> &v, const SphereRule &r) {
    check_rule(r);
    if (v.size() != r.samples.size())
        throw std::invalid_argument("Wideband spectrum does not match its angular rule.");
    for (const C &x : v)
        if (!std::isfinite(x.real()) || !std::isfinite(x.imag()))
            throw std::invalid_argument("Wideband spectrum contains a non-finite value.");
}

// exp(i a); the reference build evaluates cos and sin of one argument jointly
C phase(double a) {
    if (!std::isfinite(a))
        throw std::invalid_argument("Wideband translation phase is non-finite.");
    double s, c;
    ::sincos(a, &s, &c);
    return {c, s};
}

std::size_t tri(unsigned l, unsigned m) { return static_cast<std::size_t>(l) * (l + 1) / 2 + m; }

// Normalised associated Legendre functions (the Y_lm polar factor) of one ring.
std::vector<double> polar_factors(const AngularTransferPlan &p, double mu) {
    const unsigned L = p.bandlimit;
    std::vector<double> out(p.recurrence_first.size());
    const double rho = std::sqrt(std::max(0., 1 - mu * mu));
    double diag = 1 / std::sqrt(4 * kPi);
    for (unsigned m = 0; m <= L; ++m) {
        if (m)
            diag *= p.diagonal_scale[m] * rho;
        out[tri(m, m)] = diag;
        if (m < L)
            out[tri(m + 1, m)] = std::sqrt(2. * m + 3) * mu * diag;
        for (unsigned l = m + 2; l <= L; ++l)
            out[tri(l, m)] = p.recurrence_first[tri(l, m)] * mu * out[tri(l - 1, m)] -
                             p.recurrence_second[tri(l, m)] * out[tri(l - 2, m)];
    }
    return out;
}

// Rings of a canonical (polar-major, equispaced azimuth) rule.
std::vector<PolarSample> rings(const SphereRule &r) {
    check_rule(r);
    const unsigned na = 2 * (r.order + 1);
    std::vector<PolarSample> out;
    for (unsigned j = 0; j <= r.order; ++j) {
        const SphereSample &first = r.samples[static_cast<std::size_t>(j) * na];
        const double mu = first.direction.z, rho = std::sqrt(std::max(0., 1 - mu * mu));
        for (unsigned a = 0; a < na; ++a) {
            const SphereSample &s = r.samples[static_cast<std::size_t>(j) * na + a];
            const double ang = 2 * kPi * a / na;
            if (std::abs(s.direction.z - mu) > 1e-13 || std::abs(s.direction.x - rho * std::cos(ang)) > 1e-13 ||
                std::abs(s.direction.y - rho * std::sin(ang)) > 1e-13 ||
                std::abs(s.weight - first.weight) > 1e-13 * first.weight)
                throw std::invalid_argument("Wideband transfer requires a polar-major canonical sphere rule.");
        }
        out.push_back({mu, first.weight * na / (2 * kPi)});
    }
    return out;
}

// sum_q w_q v_q conj(Y_lm(d_q)), l <= order
std::vector<C> project(const std::vector<C> &v, const SphereRule &r, unsigned order) {
    check_values(v, r);
    if (order > r.order)
        throw std::invalid_argument("Wideband projection exceeds its angular sampling order.");
    check_order(order);
    std::vector<C> c(static_cast<std::size_t>(order + 1) * (order + 1));
    for (std::size_t q = 0; q < v.size(); ++q) {
        const auto y = spherical_harmonics(order, r.samples[q].direction);
        const C w = v[q] * r.samples[q].weight;
        for (std::size_t j = 0; j < c.size(); ++j)
            c[j] += w * std::conj(y[j]);
    }
    return c;
}

// sum_lm c_lm Y_lm(d_q) at every rule direction
std::vector<C> evaluate(const std::vector<C> &c, unsigned order, const SphereRule &r) {
    check_rule(r);
    check_order(order);
    if (c.size() != static_cast<std::size_t>(order + 1) * (order + 1))
        throw std::invalid_argument("Wideband harmonic coefficients have an invalid size.");
    std::vector<C> out(r.samples.size());
    for (std::size_t q = 0; q < out.size(); ++q) {
        const auto y = spherical_harmonics(order, r.samples[q].direction);
        for (std::size_t j = 0; j < c.size(); ++j)
            out[q] += c[j] * y[j];
    }
    return out;
}

} // namespace

SphereRule make_sphere_rule(unsigned order) {
    check_order(order);
    SphereRule rule{order, {}};
    const unsigned np = order + 1, na = 2 * np;
    rule.samples.reserve(static_cast<std::size_t>(np) * na);
    for (unsigned j = 0; j < np; ++j) {
        // Newton on P_np from the Tricomi-type initial guess cos(pi (j + 3/4) / (np + 1/2))
        double mu = std::cos(kPi * (j + .75) / (np + .5)), dp = 0;
        bool done = false;
        for (unsigned it = 0; it < 100; ++it) {
            double p0 = 0, p1 = 1;
            for (unsigned l = 1; l <= np; ++l) {
                const double p2 = ((2 * l - 1) * mu * p1 - (l - 1) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            dp = np * (mu * p1 - p0) / (mu * mu - 1);
            const double step = p1 / dp;
            mu -= step;
            if (std::abs(step) <= 4 * std::numeric_limits<double>::epsilon()) {
                done = true;
                break;
            }
        }
        if (!done)
            throw std::runtime_error("Wideband Gauss-Legendre rule did not converge.");
        const double w = 2 / ((1 - mu * mu) * dp * dp), rho = std::sqrt(std::max(0., 1 - mu * mu));
        for (unsigned m = 0; m < na; ++m) {
            const double phi = 2 * kPi * m / na;
            double sp, cp;
            ::sincos(phi, &sp, &cp); // the reference build evaluates this pair jointly
            rule.samples.push_back({{rho * cp, rho * sp, mu}, w * 2 * kPi / na});
        }
    }
    return rule;
}

std::vector<C> spherical_harmonics(unsigned order, Vec3 d) {
    const double n2 = dot(d, d);
    if (!std::isfinite(n2) || std::abs(n2 - 1) > 1e-12)
        throw std::invalid_argument("Spherical harmonic direction must be a finite unit vector.");
    check_order(order);
    std::vector<C> y(static_cast<std::size_t>(order + 1) * (order + 1));
    const double mu = std::clamp(d.z, -1., 1.), rho = std::hypot(d.x, d.y);
    const C e = rho == 0 ? C{1, 0} : C{d.x / rho, d.y / rho};
    C diag{1 / std::sqrt(4 * kPi), 0};
    for (unsigned m = 0; m <= order; ++m) {
        if (m)
            diag *= -std::sqrt((2. * m + 1) / (2. * m)) * rho * e;
        y[lm(m, static_cast<int>(m))] = diag;
        if (m < order)
            y[lm(m + 1, static_cast<int>(m))] = std::sqrt(2. * m + 3) * mu * diag;
        for (unsigned l = m + 2; l <= order; ++l) {
            const double den = static_cast<double>(l) * l - static_cast<double>(m) * m;
            const double a = std::sqrt((4. * l * l - 1) / den),
                         b = std::sqrt((2. * l + 1) * ((l - 1.) * (l - 1.) - static_cast<double>(m) * m) /
                                       ((2. * l - 3) * den));
            y[lm(l, static_cast<int>(m))] = a * mu * y[lm(l - 1, static_cast<int>(m))] - b * y[lm(l - 2, static_cast<int>(m))];
        }
        if (m)
            for (unsigned l = m; l <= order; ++l)
                y[lm(l, -static_cast<int>(m))] = (m % 2 ? -1. : 1.) * std::conj(y[lm(l, static_cast<int>(m))]);
    }
    return y;
}

RegularRadialTable regular_radial_table(unsigned order, double x) {
    check_order(order);
    if (!std::isfinite(x) || x < 0)
        throw std::invalid_argument("Regular spherical waves require a finite nonnegative argument.");
    RegularRadialTable t;
    t.values.assign(order + 1, 0.0);
    t.derivatives.assign(order + 1, 0.0);
    if (x == 0) {
        t.values[0] = 1;
        if (order)
            t.derivatives[1] = 1. / 3;
        return t;
    }
    std::vector<double> j(order + 2);
    if (x <= 1) {
        // j_l(x) = x^l / (2l+1)!! sum_n (-x^2/2)^n / (n! (2l+3)(2l+5)...(2l+2n+1))
        double lead = 1;
        for (unsigned l = 0; l <= order + 1; ++l) {
            const double lead_over_x = l ? lead / (2. * l + 1) : 0;
            if (l)
                lead *= x / (2. * l + 1);
            double term = 1, sum = 1, weighted = 0;
            for (unsigned n = 1; n < 128; ++n) {
                term *= -x * x / (2. * n * (2. * l + 2 * n + 1));
                sum += term;
                weighted += 2 * n * term;
                if (std::abs(term) <= std::numeric_limits<double>::epsilon() * std::abs(sum))
                    break;
            }
            j[l] = lead * sum;
            if (l && l <= order)
                t.derivatives[l] = lead_over_x * (l * sum + weighted);
        }
    } else {
        double sx, cx;
        ::sincos(x, &sx, &cx); // the reference build evaluates this pair jointly
        const double j0 = sx / x, j1 = (j0 - cx) / x;
        if (x > order + 32.) {
            j[0] = j0;
            j[1] = j1;
            for (unsigned l = 1; l <= order; ++l)
                j[l + 1] = ((2. * l + 1) / x) * j[l] - j[l - 1];
        } else {
            const unsigned top = std::max(order + 48u, static_cast<unsigned>(std::ceil(x + 48 + 8 * std::cbrt(x))));
            std::vector<double> mil(top + 2);
            mil[top] = 1;
            for (unsigned l = top; l > 0; --l) {
                mil[l - 1] = ((2. * l + 1) / x) * mil[l] - mil[l + 1];
                if (std::abs(mil[l - 1]) > 1e200)
                    for (std::size_t i = l - 1; i < mil.size(); ++i)
                        mil[i] *= 1e-200;
            }
            const double scale = std::abs(j0) > std::abs(j1) ? j0 / mil[0] : j1 / mil[1];
            for (unsigned l = 0; l <= order + 1; ++l)
                j[l] = scale * mil[l];
        }
        for (unsigned l = 1; l <= order; ++l)
            t.derivatives[l] = (l / x) * j[l] - j[l + 1];
    }
    t.derivatives[0] = -j[1];
    std::copy_n(j.begin(), order + 1, t.values.begin());
    for (unsigned l = 0; l <= order; ++l)
        if (!std::isfinite(t.values[l]) || !std::isfinite(t.derivatives[l]))
            throw std::runtime_error("Regular spherical-wave recursion produced a non-finite value.");
    return t;
}

AngularTransferPlan make_angular_transfer_plan(const SphereRule &source, const SphereRule &target, unsigned L) {
    check_order(L);
    if (L > source.order || L > target.order)
        throw std::invalid_argument("Wideband transfer bandlimit exceeds its source or target sampling.");
    AngularTransferPlan p;
    p.source_azimuths = 2 * (source.order + 1);
    p.target_azimuths = 2 * (target.order + 1);
    p.bandlimit = L;
    p.source_polar = rings(source);
    p.target_polar = rings(target);
    const std::size_t count = static_cast<std::size_t>(L + 1) * (L + 2) / 2;
    p.recurrence_first.assign(count, 0.0);
    p.recurrence_second.assign(count, 0.0);
    p.diagonal_scale.assign(L + 1, 0.0);
    for (unsigned m = 0; m <= L; ++m) {
        if (m)
            p.diagonal_scale[m] = -std::sqrt((2. * m + 1) / (2. * m));
        for (unsigned l = m + 2; l <= L; ++l) {
            const double den = static_cast<double>(l) * l - static_cast<double>(m) * m;
            p.recurrence_first[tri(l, m)] = std::sqrt((4. * l * l - 1) / den);
            p.recurrence_second[tri(l, m)] =
                std::sqrt((2. * l + 1) * ((l - 1.) * (l - 1.) - static_cast<double>(m) * m) / ((2. * l - 3) * den));
        }
    }
    return p;
}

std::vector<C> resample_spectrum(const std::vector<C> &values, const AngularTransferPlan &p) {
    check_order(p.bandlimit);
    const unsigned L = p.bandlimit;
    const std::size_t count = static_cast<std::size_t>(L + 1) * (L + 2) / 2;
    if (p.source_polar.size() <= L || p.target_polar.size() <= L || p.source_azimuths != 2 * p.source_polar.size() ||
        p.target_azimuths != 2 * p.target_polar.size() || p.recurrence_first.size() != count ||
        p.recurrence_second.size() != count || p.diagonal_scale.size() != L + 1 ||
        values.size() != p.source_polar.size() * p.source_azimuths)
        throw std::invalid_argument("Wideband separable transfer has an inconsistent plan or spectrum.");
    for (const C &v : values)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::invalid_argument("Wideband spectrum contains a non-finite value.");
    for (const auto *polar : {&p.source_polar, &p.target_polar})
        for (const PolarSample &s : *polar)
            if (!std::isfinite(s.cosine) || std::abs(s.cosine) > 1 || !std::isfinite(s.weight) || s.weight <= 0)
                throw std::invalid_argument("Wideband separable transfer has an invalid polar rule.");
    for (const auto *a : {&p.recurrence_first, &p.recurrence_second, &p.diagonal_scale})
        for (double v : *a)
            if (!std::isfinite(v))
                throw std::invalid_argument("Wideband separable transfer has a non-finite recurrence.");
    const int iL = static_cast<int>(L);
    // analysis: per ring, azimuthal Fourier coefficient of order m, then the polar factor
    std::vector<C> c(static_cast<std::size_t>(L + 1) * (L + 1));
    for (std::size_t j = 0; j < p.source_polar.size(); ++j) {
        const auto P = polar_factors(p, p.source_polar[j].cosine);
        for (int m = -iL; m <= iL; ++m) {
            C f{};
            for (unsigned a = 0; a < p.source_azimuths; ++a)
                f += values[j * p.source_azimuths + a] * phase(-2 * kPi * m * a / p.source_azimuths);
            f *= 2 * kPi * p.source_polar[j].weight / p.source_azimuths;
            const unsigned am = static_cast<unsigned>(std::abs(m));
            const double sign = m < 0 && am % 2 ? -1. : 1.;
            for (unsigned l = am; l <= L; ++l)
                c[lm(l, m)] += f * (sign * P[tri(l, am)]);
        }
    }
    // synthesis on the target rings
    std::vector<C> out(p.target_polar.size() * p.target_azimuths);
    for (std::size_t j = 0; j < p.target_polar.size(); ++j) {
        const auto P = polar_factors(p, p.target_polar[j].cosine);
        for (int m = -iL; m <= iL; ++m) {
            const unsigned am = static_cast<unsigned>(std::abs(m));
            const double sign = m < 0 && am % 2 ? -1. : 1.;
            C f{};
            for (unsigned l = am; l <= L; ++l)
                f += c[lm(l, m)] * (sign * P[tri(l, am)]);
            for (unsigned a = 0; a < p.target_azimuths; ++a)
                out[j * p.target_azimuths + a] += f * phase(2 * kPi * m * a / p.target_azimuths);
        }
    }
    return out;
}

std::vector<C> resample_spectrum(const std::vector<C> &values, const SphereRule &source, const SphereRule &target,
                                 unsigned L) {
    check_values(values, source);
    return resample_spectrum(values, make_angular_transfer_plan(source, target, L));
}

std::vector<C> outgoing_to_parent(const std::vector<C> &child, const SphereRule &child_rule,
                                  const SphereRule &parent_rule, unsigned L, double k, Vec3 shift) {
    auto out = resample_spectrum(child, child_rule, parent_rule, L);
    for (std::size_t q = 0; q < out.size(); ++q)
        out[q] *= phase(k * dot(parent_rule.samples[q].direction, shift));
    return out;
}

std::vector<C> incoming_to_child(const std::vector<C> &parent, const SphereRule &parent_rule,
                                 const SphereRule &child_rule, unsigned L, double k, Vec3 shift) {
    check_values(parent, parent_rule);
    std::vector<C> shifted(parent);
    for (std::size_t q = 0; q < shifted.size(); ++q)
        shifted[q] *= phase(-k * dot(parent_rule.samples[q].direction, shift));
    return resample_spectrum(shifted, parent_rule, child_rule, L);
}

std::vector<C> multipole_to_spectrum(const std::vector<C> &multipoles, unsigned order, const SphereRule &rule) {
    check_order(order);
    if (multipoles.size() != static_cast<std::size_t>(order + 1) * (order + 1))
        throw std::invalid_argument("Wideband multipoles have an invalid size.");
    std::vector<C> c(multipoles);
    C f{4 * kPi, 0}; // 4 pi i^l
    for (unsigned l = 0; l <= order; ++l) {
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            if (!std::isfinite(c[lm(l, m)].real()) || !std::isfinite(c[lm(l, m)].imag()))
                throw std::invalid_argument("Wideband multipole is non-finite.");
            c[lm(l, m)] *= f;
        }
        f *= C{0, 1};
    }
    return evaluate(c, order, rule);
}

std::vector<C> spectrum_to_local(const std::vector<C> &incoming, const SphereRule &rule, unsigned order) {
    auto c = project(incoming, rule, order);
    C f{4 * kPi, 0}; // 4 pi (-i)^l
    for (unsigned l = 0; l <= order; ++l) {
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m)
            c[lm(l, m)] *= f;
        f *= C{0, -1};
    }
    return c;
}

std::vector<C> diagonal_translation(const SphereRule &rule, unsigned order, double k, Vec3 r) {
    check_rule(rule);
    check_order(order);
    const double dist = std::sqrt(dot(r, r)), z = k * dist;
    if (!std::isfinite(dist) || !std::isfinite(z) || k <= 0 || dist <= 0 || z == 0)
        throw std::invalid_argument("Wideband diagonal translation requires finite positive kR.");
    const Vec3 u{r.x / dist, r.y / dist, r.z / dist};
    // h_l^(2)(z): h_0 = (sin z + i cos z)/z, h_1 = (1/z + i) h_0, upward three-term recurrence
    std::vector<C> h(order + 1);
    double sz, cz;
    ::sincos(z, &sz, &cz); // the reference build evaluates this pair jointly
    h[0] = C{sz, cz} / z;
    if (order)
        h[1] = C{1 / z, 1} * h[0];
    for (unsigned l = 1; l < order; ++l)
        h[l + 1] = ((2. * l + 1) / z) * h[l] - h[l - 1];
    for (const C &v : h)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag()))
            throw std::runtime_error("Unstable diagonal translation: use the low-frequency representation.");
    std::vector<C> out(rule.samples.size());
    const C pre{0, -k / (16 * kPi * kPi)};
    for (std::size_t q = 0; q < out.size(); ++q) {
        const double mu = std::clamp(dot(rule.samples[q].direction, u), -1., 1.);
        double p0 = 0, p1 = 1;
        C ph{1, 0}, sum{};
        for (unsigned l = 0; l <= order; ++l) {
            sum += (2. * l + 1) * ph * h[l] * p1; // (2l+1) (-i)^l h_l P_l(mu)
            const double p2 = ((2. * l + 1) * mu * p1 - l * p0) / (l + 1.);
            p0 = p1;
            p1 = p2;
            ph *= C{0, -1};
        }
        out[q] = pre * sum;
        if (!std::isfinite(out[q].real()) || !std::isfinite(out[q].imag()))
            throw std::runtime_error("Wideband diagonal translation is non-finite.");
    }
    return out;
}

} // namespace hsbem
