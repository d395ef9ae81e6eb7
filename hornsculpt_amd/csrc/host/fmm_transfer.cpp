#include "fmm_transfer.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <string>

namespace hsbem {
namespace {

using C = std::complex<double>;
constexpr unsigned kOrderLimit = 1024;
constexpr double kPi = std::numbers::pi;

void require_order(unsigned order, const char *what) {
    if (order > kOrderLimit)
        throw std::invalid_argument(std::string(what) + " order " + std::to_string(order) +
                                    " is above the supported limit of 1024");
}

bool finite(C v) { return std::isfinite(v.real()) && std::isfinite(v.imag()); }

std::size_t tri(unsigned l, unsigned m) { return static_cast<std::size_t>(l) * (l + 1) / 2 + m; }
std::size_t lm(unsigned l, int m) { return static_cast<std::size_t>(l) * l + l + m; }

double dot3(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

} // namespace

// n-point Gauss-Legendre rule on [-1, 1], nodes descending. Newton's method on P_n
// from Tricomi's asymptotic estimate of the i-th root; the nodes of the upper half
// are mirrored, so the rule is symmetric to the last bit.
void gauss_legendre(unsigned n, std::vector<double> &x, std::vector<double> &w) {
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    auto legendre_pair = [n](double t, double &pn, double &pn1) { // P_n(t), P_{n-1}(t)
        double prev = 1, cur = t;
        for (unsigned l = 2; l <= n; ++l) {
            const double next = ((2.0 * l - 1) * t * cur - (l - 1.0) * prev) / l;
            prev = cur;
            cur = next;
        }
        pn = n == 0 ? 1 : cur;
        pn1 = prev;
    };
    const double nn = n;
    for (unsigned i = 0; i < (n + 1) / 2; ++i) {
        const bool middle = 2 * i + 1 == n;
        double t = middle ? 0.0
                          : (1 - (nn - 1) / (8 * nn * nn * nn)) * std::cos(kPi * (4.0 * i + 3) / (4.0 * nn + 2));
        double pn = 0, pn1 = 0;
        for (int it = 0; !middle; ++it) {
            if (it == 60)
                throw std::runtime_error("Gauss-Legendre node iteration failed to converge");
            legendre_pair(t, pn, pn1);
            const double step = pn / (nn * (t * pn - pn1) / (t * t - 1));
            t -= step;
            if (std::abs(step) <= 2 * std::numeric_limits<double>::epsilon())
                break;
        }
        legendre_pair(t, pn, pn1);
        const double slope = nn * (t * pn - pn1) / (t * t - 1);
        const double weight = 2 / ((1 - t * t) * slope * slope);
        x[i] = t;
        x[n - 1 - i] = -t;
        w[i] = w[n - 1 - i] = weight;
    }
}

namespace {

// The polar rings of a product rule: cos(theta) and the (shared) sample weight.
struct Rings {
    unsigned count{}, azimuths{};
    std::vector<double> mu, weight;
};

Rings rings_of(const SphereRule &rule) {
    require_order(rule.order, "sphere rule");
    Rings r{rule.order + 1, 2 * (rule.order + 1), {}, {}};
    if (rule.samples.size() != static_cast<std::size_t>(r.count) * r.azimuths)
        throw std::invalid_argument("sphere rule sample count does not match its order");
    for (unsigned j = 0; j < r.count; ++j) {
        const SphereSample *ring = rule.samples.data() + static_cast<std::size_t>(j) * r.azimuths;
        const double mu = ring[0].direction.z, rho = std::sqrt(std::max(0.0, (1 - mu) * (1 + mu)));
        for (unsigned a = 0; a < r.azimuths; ++a) {
            const double phi = 2 * kPi * a / r.azimuths;
            const SphereSample &s = ring[a];
            const bool ok = s.weight > 0 && std::isfinite(s.weight) && std::abs(s.direction.z - mu) <= 1e-13 &&
                            std::abs(s.direction.x - rho * std::cos(phi)) <= 1e-13 &&
                            std::abs(s.direction.y - rho * std::sin(phi)) <= 1e-13 &&
                            std::abs(s.weight - ring[0].weight) <= 1e-13 * ring[0].weight;
            if (!ok)
                throw std::invalid_argument("sphere rule is not a polar-major product rule with equispaced azimuths");
        }
        r.mu.push_back(mu);
        r.weight.push_back(ring[0].weight);
    }
    return r;
}

void require_samples(const std::vector<C> &v, const Rings &r) {
    if (v.size() != static_cast<std::size_t>(r.count) * r.azimuths)
        throw std::invalid_argument("signature sample count does not match its sphere rule");
    for (const C &x : v)
        if (!finite(x))
            throw std::invalid_argument("signature samples must be finite");
}

// e^{-2 pi i t / n}, t = 0..n-1.
std::vector<C> twiddles(unsigned n) {
    std::vector<C> t(n);
    for (unsigned i = 0; i < n; ++i) {
        const double a = -2 * kPi * i / n;
        t[i] = {std::cos(a), std::sin(a)};
    }
    return t;
}

// Twiddle index step of azimuthal order m on n azimuths.
std::size_t step_of(int m, unsigned n) {
    const long long s = m % static_cast<long long>(n);
    return static_cast<std::size_t>(s < 0 ? s + n : s);
}

// (-1)^m for negative m with |m| odd (Y_{l,-|m|} = (-1)^m conj(Y_l|m|)).
double parity(int m) { return m < 0 && (-m) % 2 ? -1.0 : 1.0; }

// c_lm = sum_q w_q v_q conj(Y_lm(s_q)), l <= L: per ring the azimuthal DFT, then the
// Legendre sum.
std::vector<C> analyse(const std::vector<C> &v, const Rings &r, unsigned L) {
    std::vector<C> c(static_cast<std::size_t>(L + 1) * (L + 1));
    const std::vector<C> tw = twiddles(r.azimuths);
    const int iL = static_cast<int>(L);
    for (unsigned j = 0; j < r.count; ++j) {
        const std::vector<double> P = normalized_legendre(L, r.mu[j]);
        const C *ring = v.data() + static_cast<std::size_t>(j) * r.azimuths;
        for (int m = -iL; m <= iL; ++m) {
            const std::size_t step = step_of(m, r.azimuths);
            C f{};
            for (std::size_t a = 0, idx = 0; a < r.azimuths; ++a) {
                f += ring[a] * tw[idx];
                idx += step;
                if (idx >= r.azimuths)
                    idx -= r.azimuths;
            }
            f *= r.weight[j] * parity(m);
            const unsigned am = static_cast<unsigned>(std::abs(m));
            for (unsigned l = am; l <= L; ++l)
                c[lm(l, m)] += f * P[tri(l, am)];
        }
    }
    return c;
}

// sum_lm c_lm Y_lm(s_q) on every sample of the rings, l <= L.
std::vector<C> synthesise(const std::vector<C> &c, unsigned L, const Rings &r) {
    std::vector<C> out(static_cast<std::size_t>(r.count) * r.azimuths);
    const std::vector<C> tw = twiddles(r.azimuths);
    const int iL = static_cast<int>(L);
    for (unsigned j = 0; j < r.count; ++j) {
        const std::vector<double> P = normalized_legendre(L, r.mu[j]);
        C *ring = out.data() + static_cast<std::size_t>(j) * r.azimuths;
        for (int m = -iL; m <= iL; ++m) {
            const unsigned am = static_cast<unsigned>(std::abs(m));
            C g{};
            for (unsigned l = am; l <= L; ++l)
                g += c[lm(l, m)] * P[tri(l, am)];
            g *= parity(m);
            const std::size_t step = step_of(m, r.azimuths);
            for (std::size_t a = 0, idx = 0; a < r.azimuths; ++a) {
                ring[a] += g * std::conj(tw[idx]);
                idx += step;
                if (idx >= r.azimuths)
                    idx -= r.azimuths;
            }
        }
    }
    return out;
}

void multiply_phases(std::vector<C> &v, const SphereRule &rule, double scale, Vec3 shift) {
    for (std::size_t q = 0; q < v.size(); ++q) {
        const double angle = scale * dot3(rule.samples[q].direction, shift);
        if (!std::isfinite(angle))
            throw std::invalid_argument("plane-wave shift phase is not finite");
        v[q] *= std::polar(1.0, angle);
    }
}

} // namespace

SphereRule make_sphere_rule(unsigned order) {
    require_order(order, "sphere rule");
    const unsigned n = order + 1, na = 2 * n;
    std::vector<double> mu, w;
    gauss_legendre(n, mu, w);
    std::vector<double> cs(na), sn(na);
    for (unsigned a = 0; a < na; ++a) {
        const double phi = 2 * kPi * a / na;
        cs[a] = std::cos(phi);
        sn[a] = std::sin(phi);
    }
    SphereRule rule{order, {}};
    rule.samples.reserve(static_cast<std::size_t>(n) * na);
    for (unsigned j = 0; j < n; ++j) {
        const double rho = std::sqrt((1 - mu[j]) * (1 + mu[j])), weight = w[j] * (2 * kPi / na);
        for (unsigned a = 0; a < na; ++a)
            rule.samples.push_back({{rho * cs[a], rho * sn[a], mu[j]}, weight});
    }
    return rule;
}

std::vector<double> normalized_legendre(unsigned order, double mu) {
    require_order(order, "Legendre table");
    if (!(std::abs(mu) <= 1))
        throw std::invalid_argument("Legendre argument must lie in [-1, 1]");
    std::vector<double> p(tri(order, order) + 1);
    const double rho = std::sqrt((1 - mu) * (1 + mu));
    double sectoral = 1 / std::sqrt(4 * kPi);
    for (unsigned m = 0; m <= order; ++m) {
        if (m)
            sectoral *= -std::sqrt((2.0 * m + 1) / (2.0 * m)) * rho;
        p[tri(m, m)] = sectoral;
        if (m == order)
            break;
        p[tri(m + 1, m)] = std::sqrt(2.0 * m + 3) * mu * sectoral;
        const double mm = static_cast<double>(m) * m;
        for (unsigned l = m + 2; l <= order; ++l) {
            const double ll = static_cast<double>(l) * l, span = ll - mm;
            const double a = std::sqrt((4 * ll - 1) / span);
            const double b = std::sqrt(((l - 1.0) * (l - 1.0) - mm) * (2.0 * l + 1) / (span * (2.0 * l - 3)));
            p[tri(l, m)] = a * mu * p[tri(l - 1, m)] - b * p[tri(l - 2, m)];
        }
    }
    return p;
}

std::vector<C> spherical_harmonics(unsigned order, Vec3 d) {
    const double n2 = dot3(d, d);
    if (!std::isfinite(n2) || std::abs(n2 - 1) > 1e-12)
        throw std::invalid_argument("spherical harmonics need a finite unit direction");
    const std::vector<double> P = normalized_legendre(order, std::clamp(d.z, -1.0, 1.0));
    const double rho = std::hypot(d.x, d.y);
    const C e = rho > 0 ? C{d.x / rho, d.y / rho} : C{1, 0};
    std::vector<C> y(static_cast<std::size_t>(order + 1) * (order + 1));
    C em{1, 0}; // e^{i m phi}
    for (unsigned m = 0; m <= order; ++m) {
        const double sign = m % 2 ? -1.0 : 1.0;
        for (unsigned l = m; l <= order; ++l) {
            const C v = P[tri(l, m)] * em;
            y[lm(l, static_cast<int>(m))] = v;
            if (m)
                y[lm(l, -static_cast<int>(m))] = sign * std::conj(v);
        }
        em *= e;
    }
    return y;
}

RegularRadialTable regular_radial_table(unsigned order, double x) {
    require_order(order, "spherical Bessel table");
    if (!std::isfinite(x) || x < 0)
        throw std::invalid_argument("spherical Bessel argument must be finite and nonnegative");
    RegularRadialTable t{std::vector<double>(order + 1, 0.0), std::vector<double>(order + 1, 0.0)};
    if (x == 0) {
        t.values[0] = 1;
        if (order)
            t.derivatives[1] = 1.0 / 3;
        return t;
    }
    std::vector<double> j(order + 2);
    const double s = std::sin(x), c = std::cos(x), j0 = s / x, j1 = (s / x - c) / x;
    if (x < 1e-8) {
        // two terms of the power series are exact in double precision here
        double lead = 1; // x^l / (2l+1)!!
        for (unsigned l = 0; l <= order + 1; ++l) {
            if (l)
                lead *= x / (2.0 * l + 1);
            j[l] = lead * (1 - x * x / (2 * (2.0 * l + 3)));
        }
    } else if (x > order + 32.0) {
        // every l <= order + 1 lies below x: the upward recurrence is stable
        j[0] = j0;
        j[1] = j1;
        for (unsigned l = 1; l <= order; ++l)
            j[l + 1] = (2.0 * l + 1) / x * j[l] - j[l - 1];
    } else {
        // Miller: the downward recurrence converges to the minimal solution j_l; its
        // scale and sign follow from sum_l (2l+1) j_l^2 = 1 and the larger of j_0, j_1.
        const double reach = std::max(static_cast<double>(order + 1), x);
        const unsigned top = static_cast<unsigned>(reach + 40 + 10 * std::cbrt(reach));
        std::vector<double> f(top + 2, 0.0);
        f[top] = 1;
        for (unsigned l = top; l > 0; --l) {
            f[l - 1] = (2.0 * l + 1) / x * f[l] - f[l + 1];
            if (std::abs(f[l - 1]) > 1e100)
                for (unsigned i = l - 1; i <= top; ++i)
                    f[i] *= 1e-100;
        }
        double norm = 0;
        for (unsigned l = 0; l <= top; ++l)
            norm += (2.0 * l + 1) * f[l] * f[l];
        double scale = 1 / std::sqrt(norm);
        if ((std::abs(j0) >= std::abs(j1) ? j0 * f[0] : j1 * f[1]) < 0)
            scale = -scale;
        for (unsigned l = 0; l <= order + 1; ++l)
            j[l] = scale * f[l];
    }
    t.derivatives[0] = -j[1];
    for (unsigned l = 1; l <= order; ++l)
        t.derivatives[l] = j[l - 1] - (l + 1.0) / x * j[l];
    std::copy_n(j.begin(), order + 1, t.values.begin());
    for (unsigned l = 0; l <= order; ++l)
        if (!std::isfinite(t.values[l]) || !std::isfinite(t.derivatives[l]))
            throw std::runtime_error("spherical Bessel table is not finite");
    return t;
}

std::vector<C> resample_spectrum(const std::vector<C> &values, const SphereRule &source, const SphereRule &target,
                                 unsigned L) {
    require_order(L, "resampling bandlimit");
    const Rings from = rings_of(source), to = rings_of(target);
    if (L > source.order || L > target.order)
        throw std::invalid_argument("resampling bandlimit exceeds the source or target rule order");
    require_samples(values, from);
    return synthesise(analyse(values, from, L), L, to);
}

std::vector<C> outgoing_to_parent(const std::vector<C> &child, const SphereRule &child_rule,
                                  const SphereRule &parent_rule, unsigned L, double k, Vec3 shift) {
    std::vector<C> out = resample_spectrum(child, child_rule, parent_rule, L);
    multiply_phases(out, parent_rule, k, shift);
    return out;
}

std::vector<C> incoming_to_child(const std::vector<C> &parent, const SphereRule &parent_rule,
                                 const SphereRule &child_rule, unsigned L, double k, Vec3 shift) {
    require_samples(parent, rings_of(parent_rule));
    std::vector<C> shifted(parent);
    multiply_phases(shifted, parent_rule, -k, shift);
    return resample_spectrum(shifted, parent_rule, child_rule, L);
}

std::vector<C> multipole_to_spectrum(const std::vector<C> &multipoles, unsigned order, const SphereRule &rule) {
    require_order(order, "multipole");
    if (multipoles.size() != static_cast<std::size_t>(order + 1) * (order + 1))
        throw std::invalid_argument("multipole coefficient count does not match its order");
    std::vector<C> c(multipoles.size());
    C il{4 * kPi, 0}; // 4 pi i^l
    for (unsigned l = 0; l <= order; ++l, il *= C{0, 1})
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            if (!finite(multipoles[lm(l, m)]))
                throw std::invalid_argument("multipole coefficients must be finite");
            c[lm(l, m)] = il * multipoles[lm(l, m)];
        }
    return synthesise(c, order, rings_of(rule));
}

std::vector<C> spectrum_to_local(const std::vector<C> &incoming, const SphereRule &rule, unsigned order) {
    const Rings r = rings_of(rule);
    require_order(order, "local expansion");
    if (order > rule.order)
        throw std::invalid_argument("local expansion order exceeds the order its sphere rule resolves");
    require_samples(incoming, r);
    std::vector<C> c = analyse(incoming, r, order);
    C il{4 * kPi, 0}; // 4 pi (-i)^l
    for (unsigned l = 0; l <= order; ++l, il *= C{0, -1})
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m)
            c[lm(l, m)] *= il;
    return c;
}

std::vector<C> diagonal_translation(const SphereRule &rule, unsigned order, double k, Vec3 R) {
    rings_of(rule);
    require_order(order, "diagonal translation");
    const double dist = std::sqrt(dot3(R, R)), z = k * dist;
    if (!(k > 0) || !(dist > 0) || !std::isfinite(z) || z == 0)
        throw std::invalid_argument("diagonal translation needs k > 0 and a finite nonzero separation");
    const Vec3 axis{R.x / dist, R.y / dist, R.z / dist};
    // c_l = (2l+1) (-i)^l h_l^(2)(z); h_l by the upward recurrence (stable for h).
    std::vector<C> coef(order + 1);
    {
        const double s = std::sin(z), c = std::cos(z);
        C h_prev{s / z, c / z}; // h_0 = j_0 - i y_0
        C h = order ? C{s / z / z - c / z, c / z / z + s / z} : C{}; // h_1
        C phase{1, 0};
        for (unsigned l = 0; l <= order; ++l, phase *= C{0, -1}) {
            const C hl = l == 0 ? h_prev : h;
            if (!finite(hl))
                throw std::runtime_error("diagonal translation overflowed: the separation is too small for the "
                                         "expansion order (use the spherical representation)");
            coef[l] = (2.0 * l + 1) * phase * hl;
            if (l >= 1) {
                const C next = (2.0 * l + 1) / z * h - h_prev;
                h_prev = h;
                h = next;
            }
        }
    }
    const C pre{0, -k / (16 * kPi * kPi)};
    std::vector<C> out(rule.samples.size());
    for (std::size_t q = 0; q < out.size(); ++q) {
        const double mu = std::clamp(dot3(rule.samples[q].direction, axis), -1.0, 1.0);
        // Clenshaw: b_l = c_l + alpha_l b_{l+1} + beta_{l+1} b_{l+2},
        // alpha_l = (2l+1) mu / (l+1), beta_l = -l / (l+1); the sum is b_0.
        C b1{}, b2{};
        for (int l = static_cast<int>(order); l >= 0; --l) {
            const double alpha = (2.0 * l + 1) * mu / (l + 1.0), beta = -(l + 1.0) / (l + 2.0);
            const C b0 = coef[l] + alpha * b1 + beta * b2;
            b2 = b1;
            b1 = b0;
        }
        out[q] = pre * b1;
        if (!finite(out[q]))
            throw std::runtime_error("diagonal translation is not finite");
    }
    return out;
}

} // namespace hsbem
