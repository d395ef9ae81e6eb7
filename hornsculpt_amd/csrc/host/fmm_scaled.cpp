#include "fmm_scaled.hpp"

#include "fmm_levels.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <stdexcept>
#include <string>

namespace hsbem {
namespace {

using C = std::complex<double>;
constexpr double kPi = std::numbers::pi;

std::size_t lm(unsigned l, int m) { return static_cast<std::size_t>(l) * l + l + m; }

std::size_t waves_for(unsigned order) {
    if (order > kMaximumSphericalOrder)
        throw std::invalid_argument("scaled expansion order " + std::to_string(order) +
                                    " exceeds the supported maximum of " + std::to_string(kMaximumSphericalOrder));
    return static_cast<std::size_t>(order + 1) * (order + 1);
}

void require_scale(double k, double a) {
    if (!std::isfinite(k) || k < 0 || !std::isfinite(a) || a <= 0 || k * a > kMaximumSphericalElectricalRadius)
        throw std::invalid_argument("scaled waves need k >= 0, a finite scale a > 0 and k a within the supported range");
}

bool finite(C v) { return std::isfinite(v.real()) && std::isfinite(v.imag()); }

void require_expansion(const ScaledExpansion &e, double k) {
    require_scale(k, e.radius);
    if (e.coefficients.size() != waves_for(e.order))
        throw std::invalid_argument("scaled expansion coefficient count does not match its order");
    for (const C &v : e.coefficients)
        if (!finite(v))
            throw std::invalid_argument("scaled expansion coefficients must be finite");
}

double axis(Vec3 v, unsigned i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }

// Regular solid harmonics S_lm(p) = |p|^l Y_lm(p/|p|) (orthonormal Y, Condon-Shortley
// phase) and their gradients. Values by the Cartesian form of the Legendre
// recurrences (sectoral step with x + iy, first off-diagonal with z, three-term step
// with z and |p|^2); gradients from the degree-lowering ladder identities
//   d/dz S_lm = c_l sqrt((l-m)(l+m)) S_{l-1,m},
//   (d/dx + i d/dy) S_lm = c_l sqrt((l-m)(l-m-1)) S_{l-1,m+1},
//   (d/dx - i d/dy) S_lm = -c_l sqrt((l+m)(l+m-1)) S_{l-1,m-1},   c_l = sqrt((2l+1)/(2l-1)),
// which hold for every m and need no division by |p| (finite at the origin).
struct SolidHarmonics {
    std::vector<C> value;
    std::vector<std::array<C, 3>> gradient;
};

SolidHarmonics solid_harmonics(unsigned order, Vec3 p) {
    const std::size_t n = waves_for(order);
    SolidHarmonics s{std::vector<C>(n), std::vector<std::array<C, 3>>(n)};
    const C w{p.x, p.y};
    const double z = p.z, r2 = p.x * p.x + p.y * p.y + p.z * p.z;
    C sectoral{1 / std::sqrt(4 * kPi), 0};
    for (unsigned m = 0; m <= order; ++m) {
        const int im = static_cast<int>(m);
        if (m)
            sectoral *= -std::sqrt((2.0 * m + 1) / (2.0 * m)) * w;
        s.value[lm(m, im)] = sectoral;
        if (m < order)
            s.value[lm(m + 1, im)] = std::sqrt(2.0 * m + 3) * z * sectoral;
        const double mm = static_cast<double>(m) * m;
        for (unsigned l = m + 2; l <= order; ++l) {
            const double ll = static_cast<double>(l) * l, span = ll - mm;
            const double a = std::sqrt((4 * ll - 1) / span);
            const double b = std::sqrt(((l - 1.0) * (l - 1.0) - mm) * (2.0 * l + 1) / (span * (2.0 * l - 3)));
            s.value[lm(l, im)] = a * z * s.value[lm(l - 1, im)] - b * r2 * s.value[lm(l - 2, im)];
        }
        if (m) {
            const double sign = m % 2 ? -1.0 : 1.0;
            for (unsigned l = m; l <= order; ++l)
                s.value[lm(l, -im)] = sign * std::conj(s.value[lm(l, im)]);
        }
    }
    const C i{0, 1};
    for (unsigned l = 1; l <= order; ++l) {
        const double cl = std::sqrt((2.0 * l + 1) / (2.0 * l - 1));
        const int il = static_cast<int>(l);
        for (int m = -il; m <= il; ++m) {
            auto below = [&](int mm) { return std::abs(mm) <= il - 1 ? s.value[lm(l - 1, mm)] : C{}; };
            const C dz = cl * std::sqrt(static_cast<double>(il - m) * (il + m)) * below(m);
            const C up = cl * std::sqrt(std::max(0.0, static_cast<double>(il - m) * (il - m - 1))) * below(m + 1);
            const C down = -cl * std::sqrt(std::max(0.0, static_cast<double>(il + m) * (il + m - 1))) * below(m - 1);
            s.gradient[lm(l, m)] = {0.5 * (up + down), -0.5 * i * (up - down), dz};
        }
    }
    return s;
}

// Jhat_l(z) = (2l+1)!! j_l(z) / z^l, l = 0..order. In this normalisation the Bessel
// recurrence reads Jhat_{l-1} = Jhat_l - z^2 / ((2l+1)(2l+3)) Jhat_{l+1}; run downward
// (Miller) its solution converges to Jhat, whose scale follows from
// Jhat_0 = sin z / z or Jhat_1 = 3 j_1(z) / z, whichever j is larger in magnitude.
std::vector<double> regular_profile(unsigned order, double z) {
    if (!std::isfinite(z) || z < 0 || z > 2 * kMaximumSphericalElectricalRadius)
        throw std::invalid_argument("scaled regular waves: k r is outside the supported range");
    std::vector<double> out(order + 1, 1.0);
    if (z == 0)
        return out;
    const double reach = std::max(static_cast<double>(order), z);
    const unsigned top = order + 2 + static_cast<unsigned>(z + 30 + 10 * std::cbrt(reach));
    std::vector<double> g(top + 2, 0.0);
    g[top] = 1;
    for (unsigned l = top; l > 0; --l)
        g[l - 1] = g[l] - z * z / ((2.0 * l + 1) * (2.0 * l + 3)) * g[l + 1];
    const double j0 = std::sin(z) / z, j1 = (j0 - std::cos(z)) / z;
    const double scale = std::abs(j0) >= std::abs(j1) ? j0 / g[0] : 3 * j1 / z / g[1];
    for (unsigned l = 0; l <= order; ++l)
        out[l] = scale * g[l];
    return out;
}

// f_l(r) = (a/r)^(l+1) Hhat_l(kr) = -i (ka)^(l+1) h_l^(2)(kr) / (2l-1)!!, l = 0..order,
// with Hhat_l -> 1 as kr -> 0, by the upward recurrence of h_l (stable for h):
// f_0 = t e^{-ikr}, f_1 = t (1 + ikr) f_0, f_{l+1} = t f_l - (ka)^2 / ((2l-1)(2l+1)) f_{l-1},
// t = a/r; and f_l' from the degree-lowering identity h_l' = h_{l-1} - (l+1) h_l / (kr):
// f_0' = -f_1 / a, f_l' = k^2 a / (2l-1) f_{l-1} - (l+1) f_l / r.
void outgoing_profile(unsigned order, double k, double a, double r, std::vector<C> &f, std::vector<C> &df) {
    const double t = a / r, ka = k * a;
    f.assign(order + 2, C{});
    df.assign(order + 1, C{});
    f[0] = t * std::polar(1.0, -k * r);
    f[1] = t * C{1, k * r} * f[0];
    for (unsigned l = 1; l <= order; ++l)
        f[l + 1] = t * f[l] - ka * ka / ((2.0 * l - 1) * (2.0 * l + 1)) * f[l - 1];
    df[0] = -f[1] / a;
    for (unsigned l = 1; l <= order; ++l)
        df[l] = k * k * a / (2.0 * l - 1) * f[l - 1] - (l + 1.0) / r * f[l];
}

WaveValue combine(const std::vector<WaveValue> &basis, const std::vector<C> &coefficients) {
    WaveValue out;
    for (std::size_t j = 0; j < basis.size(); ++j) {
        out.value += coefficients[j] * basis[j].value;
        for (unsigned d = 0; d < 3; ++d)
            out.gradient[d] += coefficients[j] * basis[j].gradient[d];
    }
    return out;
}

// Multiplies degree l of an expansion by (ka)^l / (2l+1)!! (the scaled <-> unscaled
// regular normalisation).
void scale_degrees(std::vector<C> &c, unsigned order, double ka) {
    double f = 1;
    for (unsigned l = 0; l <= order; ++l) {
        if (l)
            f *= ka / (2.0 * l + 1);
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m)
            c[lm(l, m)] *= f;
    }
}

} // namespace

std::vector<WaveValue> scaled_regular_waves(unsigned order, double k, double a, Vec3 p) {
    require_scale(k, a);
    const SolidHarmonics s = solid_harmonics(order, vscale(p, 1 / a));
    const std::vector<double> J = regular_profile(order + 1, k * std::hypot(p.x, p.y, p.z));
    std::vector<WaveValue> out(s.value.size());
    // grad [Jhat_l(k|p|) S_lm(p/a)] = Jhat_l grad S / a - k^2 p Jhat_{l+1} / (2l+3) S
    for (unsigned l = 0; l <= order; ++l) {
        const double radial = -k * k * J[l + 1] / (2.0 * l + 3);
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            const std::size_t j = lm(l, m);
            out[j].value = J[l] * s.value[j];
            for (unsigned d = 0; d < 3; ++d)
                out[j].gradient[d] = J[l] / a * s.gradient[j][d] + radial * axis(p, d) * s.value[j];
        }
    }
    return out;
}

std::vector<WaveValue> scaled_outgoing_waves(unsigned order, double k, double a, Vec3 p) {
    require_scale(k, a);
    const double r = std::hypot(p.x, p.y, p.z);
    if (!std::isfinite(r) || r <= 0)
        throw std::invalid_argument("scaled outgoing waves are singular at their centre: need a finite point r > 0");
    const Vec3 u = vscale(p, 1 / r);
    // On the unit sphere S_lm(u) = Y_lm(u); grad_p Y_lm(p/|p|) = (grad S_lm(u) - l u S_lm(u)) / r.
    const SolidHarmonics s = solid_harmonics(order, u);
    std::vector<C> f, df;
    outgoing_profile(order, k, a, r, f, df);
    std::vector<WaveValue> out(s.value.size());
    for (unsigned l = 0; l <= order; ++l) {
        const double norm = 1 / (a * (2.0 * l + 1));
        const C value = norm * f[l], radial = norm * df[l], angular = value / r;
        for (int m = -static_cast<int>(l); m <= static_cast<int>(l); ++m) {
            const std::size_t j = lm(l, m);
            out[j].value = value * s.value[j];
            for (unsigned d = 0; d < 3; ++d)
                out[j].gradient[d] = radial * axis(u, d) * s.value[j] +
                                     angular * (s.gradient[j][d] - static_cast<double>(l) * axis(u, d) * s.value[j]);
        }
    }
    return out;
}

WaveValue evaluate_scaled_multipole(const ScaledExpansion &e, double k, Vec3 p) {
    require_expansion(e, k);
    return combine(scaled_outgoing_waves(e.order, k, e.radius, p), e.coefficients);
}

WaveValue evaluate_scaled_local(const ScaledExpansion &e, double k, Vec3 p) {
    require_expansion(e, k);
    return combine(scaled_regular_waves(e.order, k, e.radius, p), e.coefficients);
}

std::vector<C> scaled_multipole_to_spectrum(const ScaledExpansion &e, double k, const SphereRule &rule) {
    require_expansion(e, k);
    std::vector<C> c = e.coefficients;
    scale_degrees(c, e.order, k * e.radius);
    return multipole_to_spectrum(c, e.order, rule);
}

ScaledExpansion spectrum_to_scaled_local(const std::vector<C> &incoming, const SphereRule &rule, unsigned order,
                                         double k, double a) {
    require_scale(k, a);
    waves_for(order);
    ScaledExpansion out{order, a, spectrum_to_local(incoming, rule, order)};
    scale_degrees(out.coefficients, order, k * a);
    return out;
}

AxialPlan make_axial_plan(AxialKind kind, unsigned ps, unsigned pt, double k, double as, double at, double d,
                          unsigned polar_count, double projection_radius) {
    require_scale(k, as);
    require_scale(k, at);
    waves_for(ps);
    waves_for(pt);
    if (kind != AxialKind::multipole_to_parent && kind != AxialKind::multipole_to_local &&
        kind != AxialKind::local_to_child)
        throw std::invalid_argument("unknown coaxial translation kind");
    if (!std::isfinite(d) || d < 0)
        throw std::invalid_argument("coaxial translation distance must be finite and nonnegative");
    if (polar_count < 2 * std::max(ps, pt) + 1 || polar_count > 1024)
        throw std::invalid_argument("coaxial plan needs between 2 max(p_s, p_t) + 1 and 1024 polar nodes");
    if (!std::isfinite(projection_radius) || projection_radius < 0)
        throw std::invalid_argument("coaxial projection radius must be finite and nonnegative");
    // Sphere on which the translated field is projected: around the child for the
    // multipole shift (just outside its source ball), around the target centre otherwise.
    double rho = projection_radius;
    if (rho == 0)
        rho = kind == AxialKind::multipole_to_parent ? 1.01 * as
              : kind == AxialKind::multipole_to_local ? std::min(at, 0.5 * d)
                                                      : at;
    if (!(rho > 0) || (kind == AxialKind::multipole_to_local && rho >= d))
        throw std::invalid_argument("coaxial projection sphere must not reach the source's singular centre");

    AxialPlan plan{kind, ps, pt, polar_count, {0}, {}, rho};
    const unsigned mtop = std::min(ps, pt);
    for (unsigned m = 0; m <= mtop; ++m)
        plan.block_offsets.push_back(plan.block_offsets.back() + static_cast<std::size_t>(pt - m + 1) * (ps - m + 1));
    plan.coefficients.assign(plan.block_offsets.back(), C{});

    std::vector<double> nodes, weights;
    gauss_legendre(polar_count, nodes, weights);
    const std::vector<double> J = regular_profile(pt + 1, k * rho);
    // Rows of the Cauchy projection: ((a_t / rho)^l Jhat_l, (a_t / rho)^l rho d/drho[rho^l Jhat_l] / rho^l)
    // normalised by their squared length, so that value and radial derivative together
    // determine every degree even where Jhat_l(k rho) vanishes.
    std::vector<double> value_row(pt + 1), slope_row(pt + 1);
    for (unsigned l = 0; l <= pt; ++l) {
        const double slope = l * J[l] - (k * rho) * (k * rho) / (2.0 * l + 3) * J[l + 1];
        const double length2 = J[l] * J[l] + slope * slope;
        if (!(length2 > 0) || !std::isfinite(length2))
            throw std::runtime_error("coaxial projection: radial profile vanishes on the projection sphere");
        const double lift = std::pow(at / rho, static_cast<double>(l));
        value_row[l] = lift * J[l] / length2;
        slope_row[l] = lift * slope / length2;
    }
    std::vector<C> alpha(waves_for(pt)), beta(alpha.size()), source_value, source_slope;
    for (unsigned q = 0; q < polar_count; ++q) {
        // azimuth 0 of the ring; the coaxial azimuthal integral is exact: 2 pi
        const double mu = nodes[q], w = 2 * kPi * weights[q];
        const Vec3 u{std::sqrt((1 - mu) * (1 + mu)), 0.0, mu};
        const Vec3 on_sphere = vscale(u, rho);
        const Vec3 shifted = vadd({0.0, 0.0, d}, on_sphere);
        const std::vector<WaveValue> src =
            kind == AxialKind::multipole_to_parent ? scaled_outgoing_waves(ps, k, as, on_sphere)
            : kind == AxialKind::multipole_to_local ? scaled_outgoing_waves(ps, k, as, shifted)
                                                    : scaled_regular_waves(ps, k, as, shifted);
        source_value.assign(src.size(), C{});
        source_slope.assign(src.size(), C{});
        for (std::size_t j = 0; j < src.size(); ++j) {
            source_value[j] = src[j].value;
            source_slope[j] = src[j].gradient[0] * u.x + src[j].gradient[2] * u.z;
        }
        // Target weights: plan row (l, m) += alpha_lm * value + beta_lm * radial slope.
        if (kind == AxialKind::multipole_to_parent) {
            // Green's identity over the sphere: rho^2 (u conj(d R_lm) - conj(R_lm) d u)
            const std::vector<WaveValue> tgt = scaled_regular_waves(pt, k, at, shifted);
            for (unsigned l = 0; l <= pt; ++l)
                for (unsigned m = 0; m <= std::min(l, mtop); ++m) {
                    const std::size_t j = lm(l, static_cast<int>(m));
                    const C tgt_slope = tgt[j].gradient[0] * u.x + tgt[j].gradient[2] * u.z;
                    alpha[j] = w * rho * rho * std::conj(tgt_slope);
                    beta[j] = -w * rho * rho * std::conj(tgt[j].value);
                }
        } else {
            const std::vector<C> Y = spherical_harmonics(pt, u);
            for (unsigned l = 0; l <= pt; ++l)
                for (unsigned m = 0; m <= std::min(l, mtop); ++m) {
                    const std::size_t j = lm(l, static_cast<int>(m));
                    const C y = w * std::conj(Y[j]);
                    alpha[j] = y * value_row[l];
                    beta[j] = y * (slope_row[l] * rho);
                }
        }
        for (unsigned m = 0; m <= mtop; ++m) {
            C *block = plan.coefficients.data() + plan.block_offsets[m];
            const unsigned cols = ps - m + 1;
            for (unsigned l = m; l <= pt; ++l) {
                const std::size_t row = lm(l, static_cast<int>(m));
                C *out = block + static_cast<std::size_t>(l - m) * cols;
                for (unsigned n = m; n <= ps; ++n) {
                    const std::size_t col = lm(n, static_cast<int>(m));
                    out[n - m] += alpha[row] * source_value[col] + beta[row] * source_slope[col];
                }
            }
        }
    }
    for (const C &v : plan.coefficients)
        if (!finite(v))
            throw std::runtime_error("coaxial plan has a non-finite coefficient");
    return plan;
}

std::vector<C> apply_axial_plan(const AxialPlan &p, const std::vector<C> &source) {
    const unsigned mtop = std::min(p.source_order, p.target_order);
    if (source.size() != waves_for(p.source_order) || p.block_offsets.size() != mtop + 2 ||
        p.block_offsets.front() != 0 || p.block_offsets.back() != p.coefficients.size())
        throw std::invalid_argument("coaxial plan and source expansion do not match");
    for (unsigned m = 0; m <= mtop; ++m)
        if (p.block_offsets[m + 1] < p.block_offsets[m] ||
            p.block_offsets[m + 1] - p.block_offsets[m] !=
                static_cast<std::size_t>(p.target_order - m + 1) * (p.source_order - m + 1))
            throw std::invalid_argument("coaxial plan block " + std::to_string(m) + " has the wrong size");
    for (const C &v : source)
        if (!finite(v))
            throw std::invalid_argument("coaxial translation source must be finite");
    std::vector<C> out(waves_for(p.target_order));
    for (unsigned m = 0; m <= mtop; ++m) {
        const C *block = p.coefficients.data() + p.block_offsets[m];
        const unsigned cols = p.source_order - m + 1;
        // the +m and -m orders share the block
        for (const int sm : {static_cast<int>(m), -static_cast<int>(m)}) {
            if (m == 0 && sm < 0)
                break;
            for (unsigned l = m; l <= p.target_order; ++l) {
                const C *row = block + static_cast<std::size_t>(l - m) * cols;
                C sum{};
                for (unsigned n = m; n <= p.source_order; ++n)
                    sum += row[n - m] * source[lm(n, sm)];
                out[lm(l, sm)] = sum;
            }
        }
    }
    for (const C &v : out)
        if (!finite(v))
            throw std::runtime_error("coaxial translation produced a non-finite coefficient");
    return out;
}

} // namespace hsbem
