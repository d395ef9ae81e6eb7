#include "fmm_levels.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <limits>
#include <math.h>
#include <stdexcept>

namespace hsbem {
namespace {

constexpr unsigned kMaximumAngularOrder = 1024;
constexpr double kInf = std::numeric_limits<double>::infinity();

void require_spherical_domain(const WidebandLevelPlan &p) {
    if (p.electrical_radius > kMaximumSphericalElectricalRadius || p.signature_order > kMaximumSphericalOrder)
        throw std::runtime_error("spherical expansion plan: the conditioned order is beyond the supported range");
}

} // namespace

WidebandLevelPlan make_wideband_level_plan(double radius, double k, double delta, double eps) {
    if (!std::isfinite(radius) || radius < 0 || !std::isfinite(k) || k <= 0 || !std::isfinite(delta) || delta <= 1 ||
        !std::isfinite(eps) || eps <= 0 || eps >= 1)
        throw std::invalid_argument("level planning needs finite positive box sizes, wavenumber and accuracy");
    WidebandLevelPlan p;
    p.active_radius = radius;
    p.electrical_radius = k * radius;
    p.accuracy_seed = eps;
    if (!std::isfinite(p.electrical_radius))
        throw std::invalid_argument("level planning: electrical radius is not finite");
    if (radius == 0)
        return p;
    const double kr = p.electrical_radius;
    const double low = 1 - (std::log(eps) + 1.5 * std::log1p(-1 / delta)) / std::log(delta);
    const double high = kr + .5 * std::pow(3 * -std::log(eps), 2. / 3) * std::cbrt(kr);
    const double top = std::max(low, high);
    const double order = top * std::pow(std::pow(low / top, 4) + std::pow(high / top, 4), .25);
    if (!std::isfinite(order) || std::ceil(order) > kMaximumAngularOrder / 2)
        throw std::runtime_error("level planning: plane-wave order is beyond the supported range");
    p.signature_order = static_cast<unsigned>(std::ceil(order));
    p.translation_order = p.quadrature_order = 2 * p.signature_order;
    p.spherical_coefficients = static_cast<std::uint64_t>(p.signature_order + 1) * (p.signature_order + 1);
    const auto q = static_cast<std::uint64_t>(p.quadrature_order + 1);
    p.angular_samples = 2 * q * q;
    p.representation = order - high <= 2 ? LevelRepresentation::plane_wave : LevelRepresentation::regular_spherical;
    if (p.representation == LevelRepresentation::regular_spherical &&
        (p.electrical_radius > kMaximumSphericalElectricalRadius || p.signature_order > kMaximumSphericalOrder))
        throw std::runtime_error("Wideband LF seed requires a spherical domain beyond the supported numerical/resource "
                                 "contract.");
    return p;
}

double enforce_diagonal_stability(WidebandLevelPlan &p, double k, double distance) {
    if (p.representation != LevelRepresentation::plane_wave || distance == 0)
        return 0;
    if (!std::isfinite(distance) || distance <= 0 || !std::isfinite(k) || k <= 0)
        throw std::invalid_argument("diagonal-translation check needs a positive separation");
    const double z = k * distance;
    if (!std::isfinite(z) || !(z > 0))
        throw std::invalid_argument("diagonal-translation check: electrical separation is not finite");
    // h_l^(2)(z) by upward recurrence from h_0 = (sin z + i cos z)/z, h_1 = h_0 (1/z + i).
    using C = std::complex<double>;
    double sz, cz;
    ::sincos(z, &sz, &cz); // joint routine, as the reference build evaluates this pair
    C h_prev{sz / z, cz / z};
    C h = h_prev * C{1 / z, 1};
    double indicator = z * std::abs(h_prev);
    for (unsigned l = 1; l <= p.translation_order; ++l) {
        indicator += z * (2 * l + 1) * std::abs(h);
        if (!std::isfinite(indicator))
            break;
        const C next = (static_cast<double>(2 * l + 1) / z) * h - h_prev;
        h_prev = h;
        h = next;
    }
    if (!std::isfinite(indicator) || std::numeric_limits<double>::epsilon() * indicator > .01 * p.accuracy_seed) {
        require_spherical_domain(p);
        p.representation = LevelRepresentation::regular_spherical;
    }
    return indicator;
}

SupportExpansionPlan make_support_expansion_plan(const SupportTree &tree, const std::vector<std::uint32_t> &face_column,
                                                 double k, double seed) {
    if (face_column.size() != tree.flux_leaf.size() || !std::isfinite(k) || k <= 0 || !std::isfinite(seed) ||
        seed <= 0 || seed >= 1)
        throw std::invalid_argument("expansion plan: invalid geometry or accuracy seed");
    const std::size_t nn = tree.nodes.size();
    SupportExpansionPlan out;
    out.levels.resize(tree.levels.size());
    out.active.assign(nn, 0);
    // nodes holding an active flux panel (the path to the root of each one)
    std::vector<char> flux(nn, 0);
    for (std::size_t panel = 0; panel < face_column.size(); ++panel) {
        if (face_column[panel] == UINT32_MAX)
            continue;
        for (std::uint32_t n = tree.flux_leaf[panel]; n != kNoSupportNode; n = tree.nodes[n].parent) {
            if (n >= nn)
                throw std::invalid_argument("expansion plan: a flux leaf index is out of range");
            if (flux[n])
                break;
            flux[n] = 1;
        }
    }
    // delta[node] = {pressure source, flux source, pressure target}
    std::vector<std::array<double, 3>> delta(nn, {kInf, kInf, kInf});
    auto visit = [&](const SupportPair &pr, bool is_flux) {
        if (pr.source >= nn || pr.target >= nn)
            throw std::invalid_argument("expansion plan: a far-pair node index is out of range");
        if (is_flux && !flux[pr.source])
            return;
        const SupportNode &s = tree.nodes[pr.source], &t = tree.nodes[pr.target];
        const double rs = is_flux ? s.flux_radius : s.pressure_radius, rt = t.pressure_radius;
        const double d = std::hypot(t.center.x - s.center.x, t.center.y - s.center.y, t.center.z - s.center.z);
        if (s.level != t.level || s.level >= tree.levels.size() || !(rs > 0) || !(rt > 0) || !std::isfinite(d) ||
            !(d > rs + rt))
            throw std::invalid_argument("expansion plan: far pairs must be separated boxes of one level");
        double &ds = delta[pr.source][is_flux ? 1 : 0];
        ds = std::min(ds, (d - rt) / rs);
        delta[pr.target][2] = std::min(delta[pr.target][2], (d - rs) / rt);
        double &closest = out.levels[s.level].minimum_far_distance;
        closest = closest > 0 ? std::min(closest, d) : d;
    };
    for (const SupportPair &p : tree.pressure_pressure.far)
        visit(p, false);
    for (const SupportPair &p : tree.pressure_flux.far)
        visit(p, true);
    for (std::size_t l = 0; l < tree.levels.size(); ++l) {
        SupportLevelExpansion &lev = out.levels[l];
        lev.source_separation = lev.target_separation = kInf;
        for (std::uint32_t id : tree.levels[l].nodes) {
            if (id >= nn || tree.nodes[id].level != l)
                throw std::invalid_argument("expansion plan: level list is inconsistent");
            const SupportNode &n = tree.nodes[id];
            if (n.parent != kNoSupportNode) {
                if (n.parent >= nn || tree.nodes[n.parent].level >= l)
                    throw std::invalid_argument("expansion plan: a parent does not precede its child");
                const SupportNode &pa = tree.nodes[n.parent];
                const double shift = std::hypot(n.center.x - pa.center.x, n.center.y - pa.center.y, n.center.z - pa.center.z);
                auto nested = [&](double rc, double rp) {
                    return shift + rc <= rp * (1 + 32 * std::numeric_limits<double>::epsilon());
                };
                if ((n.pressure_count && !nested(n.pressure_radius, pa.pressure_radius)) ||
                    (n.flux_count && flux[id] && !nested(n.flux_radius, pa.flux_radius)))
                    throw std::invalid_argument("expansion plan: a child's ball is not inside its parent's");
                for (int c = 0; c < 3; ++c)
                    delta[id][c] = std::min(delta[id][c], delta[n.parent][c]);
            }
            if (n.pressure_count) {
                if (std::isfinite(delta[id][0])) {
                    out.active[id] |= kPressureOutgoing;
                    lev.source_separation = std::min(lev.source_separation, delta[id][0]);
                }
                if (std::isfinite(delta[id][2])) {
                    out.active[id] |= kPressureIncoming;
                    lev.target_separation = std::min(lev.target_separation, delta[id][2]);
                }
                if (out.active[id] & (kPressureOutgoing | kPressureIncoming))
                    lev.max_active_radius = std::max(lev.max_active_radius, n.pressure_radius);
            }
            if (n.flux_count && flux[id] && std::isfinite(delta[id][1])) {
                out.active[id] |= kFluxOutgoing;
                lev.source_separation = std::min(lev.source_separation, delta[id][1]);
                lev.max_active_radius = std::max(lev.max_active_radius, n.flux_radius);
            }
        }
    }
    // coarseward monotone envelope of the active radius
    double envelope = 0;
    for (auto it = out.levels.rbegin(); it != out.levels.rend(); ++it)
        if (it->max_active_radius > 0) {
            envelope = std::max(envelope, it->max_active_radius);
            it->expansion.active_radius = envelope;
        }
    double policy = kInf;
    bool spherical = false;
    for (SupportLevelExpansion &lev : out.levels) {
        if (!(lev.max_active_radius > 0)) {
            lev.source_separation = lev.target_separation = 0;
            continue;
        }
        policy = std::min({policy, lev.source_separation, lev.target_separation});
        if (!std::isfinite(policy) || !(policy > 1))
            throw std::invalid_argument("expansion plan: convergence ratio is not finite");
        lev.policy_separation = policy;
        lev.expansion = make_wideband_level_plan(lev.expansion.active_radius, k, policy, seed);
        lev.diagonal_condition = enforce_diagonal_stability(lev.expansion, k, lev.minimum_far_distance);
        if (spherical) { // once a level is spherical, every finer level stays spherical
            require_spherical_domain(lev.expansion);
            lev.expansion.representation = LevelRepresentation::regular_spherical;
        }
        spherical |= lev.expansion.representation == LevelRepresentation::regular_spherical;
        if (!std::isfinite(lev.source_separation))
            lev.source_separation = 0;
        if (!std::isfinite(lev.target_separation))
            lev.target_separation = 0;
    }
    return out;
}

} // namespace hsbem
