#include "support_tree.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <stdexcept>

namespace hsbem {
namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();

SupportBox empty_box() { return {{kInf, kInf, kInf}, {-kInf, -kInf, -kInf}, 0}; }

void grow(SupportBox &a, const SupportBox &b) {
    a.lower = {std::min(a.lower.x, b.lower.x), std::min(a.lower.y, b.lower.y), std::min(a.lower.z, b.lower.z)};
    a.upper = {std::max(a.upper.x, b.upper.x), std::max(a.upper.y, b.upper.y), std::max(a.upper.z, b.upper.z)};
    a.max_edge = std::max(a.max_edge, b.max_edge);
}

// Euclidean gap between two boxes (0 when they touch or overlap).
double box_gap(const SupportBox &a, const SupportBox &b) {
    const double gx = std::max({0., a.lower.x - b.upper.x, b.lower.x - a.upper.x});
    const double gy = std::max({0., a.lower.y - b.upper.y, b.lower.y - a.upper.y});
    const double gz = std::max({0., a.lower.z - b.upper.z, b.lower.z - a.upper.z});
    return std::hypot(gx, gy, gz);
}

double distance(Vec3 a, Vec3 b) { return std::hypot(a.x - b.x, a.y - b.y, a.z - b.z); }

class TreeBuilder {
  public:
    TreeBuilder(const Space &space, const SupportTreeSettings &settings) : sp_(space), cfg_(settings) {}

    SupportTree run() {
        const std::size_t nv = sp_.vertices.size(), nt = sp_.panels.size();
        if (!nv || !nt || sp_.incidence_offsets.size() != nv + 1 || nv >= kNoSupportNode || nt >= kNoSupportNode)
            throw std::invalid_argument("octree: empty or invalid geometry");
        t_.separation = cfg_.separation;
        t_.near_ratio = cfg_.near_ratio;
        // DP0 supports: one panel each; P1 supports: the incident panel star.
        SupportBox all = empty_box();
        t_.flux_supports.reserve(nt);
        for (const Panel &p : sp_.panels) {
            SupportBox b = empty_box();
            for (const Vec3 &v : p.vertices) {
                if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
                    throw std::invalid_argument("octree: coordinates must be finite");
                grow(b, {v, v, p.max_edge});
            }
            t_.flux_supports.push_back({p.centroid, b});
            grow(all, b);
        }
        t_.pressure_supports.reserve(nv);
        for (std::size_t v = 0; v < nv; ++v) {
            const std::size_t lo = sp_.incidence_offsets[v], hi = sp_.incidence_offsets[v + 1];
            if (lo >= hi || hi > sp_.incidences.size())
                throw std::invalid_argument("octree: a node's panel list is invalid");
            SupportBox b = empty_box();
            for (std::size_t j = lo; j < hi; ++j) {
                const std::uint32_t panel = sp_.incidences[j].panel;
                if (panel >= nt)
                    throw std::invalid_argument("octree: panel index out of range");
                grow(b, t_.flux_supports[panel].box);
            }
            t_.pressure_supports.push_back({sp_.vertices[v], b});
        }
        t_.pressure_order.resize(nv);
        t_.flux_order.resize(nt);
        std::iota(t_.pressure_order.begin(), t_.pressure_order.end(), 0u);
        std::iota(t_.flux_order.begin(), t_.flux_order.end(), 0u);
        t_.pressure_leaf.resize(nv);
        t_.flux_leaf.resize(nt);
        const Vec3 root{std::midpoint(all.lower.x, all.upper.x), std::midpoint(all.lower.y, all.upper.y),
                        std::midpoint(all.lower.z, all.upper.z)};
        const double half =
            .5 * std::max({all.upper.x - all.lower.x, all.upper.y - all.lower.y, all.upper.z - all.lower.z});
        if (!(half > 0) || !std::isfinite(half))
            throw std::invalid_argument("octree: root box is not finite");
        subdivide(root, std::nextafter(half, kInf), {0, 0, 0}, 0, kNoSupportNode, 0, static_cast<std::uint32_t>(nv), 0,
                  static_cast<std::uint32_t>(nt));
        nest_radii();
        dual_traverse(0, 0, false, t_.pressure_pressure);
        dual_traverse(0, 0, true, t_.pressure_flux);
        mark_active(t_.pressure_pressure, kPressureOutgoing);
        mark_active(t_.pressure_flux, kFluxOutgoing);
        offset_classes();
        for (const SupportNode &n : t_.nodes) {
            SupportLevelStats &l = t_.levels[n.level];
            l.pressure_outgoing += (n.active & kPressureOutgoing) != 0;
            l.flux_outgoing += (n.active & kFluxOutgoing) != 0;
            l.pressure_incoming += (n.active & kPressureIncoming) != 0;
            if (n.active & (kPressureOutgoing | kPressureIncoming))
                l.max_active_radius = std::max(l.max_active_radius, n.pressure_radius);
            if (n.active & kFluxOutgoing)
                l.max_active_radius = std::max(l.max_active_radius, n.flux_radius);
        }
        return std::move(t_);
    }

  private:
    // Stable counting sort of order[first, first+count) by octant of the anchor.
    std::array<std::uint32_t, 9> sort_octants(std::vector<std::uint32_t> &order, const std::vector<SupportAnchor> &sup,
                                              std::uint32_t first, std::uint32_t count, Vec3 c) {
        auto octant = [&](std::uint32_t id) {
            const Vec3 &a = sup[id].anchor;
            return static_cast<unsigned>(a.x >= c.x) | static_cast<unsigned>(a.y >= c.y) << 1 |
                   static_cast<unsigned>(a.z >= c.z) << 2;
        };
        std::array<std::uint32_t, 9> start{};
        for (std::uint32_t i = first; i < first + count; ++i)
            ++start[octant(order[i]) + 1];
        start[0] = first;
        for (int o = 1; o < 9; ++o)
            start[o] += start[o - 1];
        std::array<std::uint32_t, 9> fill = start;
        std::vector<std::uint32_t> out(count);
        for (std::uint32_t i = first; i < first + count; ++i) {
            const unsigned o = octant(order[i]);
            out[fill[o]++ - first] = order[i];
        }
        std::copy(out.begin(), out.end(), order.begin() + first);
        return start;
    }

    // Radius about c of the ball holding every vertex of the complete supports.
    double vertex_ball(Vec3 c, std::uint32_t first, std::uint32_t count, bool flux) const {
        double r = 0;
        auto panel = [&](std::uint32_t p) {
            for (const Vec3 &v : sp_.panels[p].vertices)
                r = std::max(r, distance(v, c));
        };
        for (std::uint32_t i = first; i < first + count; ++i) {
            if (flux) {
                panel(t_.flux_order[i]);
            } else {
                const std::uint32_t v = t_.pressure_order[i];
                for (std::size_t j = sp_.incidence_offsets[v]; j < sp_.incidence_offsets[v + 1]; ++j)
                    panel(sp_.incidences[j].panel);
            }
        }
        return count ? std::nextafter(r, kInf) : 0.0;
    }

    // Preorder node creation: a node is numbered before its children, children
    // in octant order.
    std::uint32_t subdivide(Vec3 c, double half, std::array<std::uint32_t, 3> coords, std::uint32_t level,
                            std::uint32_t parent, std::uint32_t fp, std::uint32_t np, std::uint32_t fq,
                            std::uint32_t nq) {
        if (t_.nodes.size() >= kNoSupportNode)
            throw std::overflow_error("octree: too many boxes for 32-bit indices");
        SupportNode n;
        n.center = c;
        n.half_width = half;
        n.coords = coords;
        n.level = level;
        n.parent = parent;
        n.first_pressure = fp;
        n.pressure_count = np;
        n.first_flux = fq;
        n.flux_count = nq;
        n.pressure_box = empty_box();
        n.flux_box = empty_box();
        for (std::uint32_t i = fp; i < fp + np; ++i)
            grow(n.pressure_box, t_.pressure_supports[t_.pressure_order[i]].box);
        for (std::uint32_t i = fq; i < fq + nq; ++i)
            grow(n.flux_box, t_.flux_supports[t_.flux_order[i]].box);
        const bool leaf = (np <= cfg_.leaf_capacity && nq <= cfg_.leaf_capacity) || level == cfg_.maximum_depth;
        if (leaf) {
            n.pressure_radius = vertex_ball(c, fp, np, false);
            n.flux_radius = vertex_ball(c, fq, nq, true);
        }
        const auto id = static_cast<std::uint32_t>(t_.nodes.size());
        t_.nodes.push_back(n);
        if (t_.levels.size() <= level)
            t_.levels.resize(level + 1);
        t_.levels[level].nodes.push_back(id);
        t_.levels[level].max_radius = std::max({t_.levels[level].max_radius, n.pressure_radius, n.flux_radius});
        if (leaf) {
            for (std::uint32_t i = fp; i < fp + np; ++i)
                t_.pressure_leaf[t_.pressure_order[i]] = id;
            for (std::uint32_t i = fq; i < fq + nq; ++i)
                t_.flux_leaf[t_.flux_order[i]] = id;
            return id;
        }
        const auto ps = sort_octants(t_.pressure_order, t_.pressure_supports, fp, np, c);
        const auto qs = sort_octants(t_.flux_order, t_.flux_supports, fq, nq, c);
        const double h = half * .5;
        for (unsigned o = 0; o < 8; ++o) {
            if (ps[o] == ps[o + 1] && qs[o] == qs[o + 1])
                continue;
            const Vec3 cc{c.x + ((o & 1) ? h : -h), c.y + ((o & 2) ? h : -h), c.z + ((o & 4) ? h : -h)};
            std::array<std::uint32_t, 3> k = coords;
            for (unsigned a = 0; a < 3; ++a)
                k[a] = 2 * coords[a] + ((o >> a) & 1);
            const std::uint32_t child =
                subdivide(cc, h, k, level + 1, id, ps[o], ps[o + 1] - ps[o], qs[o], qs[o + 1] - qs[o]);
            t_.nodes[id].children[o] = child;
        }
        return id;
    }

    // Interior balls enclose their children's balls (outward-rounded); nodes
    // come after their parent in preorder, so one reverse sweep nests them.
    void nest_radii() {
        for (std::size_t i = t_.nodes.size(); i-- > 1;) {
            const SupportNode &ch = t_.nodes[i];
            SupportNode &pa = t_.nodes[ch.parent];
            const double d = distance(ch.center, pa.center);
            if (ch.pressure_count)
                pa.pressure_radius = std::max(pa.pressure_radius, std::nextafter(d + ch.pressure_radius, kInf));
            if (ch.flux_count)
                pa.flux_radius = std::max(pa.flux_radius, std::nextafter(d + ch.flux_radius, kInf));
        }
        for (auto &l : t_.levels)
            l.max_radius = 0;
        for (const SupportNode &n : t_.nodes)
            t_.levels[n.level].max_radius = std::max({t_.levels[n.level].max_radius, n.pressure_radius, n.flux_radius});
    }

    void dual_traverse(std::uint32_t ti, std::uint32_t si, bool flux, SupportLists &out) {
        const SupportNode &t = t_.nodes[ti], &s = t_.nodes[si];
        const std::uint32_t ns = flux ? s.flux_count : s.pressure_count;
        if (!t.pressure_count || !ns)
            return;
        const SupportBox &sb = flux ? s.flux_box : s.pressure_box;
        const double sr = flux ? s.flux_radius : s.pressure_radius;
        const double d = distance(t.center, s.center), g = box_gap(t.pressure_box, sb);
        const bool far = t.level == s.level && g > 0 &&
                         g > cfg_.near_ratio * std::max(t.pressure_box.max_edge, sb.max_edge) &&
                         d > cfg_.separation * (t.pressure_radius + sr);
        const bool tl = t.leaf(), sl = s.leaf();
        if (far || (tl && sl)) {
            if (interactions_ == cfg_.maximum_interactions)
                throw std::runtime_error("octree: interaction lists exceed their memory budget");
            ++interactions_;
            const std::uint64_t pairs = static_cast<std::uint64_t>(t.pressure_count) * ns;
            if (far) {
                out.far.push_back({ti, si});
                out.represented_far += pairs;
            } else {
                out.near.push_back({ti, si});
                out.represented_near += pairs;
            }
            return;
        }
        // Descend the coarser (or non-leaf) side; a coarse leaf never becomes a
        // far endpoint against finer nodes (its products stay near).
        const std::array<std::uint32_t, 8> tc = t.children, sc = s.children;
        const std::uint32_t tlev = t.level, slev = s.level;
        if (!tl && (sl || tlev < slev)) {
            for (auto c : tc)
                if (c != kNoSupportNode)
                    dual_traverse(c, si, flux, out);
        } else if (!sl && (tl || slev < tlev)) {
            for (auto c : sc)
                if (c != kNoSupportNode)
                    dual_traverse(ti, c, flux, out);
        } else {
            for (auto a : tc)
                if (a != kNoSupportNode)
                    for (auto b : sc)
                        if (b != kNoSupportNode)
                            dual_traverse(a, b, flux, out);
        }
    }

    void activate(std::uint32_t id, unsigned channel) {
        SupportNode &n = t_.nodes[id];
        const bool has = channel == kFluxOutgoing ? n.flux_count != 0 : n.pressure_count != 0;
        if (!has || (n.active & channel))
            return;
        n.active |= channel;
        const std::array<std::uint32_t, 8> ch = n.children;
        for (auto c : ch)
            if (c != kNoSupportNode)
                activate(c, channel);
    }

    void mark_active(SupportLists &lists, unsigned outgoing) {
        std::map<std::pair<std::uint32_t, std::uint32_t>, SupportLevelPair> by_level;
        for (const SupportPair &p : lists.far) {
            activate(p.target, kPressureIncoming);
            activate(p.source, outgoing);
            const SupportNode &t = t_.nodes[p.target], &s = t_.nodes[p.source];
            SupportLevelStats &l = t_.levels[t.level];
            if (outgoing == kPressureOutgoing)
                ++l.pp_far;
            else
                ++l.pf_far;
            SupportLevelPair &e = by_level[{t.level, s.level}];
            e.target_level = t.level;
            e.source_level = s.level;
            ++e.interactions;
            e.represented_pairs += static_cast<std::uint64_t>(t.pressure_count) *
                                   (outgoing == kPressureOutgoing ? s.pressure_count : s.flux_count);
        }
        for (const auto &[key, e] : by_level)
            lists.far_levels.push_back(e);
    }

    // Exact integer dyadic offsets of far pairs per level: offset, |offset|^2
    // and polar classes (sign z, z^2 / g, |o|^2 / g with g = gcd(z^2, |o|^2)).
    void offset_classes() {
        using Offset = std::array<std::int64_t, 3>;
        const std::size_t nl = t_.levels.size();
        std::vector<std::set<Offset>> pp(nl), pf(nl);
        auto collect = [&](const SupportLists &lists, std::vector<std::set<Offset>> &sets) {
            for (const SupportPair &p : lists.far) {
                const SupportNode &t = t_.nodes[p.target], &s = t_.nodes[p.source];
                Offset o{};
                for (int a = 0; a < 3; ++a)
                    o[a] = static_cast<std::int64_t>(t.coords[a]) - s.coords[a];
                sets[t.level].insert(o);
            }
        };
        collect(t_.pressure_pressure, pp);
        collect(t_.pressure_flux, pf);
        auto squared = [](const std::set<Offset> &set) {
            std::set<std::uint64_t> r;
            for (const Offset &o : set)
                r.insert(static_cast<std::uint64_t>(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
            return r;
        };
        auto polar = [](const std::set<Offset> &set) {
            std::set<std::array<std::uint64_t, 3>> r;
            for (const Offset &o : set) {
                const auto z2 = static_cast<std::uint64_t>(o[2] * o[2]);
                const auto n2 = static_cast<std::uint64_t>(o[0] * o[0] + o[1] * o[1]) + z2;
                const std::uint64_t g = std::gcd(z2, n2);
                r.insert({o[2] < 0 ? 0u : o[2] > 0 ? 2u : 1u, z2 / g, n2 / g});
            }
            return r;
        };
        for (std::size_t l = 0; l < nl; ++l) {
            SupportLevelStats &s = t_.levels[l];
            s.pp_offsets = pp[l].size();
            s.pf_offsets = pf[l].size();
            auto pol = polar(pp[l]);
            const auto polf = polar(pf[l]);
            s.pp_polar_classes = pol.size();
            s.pf_polar_classes = polf.size();
            pol.insert(polf.begin(), polf.end());
            s.union_polar_classes = pol.size();
            auto sq = squared(pp[l]);
            const auto sqf = squared(pf[l]);
            s.pp_distance_classes = sq.size();
            s.pf_distance_classes = sqf.size();
            sq.insert(sqf.begin(), sqf.end());
            s.squared_offsets.assign(sq.begin(), sq.end());
            pp[l].insert(pf[l].begin(), pf[l].end());
            s.union_offsets = pp[l].size();
        }
    }

    const Space &sp_;
    SupportTreeSettings cfg_;
    SupportTree t_;
    std::uint64_t interactions_{};
};

} // namespace

SupportTree make_support_tree(const Space &space, const SupportTreeSettings &settings) {
    if (!settings.leaf_capacity || settings.maximum_depth > 30 || !std::isfinite(settings.separation) ||
        settings.separation <= 1 || !std::isfinite(settings.near_ratio) || settings.near_ratio < 0)
        throw std::invalid_argument("octree: invalid leaf capacity, depth, separation or near ratio");
    return TreeBuilder(space, settings).run();
}

} // namespace hsbem
