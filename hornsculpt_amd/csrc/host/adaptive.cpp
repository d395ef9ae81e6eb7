// Adaptive near-pair rule. The arithmetic follows pair_quadrature.cpp step for
// step (same point maps, same complex kernel expressions, same heap order), so
// the patch lists it produces are identical to the reference planner's.
#include "adaptive.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace hsbem {
namespace {

using Moments = std::array<std::complex<double>, 36>;
using Bary = std::array<double, 3>;

Vec3 local_position(const Panel &p, const double *b) {
    return vadd(vscale(vsub(p.vertices[1], p.vertices[0]), b[1]), vscale(vsub(p.vertices[2], p.vertices[0]), b[2]));
}

// Closest-point distances after Ericson, Real-Time Collision Detection, 5.1.5 / 5.1.9.
double point_triangle_d2(Vec3 p, const std::array<Vec3, 3> &t) {
    const Vec3 a = t[0], b = t[1], c = t[2];
    const Vec3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(p, a);
    const double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0 && d2 <= 0)
        return vnorm2(ap);
    const Vec3 bp = vsub(p, b);
    const double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0 && d4 <= d3)
        return vnorm2(bp);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0)
        return vnorm2(vsub(p, vadd(a, vscale(ab, d1 / (d1 - d3)))));
    const Vec3 cp = vsub(p, c);
    const double d5 = vdot(ab, cp), d6 = vdot(ac, cp);
    if (d6 >= 0 && d5 <= d6)
        return vnorm2(cp);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0)
        return vnorm2(vsub(p, vadd(a, vscale(ac, d2 / (d2 - d6)))));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return vnorm2(vsub(p, vadd(b, vscale(vsub(c, b), w))));
    }
    const double inv = 1 / (va + vb + vc);
    return vnorm2(vsub(p, vadd(a, vadd(vscale(ab, vb * inv), vscale(ac, vc * inv)))));
}

double segment_d2(Vec3 p, Vec3 q, Vec3 a, Vec3 b) {
    const Vec3 u = vsub(q, p), v = vsub(b, a), w = vsub(p, a);
    const double aa = vdot(u, u), bb = vdot(u, v), cc = vdot(v, v), dd = vdot(u, w), ee = vdot(v, w);
    const double det = aa * cc - bb * bb;
    double s = det > aa * cc * 1e-14 ? std::clamp((bb * ee - cc * dd) / det, 0.0, 1.0) : 0;
    double t = (bb * s + ee) / cc;
    if (t < 0) {
        t = 0;
        s = std::clamp(-dd / aa, 0.0, 1.0);
    } else if (t > 1) {
        t = 1;
        s = std::clamp((bb - dd) / aa, 0.0, 1.0);
    }
    return vnorm2(vsub(vadd(w, vscale(u, s)), vscale(v, t)));
}

double triangle_distance(const std::array<Vec3, 3> &a, const std::array<Vec3, 3> &b) {
    double d2 = std::numeric_limits<double>::infinity();
    for (unsigned i = 0; i < 3; ++i) {
        d2 = std::min({d2, point_triangle_d2(a[i], b), point_triangle_d2(b[i], a)});
        for (unsigned j = 0; j < 3; ++j)
            d2 = std::min(d2, segment_d2(a[i], a[(i + 1) % 3], b[j], b[(j + 1) % 3]));
    }
    return std::sqrt(d2);
}

Bary midpoint(const Bary &a, const Bary &b) { return {(a[0] + b[0]) * .5, (a[1] + b[1]) * .5, (a[2] + b[2]) * .5}; }

std::array<Vec3, 3> physical_corners(const Panel &p, const std::array<Bary, 3> &v, Vec3 origin) {
    const Vec3 off = vsub(p.vertices[0], origin);
    return {vadd(off, local_position(p, v[0].data())), vadd(off, local_position(p, v[1].data())),
            vadd(off, local_position(p, v[2].data()))};
}

double diameter(const std::array<Vec3, 3> &p) {
    return std::sqrt(std::max({vnorm2(vsub(p[0], p[1])), vnorm2(vsub(p[1], p[2])), vnorm2(vsub(p[2], p[0]))}));
}

// G = exp(-ikr)/r and the dimensionless gradient components -(1+ikr) G scale d/r^2.
inline std::array<std::complex<double>, 4> kernels(Vec3 d, double k, double scale) {
    const double r = std::sqrt(vnorm2(d));
    if (!(r > 0))
        throw std::invalid_argument("near pair: the two panels overlap or repeat a vertex position");
    const auto g = std::exp(std::complex<double>(0, -k * r)) / r;
    const auto dg = -g * std::complex<double>(1, k * r) * (scale / (r * r));
    return {g, dg * d.x, dg * d.y, dg * d.z};
}

struct SubTri {
    std::array<Bary, 3> v;
    double fraction;
};

Moments tensor_moments(const Panel &test, const Panel &trial, const SubTri &tx, const SubTri &ty,
                       const std::vector<TrianglePoint> &rule, double k, double scale) {
    Moments m{};
    const Vec3 base = vsub(test.vertices[0], trial.vertices[0]);
    for (const auto &x : rule)
        for (const auto &y : rule) {
            double pt[3]{}, ps[3]{};
            const double w = x.weight * y.weight * tx.fraction * ty.fraction;
            for (unsigned i = 0; i < 3; ++i)
                for (unsigned j = 0; j < 3; ++j) {
                    pt[i] += x.bary[j] * tx.v[j][i];
                    ps[i] += y.bary[j] * ty.v[j][i];
                }
            const auto kv = kernels(vadd(base, vsub(local_position(test, pt), local_position(trial, ps))), k, scale);
            for (unsigned i = 0; i < 3; ++i)
                for (unsigned j = 0; j < 3; ++j) {
                    const double wij = w * pt[i] * ps[j];
                    for (unsigned c = 0; c < 4; ++c)
                        m[c * 9 + i * 3 + j] += kv[c] * wij;
                }
        }
    return m;
}

FanPanel fan_panel(const Panel &p) { return {{p.vertices[0], p.vertices[1], p.vertices[2]}, p.normal, p.area}; }

Moments fan_moments(const Panel &test, const Panel &trial, const SubTri &tx, const std::vector<TrianglePoint> &outer,
                    unsigned order, unsigned intervals, double k, double scale) {
    PairPatch patch{};
    patch.projection = 1;
    patch.area_fraction = tx.fraction;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j)
            patch.test[3 * i + j] = tx.v[i][j];
    static const auto g14 = gauss_legendre01(14), g20 = gauss_legendre01(20);
    const auto &base = order == 14 ? g14 : g20;
    std::vector<GaussPoint> gauss;
    gauss.reserve(base.size() * intervals);
    for (unsigned iv = 0; iv < intervals; ++iv)
        for (const auto &g : base)
            gauss.push_back({(iv + g.position) / intervals, g.weight / intervals});
    const FanPanel ft = fan_panel(test), fs = fan_panel(trial);
    Moments m{};
    const Vec3 off = vsub(test.vertices[0], trial.vertices[0]);
    for (const auto &x : outer) {
        const FanOuter o = fan_outer(x, patch, ft, fs);
        const Vec3 lt = local_position(test, o.test);
        std::array<std::complex<double>, 12> inner{};
        for (unsigned e = 0; e < 3; ++e) {
            const FanEdge edge = fan_edge(o, e, fs);
            for (const auto &u : gauss) {
                const FanRay ray = fan_ray(o, edge, u);
                for (const auto &v : gauss) {
                    const PairPoint p = fan_point(o, ray, v);
                    const auto kv = kernels(vadd(off, vsub(lt, local_position(trial, p.trial))), k, scale);
                    for (unsigned j = 0; j < 3; ++j) {
                        const double w = p.weight * p.trial[j];
                        for (unsigned c = 0; c < 4; ++c)
                            inner[3 * c + j] += kv[c] * w;
                    }
                }
            }
        }
        for (unsigned i = 0; i < 3; ++i)
            for (unsigned j = 0; j < 3; ++j)
                for (unsigned c = 0; c < 4; ++c)
                    m[9 * c + 3 * i + j] += o.test[i] * inner[3 * c + j];
    }
    return m;
}

// Split the test triangle along the three projected trial edge lines and fan
// triangulate every piece (both sides are kept; nothing is clipped away).
std::vector<SubTri> projection_partition(const Panel &test, const Panel &trial) {
    std::array<Bary, 3> proj{};
    for (unsigned i = 0; i < 3; ++i) {
        const Vec3 rel = vsub(test.vertices[i], trial.vertices[0]);
        proj[i][1] = vdot(rel, trial.gradients[1]);
        proj[i][2] = vdot(rel, trial.gradients[2]);
        proj[i][0] = 1 - proj[i][1] - proj[i][2];
    }
    using Polygon = std::vector<Bary>;
    std::vector<Polygon> polys{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    for (unsigned edge = 0; edge < 3; ++edge) {
        auto value = [&](const Bary &p) { return p[0] * proj[0][edge] + p[1] * proj[1][edge] + p[2] * proj[2][edge]; };
        std::vector<Polygon> next;
        for (const auto &poly : polys) {
            bool pos = false, neg = false;
            for (const auto &p : poly) {
                pos |= value(p) > 0;
                neg |= value(p) < 0;
            }
            if (!(pos && neg)) {
                next.push_back(poly);
                continue;
            }
            Polygon plus, minus;
            for (unsigned i = 0; i < poly.size(); ++i) {
                const Bary a = poly[i], b = poly[(i + 1) % poly.size()];
                const double va = value(a), vb = value(b);
                if (va >= 0)
                    plus.push_back(a);
                if (va <= 0)
                    minus.push_back(a);
                if ((va < 0 && vb > 0) || (va > 0 && vb < 0)) {
                    const double t = va / (va - vb);
                    Bary x{};
                    for (unsigned j = 0; j < 3; ++j)
                        x[j] = a[j] + t * (b[j] - a[j]);
                    plus.push_back(x);
                    minus.push_back(x);
                }
            }
            if (plus.size() >= 3)
                next.push_back(std::move(plus));
            if (minus.size() >= 3)
                next.push_back(std::move(minus));
        }
        polys = std::move(next);
    }
    std::vector<SubTri> out;
    double area = 0;
    for (const auto &poly : polys)
        for (unsigned i = 1; i + 1 < poly.size(); ++i) {
            const Bary a = poly[0], b = poly[i], c = poly[i + 1];
            const double f = std::abs((b[1] - a[1]) * (c[2] - a[2]) - (b[2] - a[2]) * (c[1] - a[1]));
            if (f > 0) {
                out.push_back({{a, b, c}, f});
                area += f;
            }
        }
    if (std::abs(area - 1) > 1024 * std::numeric_limits<double>::epsilon())
        throw std::runtime_error("near pair: the projected fan partition does not cover the panel's area");
    return out;
}

bool use_projection(const Panel &test, const Panel &trial, double scale) {
    if (panel_distance(test, trial) >= .025 * scale)
        return false;
    const double roundoff = 128 * std::numeric_limits<double>::epsilon() * scale;
    double lo = std::numeric_limits<double>::infinity(), hi = -lo;
    for (unsigned i = 0; i < 3; ++i) {
        const double h = vdot(vsub(test.vertices[i], trial.vertices[0]), trial.normal);
        lo = std::min(lo, h);
        hi = std::max(hi, h);
    }
    return (lo >= -roundoff && hi > roundoff) || (hi <= roundoff && lo < -roundoff);
}

double max_norm(const Moments &m) {
    double out = 0;
    for (const auto &v : m) {
        const double a = std::abs(v);
        if (!std::isfinite(a))
            throw std::runtime_error("near pair: a quadrature moment is not finite");
        out = std::max(out, a);
    }
    return out;
}

} // namespace

double panel_distance(const Panel &test, const Panel &trial) {
    std::array<Vec3, 3> a{}, b{};
    for (unsigned i = 0; i < 3; ++i) {
        a[i] = vsub(test.vertices[i], test.vertices[0]);
        b[i] = vsub(trial.vertices[i], test.vertices[0]);
    }
    return triangle_distance(a, b);
}

bool is_near_pair(const Panel &test, const Panel &trial, double ratio) {
    if (!std::isfinite(ratio) || ratio < 0)
        throw std::invalid_argument("near distance ratio must be finite and not negative");
    if (ratio == 0)
        return false;
    if (classify_pair_ids(test.nodes.data(), trial.nodes.data()).topology != Topology::disjoint)
        return false;
    const double size = std::max(test.max_edge, trial.max_edge);
    const double bound = ratio * size + test.max_edge + trial.max_edge;
    if (vnorm2(vsub(test.centroid, trial.centroid)) >= bound * bound)
        return false;
    return panel_distance(test, trial) < ratio * size;
}

PatchEstimate estimate_patch(const Panel &test, const Panel &trial, const PairPatch &patch, double k) {
    if ((patch.degree != 8 && patch.degree != 14) || !std::isfinite(k) || k < 0)
        throw std::invalid_argument("Invalid adaptive patch estimate.");
    static const auto r6 = triangle_rule(6), r8 = triangle_rule(8), r12 = triangle_rule(12), r14 = triangle_rule(14);
    SubTri a{}, b{};
    a.fraction = patch.area_fraction;
    b.fraction = 1;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j) {
            a.v[i][j] = patch.test[3 * i + j];
            b.v[i][j] = patch.trial[3 * i + j];
        }
    const double scale = std::max(test.max_edge, trial.max_edge);
    const auto &lo_rule = patch.degree == 8 ? r6 : r12;
    const auto &hi_rule = patch.degree == 8 ? r8 : r14;
    const Moments lo = patch.projection ? fan_moments(test, trial, a, lo_rule, 20, patch.projection_intervals, k, scale)
                                        : tensor_moments(test, trial, a, b, lo_rule, k, scale);
    const Moments hi = patch.projection ? fan_moments(test, trial, a, hi_rule, 20, patch.projection_intervals, k, scale)
                                        : tensor_moments(test, trial, a, b, hi_rule, k, scale);
    Moments diff{};
    for (unsigned i = 0; i < 36; ++i)
        diff[i] = hi[i] - lo[i];
    (void)max_norm(hi);
    const double outer = max_norm(diff);
    double inner = 0;
    if (patch.projection) {
        const Moments coarse = fan_moments(test, trial, a, hi_rule, 14, patch.projection_intervals, k, scale);
        for (unsigned i = 0; i < 36; ++i)
            diff[i] = hi[i] - coarse[i];
        inner = max_norm(diff);
    }
    return {hi, outer + inner, inner, outer};
}

// ---------------------------------------------------------------------------

AdaptivePair::Leaf AdaptivePair::make_leaf(const Sub &a, const Sub &b, unsigned depth, unsigned degree,
                                           unsigned intervals) const {
    Leaf l;
    l.test = a;
    l.trial = b;
    l.test_diameter = diameter(physical_corners(test_, a.v, test_.vertices[0]));
    l.trial_diameter = diameter(physical_corners(trial_, b.v, test_.vertices[0]));
    l.depth = depth;
    l.degree = degree;
    l.intervals = intervals;
    return l;
}

PairPatch AdaptivePair::to_patch(const Leaf &l) const {
    PairPatch p{};
    p.degree = l.degree;
    p.projection = projection_ ? 1 : 0;
    p.projection_intervals = l.intervals;
    p.area_fraction = l.test.fraction * l.trial.fraction;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j) {
            p.test[3 * i + j] = l.test.v[i][j];
            p.trial[3 * i + j] = l.trial.v[i][j];
        }
    return p;
}

AdaptivePair::AdaptivePair(const Panel &test, const Panel &trial, double tolerance, unsigned max_depth, double k)
    : test_(test), trial_(trial), tolerance_(tolerance), k_(k), max_depth_(max_depth) {
    if (!std::isfinite(k) || k < 0 || !std::isfinite(tolerance) || tolerance <= 0 || max_depth > 16 ||
        classify_pair_ids(test.nodes.data(), trial.nodes.data()).topology != Topology::disjoint)
        throw std::invalid_argument("near pair: invalid refinement settings or the panels share a vertex");
    projection_ = use_projection(test, trial, std::max(test.max_edge, trial.max_edge));
    promote_far_root_ = !projection_ && !is_near_pair(test, trial, 1.0);
    if (projection_) {
        for (const auto &part : projection_partition(test, trial)) {
            Sub s;
            s.v = part.v;
            s.fraction = part.fraction;
            initial_.push_back(s);
        }
        if (initial_.size() > 16)
            throw std::logic_error("near pair: the projected partition has too many pieces");
    } else {
        initial_.push_back(Sub{});
    }
    request_initial();
}

void AdaptivePair::request_initial() {
    requested_leaves_.clear();
    while (next_initial_ < initial_.size() && requested_leaves_.size() < 4)
        requested_leaves_.push_back(make_leaf(initial_[next_initial_++], Sub{}, 0, 8));
    requested_patches_.clear();
    for (const auto &l : requested_leaves_)
        requested_patches_.push_back(to_patch(l));
}

void AdaptivePair::include(Leaf l) {
    absolute_error_ += l.error;
    if (!std::isfinite(absolute_error_))
        throw std::runtime_error("near pair: the accumulated error estimate is not finite");
    for (unsigned i = 0; i < 36; ++i)
        integral_[i] += l.integral[i];
    const bool outer_limit = l.depth == max_depth_ && l.degree == 14;
    const unsigned next_inner = l.intervals * 2;
    const bool inner_limit = next_inner > (1u << max_depth_) ||
                             std::uint64_t(192) * 400 * next_inner * next_inner > std::numeric_limits<unsigned>::max();
    const auto by_error = [](const Leaf &a, const Leaf &b) { return a.error < b.error; };
    if ((!projection_ && outer_limit) ||
        (projection_ && (l.inner_error > l.outer_error ? inner_limit : outer_limit))) {
        at_limit_.push_back(std::move(l));
    } else {
        pending_.push_back(std::move(l));
        std::push_heap(pending_.begin(), pending_.end(), by_error);
    }
}

void AdaptivePair::advance() {
    const auto by_error = [](const Leaf &a, const Leaf &b) { return a.error < b.error; };
    requested_leaves_.clear();
    requested_patches_.clear();
    if (next_initial_ < initial_.size()) {
        request_initial();
        return;
    }
    if (absolute_error_ <= tolerance_ * .2 * max_norm(integral_)) {
        auto append = [&](const Leaf &l) {
            result_.maximum_depth = std::max(result_.maximum_depth, l.depth);
            result_.patches.push_back(to_patch(l));
        };
        for (const auto &l : at_limit_)
            append(l);
        while (!pending_.empty()) {
            append(pending_.front());
            std::pop_heap(pending_.begin(), pending_.end(), by_error);
            pending_.pop_back();
        }
        const double ref = max_norm(integral_);
        if (!(ref > 0))
            throw std::runtime_error("near pair: no finite nonzero scale to measure the error against");
        result_.estimated_relative_error = absolute_error_ / ref;
        if (!std::isfinite(result_.estimated_relative_error) || result_.estimated_relative_error > tolerance_)
            throw std::runtime_error("near pair: refinement ended above the requested tolerance");
        finished_ = true;
        return;
    }
    if (pending_.empty())
        throw std::runtime_error(
            "Near-pair quadrature did not attain the requested error tolerance before the subdivision limit.");
    const Leaf worst = pending_.front();
    std::pop_heap(pending_.begin(), pending_.end(), by_error);
    pending_.pop_back();
    absolute_error_ = std::max(0.0, absolute_error_ - worst.error);
    for (unsigned i = 0; i < 36; ++i)
        integral_[i] -= worst.integral[i];
    if (projection_ && worst.inner_error > worst.outer_error) {
        requested_leaves_.push_back(make_leaf(worst.test, worst.trial, worst.depth, worst.degree, worst.intervals * 2));
    } else if (worst.degree == 8 && (worst.depth == max_depth_ || (promote_far_root_ && worst.depth == 0))) {
        requested_leaves_.push_back(make_leaf(worst.test, worst.trial, worst.depth, 14, worst.intervals));
    } else {
        if (pending_.size() + at_limit_.size() + 4 > 50000)
            throw std::runtime_error("near pair: refinement needs more than 50000 patches");
        auto split = [](const Sub &s) {
            const Bary a = s.v[0], b = s.v[1], c = s.v[2];
            const Bary ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            const double f = s.fraction * .25;
            return std::array<Sub, 4>{{{{a, ab, ca}, f}, {{ab, b, bc}, f}, {{ca, bc, c}, f}, {{ab, bc, ca}, f}}};
        };
        if (projection_ || worst.test_diameter >= worst.trial_diameter) {
            for (const auto &child : split(worst.test))
                requested_leaves_.push_back(make_leaf(child, worst.trial, worst.depth + 1, 8, worst.intervals));
        } else {
            for (const auto &child : split(worst.trial))
                requested_leaves_.push_back(make_leaf(worst.test, child, worst.depth + 1, 8));
        }
    }
    for (const auto &l : requested_leaves_)
        requested_patches_.push_back(to_patch(l));
}

void AdaptivePair::commit(const std::vector<PatchEstimate> &estimates) {
    if (finished_ || estimates.size() != requested_leaves_.size())
        throw std::invalid_argument("near pair: internal patch bookkeeping is inconsistent");
    for (const auto &e : estimates) {
        if (!(e.absolute_error >= 0) || !std::isfinite(e.absolute_error) || !(e.inner_error >= 0) ||
            !std::isfinite(e.inner_error) || !(e.outer_error >= 0) || !std::isfinite(e.outer_error))
            throw std::runtime_error("near pair: an error estimate is not finite");
        (void)max_norm(e.integral);
    }
    const auto leaves = requested_leaves_;
    for (std::size_t i = 0; i < leaves.size(); ++i) {
        Leaf l = leaves[i];
        l.integral = estimates[i].integral;
        l.error = estimates[i].absolute_error;
        l.inner_error = estimates[i].inner_error;
        l.outer_error = estimates[i].outer_error;
        include(std::move(l));
    }
    advance();
}

PatchEstimator host_patch_estimator(const Space &space, double k, unsigned threads) {
    return [&space, k, threads](const std::vector<PatchRequest> &requests, std::vector<PatchEstimate> &estimates) {
        estimates.resize(requests.size());
        const unsigned workers = static_cast<unsigned>(std::min<std::size_t>(
            threads ? threads : std::max(1u, std::thread::hardware_concurrency()), requests.size()));
        std::atomic<std::size_t> next{0};
        std::exception_ptr failure;
        std::mutex failure_mutex;
        auto run = [&] {
            for (;;) {
                const std::size_t i = next.fetch_add(1, std::memory_order_relaxed);
                if (i >= requests.size())
                    return;
                try {
                    const PatchRequest &r = requests[i];
                    estimates[i] = estimate_patch(space.panels[r.test], space.panels[r.trial], r.patch, k);
                } catch (const std::exception &e) {
                    std::lock_guard lock(failure_mutex);
                    if (!failure)
                        failure = std::make_exception_ptr(PatchEstimateError(i, e.what()));
                    next = requests.size();
                    return;
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned w = 1; w < workers; ++w)
            pool.emplace_back(run);
        run();
        for (auto &t : pool)
            t.join();
        if (failure)
            std::rethrow_exception(failure);
    };
}

AdaptiveResult adaptive_pair_rule(const Panel &test, const Panel &trial, double tolerance, unsigned max_depth,
                                  double k) {
    AdaptivePair state(test, trial, tolerance, max_depth, k);
    std::vector<PatchEstimate> estimates;
    while (!state.done()) {
        estimates.clear();
        for (const auto &p : state.requests())
            estimates.push_back(estimate_patch(test, trial, p, k));
        state.commit(estimates);
    }
    return state.take();
}

} // namespace hsbem
