#include "plan.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <limits>
#include <memory>
#include <mutex>
#include <numbers>
#include <set>
#include <sstream>
#include <stdexcept>
#include <thread>
#include <tuple>

namespace hsbem {
namespace {

using Pair = std::pair<std::uint32_t, std::uint32_t>;

std::uint32_t checked32(std::size_t v) {
    if (v > std::numeric_limits<std::uint32_t>::max())
        throw std::runtime_error("interaction plan is too large for 32-bit indices");
    return static_cast<std::uint32_t>(v);
}

// Classification with the consistency checks of the reference classifyPair
// (pair_quadrature.cpp:771-784).
Permutation classify_checked(const Panel &t, const Panel &s) {
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = i + 1; j < 3; ++j)
            if (t.nodes[i] == t.nodes[j] || s.nodes[i] == s.nodes[j])
                throw std::invalid_argument("a triangle repeats a vertex index");
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j)
            if (t.nodes[i] == s.nodes[j] && vnorm2(vsub(t.vertices[i], s.vertices[j])) != 0)
                throw std::invalid_argument("one vertex index is used with two different positions");
    return classify_pair_ids(t.nodes.data(), s.nodes.data());
}

// Disjoint candidate pairs (a<b) whose vertex-0 separation is within
// ratio*h + h_a + h_b on every axis, found on a uniform grid of 27 neighbours.
std::vector<Pair> spatial_candidates(const Space &space, double ratio) {
    double h_max = 0;
    for (const auto &p : space.panels) {
        if (!(p.max_edge > 0) || !std::isfinite(p.max_edge))
            throw std::invalid_argument("interaction plan: a panel has an invalid size");
        h_max = std::max(h_max, p.max_edge);
    }
    const Vec3 origin = space.panels.front().vertices[0];
    double extent = 0;
    for (const auto &p : space.panels) {
        const Vec3 v = p.vertices[0];
        extent = std::max({extent, std::abs(v.x - origin.x), std::abs(v.y - origin.y), std::abs(v.z - origin.z)});
    }
    const double width = (ratio + 2) * h_max + 128 * std::numeric_limits<double>::epsilon() * extent;
    if (!(width > 0) || !std::isfinite(width))
        throw std::invalid_argument("interaction plan: candidate grid spacing is not finite");
    using Cell = std::tuple<std::int64_t, std::int64_t, std::int64_t>;
    auto cell_of = [&](Vec3 v) {
        auto axis = [&](double value, double base) {
            const double c = std::floor((value - base) / width);
            constexpr double limit = static_cast<double>(std::numeric_limits<std::int64_t>::max()) / 2;
            if (!std::isfinite(c) || c < -limit || c > limit)
                throw std::invalid_argument("interaction plan: candidate grid index out of range");
            return static_cast<std::int64_t>(c);
        };
        return Cell{axis(v.x, origin.x), axis(v.y, origin.y), axis(v.z, origin.z)};
    };
    const std::size_t n = space.panels.size();
    std::vector<std::pair<Cell, std::uint32_t>> grid(n);
    std::vector<Cell> where(n);
    for (std::size_t i = 0; i < n; ++i) {
        where[i] = cell_of(space.panels[i].vertices[0]);
        grid[i] = {where[i], checked32(i)};
    }
    std::sort(grid.begin(), grid.end());
    std::vector<Pair> out;
    for (std::size_t i = 0; i < n; ++i) {
        const auto [cx, cy, cz] = where[i];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const Cell key{cx + dx, cy + dy, cz + dz};
                    auto it = std::lower_bound(grid.begin(), grid.end(), key,
                                               [](const auto &e, const Cell &c) { return e.first < c; });
                    for (; it != grid.end() && it->first == key; ++it) {
                        const std::uint32_t j = it->second;
                        if (j <= i)
                            continue;
                        const Panel &a = space.panels[i], &b = space.panels[j];
                        const double bound = ratio * std::max(a.max_edge, b.max_edge) + a.max_edge + b.max_edge;
                        if (std::abs(a.vertices[0].x - b.vertices[0].x) > bound ||
                            std::abs(a.vertices[0].y - b.vertices[0].y) > bound ||
                            std::abs(a.vertices[0].z - b.vertices[0].z) > bound)
                            continue;
                        if (classify_pair_ids(a.nodes.data(), b.nodes.data()).topology == Topology::disjoint)
                            out.emplace_back(checked32(i), j);
                    }
                }
    }
    std::sort(out.begin(), out.end());
    return out;
}

struct PlannedPair {
    Pair pair;
    AdaptiveResult rule;
};

} // namespace

double planning_wavenumber(const std::vector<double> &solved, const std::vector<double> &planned, double c) {
    if (solved.empty() || !std::isfinite(c) || c <= 0)
        throw std::invalid_argument("interaction planning needs frequencies and a sound speed");
    std::set<double> raster;
    double f_max = 0;
    for (double f : planned) {
        if (!std::isfinite(f) || f <= 0 || !raster.insert(f).second)
            throw std::invalid_argument("planning frequencies must be finite, positive and distinct");
        f_max = std::max(f_max, f);
    }
    for (double f : solved) {
        if (!std::isfinite(f) || f <= 0 || (!planned.empty() && !raster.count(f)))
            throw std::invalid_argument("each solved frequency must be one of the planning frequencies");
        if (planned.empty())
            f_max = std::max(f_max, f);
    }
    const double k = 2 * std::numbers::pi * f_max / c;
    if (!std::isfinite(k))
        throw std::invalid_argument("planning wavenumber is not finite");
    return k;
}

namespace {

// The message of a near pair's refusal, with the pair's identity.
std::string near_failure(const Space &space, Pair pr, double k_max, unsigned depth, const char *what) {
    const Panel &a = space.panels[pr.first], &b = space.panels[pr.second];
    std::ostringstream msg;
    msg << "near panels " << pr.first << '/' << pr.second << " (faces " << a.face_id << '/' << b.face_id
        << ", k=" << k_max << ", gap=" << panel_distance(a, b) << " m, depth=" << depth << "): " << what;
    return msg.str();
}

// Near pairs refined together (AdaptivePairBatch, adaptive_patch_estimator.hpp:62-86):
// up to kBatchPairs pairs at a time; each round collects the pending patches of every
// unfinished pair, estimates them in one call and commits each pair's estimates
// (on up to `workers` threads: the pairs are independent). `done` receives each pair
// and its rule.
constexpr std::size_t kBatchPairs = std::size_t(1) << 15;

// fn(i) for i < count on up to `workers` threads; the first exception thrown is rethrown
// (the remaining items are skipped).
template <class Fn> void parallel_items(std::size_t count, unsigned workers, Fn &&fn) {
    std::atomic<std::size_t> next{0};
    std::exception_ptr failure;
    std::mutex failure_mutex;
    auto run = [&] {
        for (;;) {
            const std::size_t i = next.fetch_add(1, std::memory_order_relaxed);
            if (i >= count)
                return;
            try {
                fn(i);
            } catch (...) {
                std::lock_guard lock(failure_mutex);
                if (!failure)
                    failure = std::current_exception();
                next = count;
                return;
            }
        }
    };
    const unsigned threads = static_cast<unsigned>(std::min<std::size_t>(std::max(workers, 1u), count));
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < threads; ++w)
        pool.emplace_back(run);
    run();
    for (auto &t : pool)
        t.join();
    if (failure)
        std::rethrow_exception(failure);
}

template <class Done>
void refine_batched(const Space &space, const QuadratureSettings &s, double k_max, std::vector<Pair> pairs,
                    const PatchEstimator &estimator, unsigned workers, Done &&done) {
    std::sort(pairs.begin(), pairs.end());
    std::vector<PatchRequest> requests;
    std::vector<PatchEstimate> estimates;
    for (std::size_t first = 0; first < pairs.size(); first += kBatchPairs) {
        const std::size_t count = std::min(kBatchPairs, pairs.size() - first);
        std::vector<std::unique_ptr<AdaptivePair>> states(count);
        parallel_items(count, workers, [&](std::size_t i) {
            const Pair pr = pairs[first + i];
            try {
                states[i] = std::make_unique<AdaptivePair>(space.panels[pr.first], space.panels[pr.second],
                                                           s.near_tolerance, s.near_max_depth, k_max);
            } catch (const std::exception &e) {
                throw std::runtime_error(near_failure(space, pr, k_max, s.near_max_depth, e.what()));
            }
        });
        std::vector<std::size_t> active(count), offsets;
        for (std::size_t i = 0; i < count; ++i)
            active[i] = i;
        while (!active.empty()) {
            requests.clear();
            offsets.assign(1, 0);
            for (const std::size_t i : active) {
                const Pair pr = pairs[first + i];
                for (const auto &patch : states[i]->requests())
                    requests.push_back({pr.first, pr.second, patch});
                offsets.push_back(requests.size());
            }
            estimates.assign(requests.size(), PatchEstimate{});
            try {
                estimator(requests, estimates);
            } catch (const PatchEstimateError &e) {
                const auto at = std::upper_bound(offsets.begin(), offsets.end(), e.request) - offsets.begin() - 1;
                throw std::runtime_error(
                    near_failure(space, pairs[first + active[at]], k_max, s.near_max_depth, e.what()));
            }
            parallel_items(active.size(), workers, [&](std::size_t a) {
                const std::size_t i = active[a];
                const std::vector<PatchEstimate> mine(estimates.begin() + static_cast<std::ptrdiff_t>(offsets[a]),
                                                      estimates.begin() + static_cast<std::ptrdiff_t>(offsets[a + 1]));
                try {
                    states[i]->commit(mine);
                } catch (const std::exception &e) {
                    throw std::runtime_error(near_failure(space, pairs[first + i], k_max, s.near_max_depth, e.what()));
                }
            });
            std::vector<std::size_t> still;
            for (const std::size_t i : active)
                if (!states[i]->done())
                    still.push_back(i);
            active = std::move(still);
        }
        for (std::size_t i = 0; i < count; ++i)
            done(pairs[first + i], states[i]->take());
    }
}

} // namespace

InteractionPlan make_interaction_plan(const Space &space, const QuadratureSettings &s, double k_max,
                                      unsigned threads, const PatchEstimator *estimator) {
    if (space.panels.empty() || space.vertices.empty() || !std::isfinite(k_max) || k_max < 0 ||
        !std::isfinite(s.near_distance_ratio) || s.near_distance_ratio < 0 || !std::isfinite(s.near_tolerance) ||
        s.near_tolerance <= 0 || s.near_max_depth > 16 || (s.regular_order != 2 && s.regular_order != 4) ||
        s.singular_order < 2 || s.singular_order > 20 || space.incidence_offsets.size() != space.vertices.size() + 1)
        throw std::invalid_argument("invalid interaction-plan input or quadrature settings");
    const std::size_t n = space.panels.size();
    checked32(n);
    InteractionPlan plan;
    std::vector<PanelBounds> bounds;
    bounds.reserve(n);
    // Same comparison as the qualifier (multiplication, not a divided threshold).
    bool all_pairs = !((1e-6 + 4 * 3.5e-8) <= .2 * s.near_tolerance);
    for (const auto &p : space.panels) {
        bounds.push_back(make_panel_bounds(p));
        all_pairs = all_pairs || !bounds.back().qualified_shape || k_max * p.max_edge > 2;
    }
    plan.all_pairs = all_pairs;
    if (all_pairs)
        plan.whole_panel_degrees.assign(n * (n - 1) / 2, 0);
    plan.near_offsets.assign(n + 1, 0);
    for (std::size_t t = 0; t < n; ++t) {
        std::vector<std::uint32_t> adjacent;
        for (const auto node : space.panels[t].nodes)
            for (auto o = space.incidence_offsets[node]; o < space.incidence_offsets[node + 1]; ++o)
                adjacent.push_back(space.incidences[o].panel);
        std::sort(adjacent.begin(), adjacent.end());
        adjacent.erase(std::unique(adjacent.begin(), adjacent.end()), adjacent.end());
        for (const auto u : adjacent)
            plan.touching.push_back({checked32(t), u, classify_checked(space.panels[t], space.panels[u])});
    }
    const std::vector<Pair> pairs =
        all_pairs ? std::vector<Pair>{} : spatial_candidates(space, std::max(2., s.near_distance_ratio));
    plan.candidate_pairs = pairs.size();
    // Up to 16 host threads (the reference caps its planner at 8): one MI355X's share of a
    // host's cores; the result is sorted, so it does not depend on the thread count.
    unsigned workers = threads ? threads : std::max(1u, std::thread::hardware_concurrency());
    workers = static_cast<unsigned>(std::min<std::size_t>({workers, 16, all_pairs ? n : pairs.size()}));
    if (!workers)
        return plan;

    std::vector<PlannedPair> planned;
    std::mutex out_mutex;
    std::atomic<std::size_t> next{0}, streamed{0};
    std::atomic<bool> failed{false};
    std::exception_ptr failure;
    std::mutex failure_mutex;
    std::vector<Pair> batched;
    auto record = [&](Pair pr, AdaptiveResult rule) {
        std::lock_guard lock(out_mutex);
        if (all_pairs && rule.maximum_depth == 0 && rule.patches.size() == 1 && !rule.patches.front().projection) {
            plan.whole_panel_degrees[unordered_pair_index(pr.first, pr.second)] =
                static_cast<std::uint8_t>(rule.patches.front().degree);
            plan.maximum_estimated_relative_error =
                std::max(plan.maximum_estimated_relative_error, rule.estimated_relative_error);
            return;
        }
        planned.push_back({pr, std::move(rule)});
    };
    auto process = [&](Pair pr) {
        const Panel &a = space.panels[pr.first], &b = space.panels[pr.second];
        if (classify_checked(a, b).topology != Topology::disjoint)
            return;
        if (all_pairs)
            ++streamed;
        const unsigned degree =
            qualified_regular_degree(bounds[pr.first], bounds[pr.second], k_max, s.near_tolerance, s.regular_order);
        const double h = std::max(a.max_edge, b.max_edge);
        const double limit = bounds[pr.first].radius + bounds[pr.second].radius + s.near_distance_ratio * h;
        const double dx = a.centroid.x - b.centroid.x, dy = a.centroid.y - b.centroid.y,
                     dz = a.centroid.z - b.centroid.z;
        const bool geometric_near = degree && s.near_distance_ratio > 0 && dx * dx + dy * dy + dz * dz < limit * limit &&
                                    is_near_pair(a, b, s.near_distance_ratio);
        if (degree && !geometric_near)
            return;
        if (estimator) { // refined in batches below
            std::lock_guard lock(out_mutex);
            batched.push_back(pr);
            return;
        }
        AdaptiveResult rule;
        try {
            rule = adaptive_pair_rule(a, b, s.near_tolerance, s.near_max_depth, k_max);
        } catch (const std::exception &e) {
            throw std::runtime_error(near_failure(space, pr, k_max, s.near_max_depth, e.what()));
        }
        record(pr, std::move(rule));
    };
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < workers; ++w)
        pool.emplace_back([&] {
            try {
                for (;;) {
                    if (failed.load(std::memory_order_relaxed))
                        return;
                    const std::size_t i = next.fetch_add(1, std::memory_order_relaxed);
                    if (i >= (all_pairs ? n : pairs.size()))
                        return;
                    if (all_pairs) {
                        for (std::size_t j = i + 1; j < n; ++j)
                            process({checked32(i), checked32(j)});
                    } else {
                        process(pairs[i]);
                    }
                }
            } catch (...) {
                std::lock_guard lock(failure_mutex);
                if (!failure)
                    failure = std::current_exception();
                failed = true;
            }
        });
    for (auto &t : pool)
        t.join();
    if (failure)
        std::rethrow_exception(failure);
    if (estimator)
        refine_batched(space, s, k_max, std::move(batched), *estimator, workers, record);
    if (all_pairs)
        plan.candidate_pairs = streamed.load();
    std::sort(planned.begin(), planned.end(), [](const auto &x, const auto &y) { return x.pair < y.pair; });
    for (auto &e : planned) {
        const auto first = checked32(plan.patches.size());
        const auto count = checked32(e.rule.patches.size());
        plan.patches.insert(plan.patches.end(), e.rule.patches.begin(), e.rule.patches.end());
        checked32(plan.patches.size());
        plan.near.push_back({e.pair.first, e.pair.second, first, count, 0});
        plan.near.push_back({e.pair.second, e.pair.first, first, count, 1});
        plan.maximum_depth = std::max(plan.maximum_depth, e.rule.maximum_depth);
        plan.maximum_estimated_relative_error =
            std::max(plan.maximum_estimated_relative_error, e.rule.estimated_relative_error);
    }
    checked32(plan.near.size());
    std::sort(plan.near.begin(), plan.near.end(),
              [](const auto &x, const auto &y) { return std::pair{x.test, x.trial} < std::pair{y.test, y.trial}; });
    plan.near_trials.reserve(plan.near.size());
    for (const auto &j : plan.near) {
        ++plan.near_offsets[j.test + 1];
        plan.near_trials.push_back(j.trial);
    }
    for (std::size_t r = 1; r < plan.near_offsets.size(); ++r)
        plan.near_offsets[r] += plan.near_offsets[r - 1];
    return plan;
}

} // namespace hsbem
