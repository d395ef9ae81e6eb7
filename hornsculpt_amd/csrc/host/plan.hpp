// Dense interaction plan: which ordered panel pairs are touching (Sauter-Schwab),
// near (adaptive patches) or regular (qualified tensor rule at the solve k).
//
// Mirrors modules/bem/galerkin/interaction_plan.hpp:19-71 and
// interaction_plan.cpp:57-433 (spatial candidates on a grid of width
// (max(2,ratio)+2)*h_max, per-pair qualification at the planning wavenumber,
// near jobs in both directions sharing one patch list, sorted near CSR, and
// the all-pairs mode with one accepted whole-panel degree byte per pair).
#pragma once

#include "adaptive.hpp"
#include "qualification.hpp"

#include <cstdint>
#include <vector>

namespace hsbem {

struct QuadratureSettings {
    unsigned regular_order{4};        // 2 or 4 (minimum regular degree)
    unsigned singular_order{4};       // Gauss order per Sauter-Schwab coordinate, 2..20
    double near_distance_ratio{1.0};
    double near_tolerance{1e-4};
    unsigned near_max_depth{6};
};

struct TouchingJob {
    std::uint32_t test{}, trial{};
    Permutation permutation;
};

struct NearJob {
    std::uint32_t test{}, trial{}, first_patch{}, patch_count{}, swap{};
};

struct InteractionPlan {
    std::vector<TouchingJob> touching;
    std::vector<NearJob> near;
    std::vector<PairPatch> patches;
    std::vector<std::uint8_t> whole_panel_degrees; // all-pairs mode only; index = pair_index(a,b)
    std::vector<std::uint32_t> near_offsets, near_trials;
    std::size_t candidate_pairs{};
    unsigned maximum_depth{};
    double maximum_estimated_relative_error{};
    bool all_pairs{};
};

HS_HD inline std::size_t unordered_pair_index(std::uint32_t a, std::uint32_t b) {
    const std::size_t lo = a < b ? a : b, hi = a < b ? b : a;
    return hi * (hi - 1) / 2 + lo;
}

// Throws std::invalid_argument on invalid settings and std::runtime_error when a
// near pair cannot reach its tolerance. threads=0 uses up to sixteen host threads.
// With an estimator (built for k_max), the near pairs are refined together in
// batches, each round estimating every pending patch of every unfinished pair in
// one call; without one, each pair runs to completion on host estimates. Both give
// the same plan from the same estimates.
InteractionPlan make_interaction_plan(const Space &space, const QuadratureSettings &settings, double k_max,
                                      unsigned threads = 0, const PatchEstimator *estimator = nullptr);

// Planning wavenumber: 2 pi max(f)/c over the planning raster when given, else
// over the solved frequencies (planning_frequencies.hpp:14-38).
double planning_wavenumber(const std::vector<double> &solved, const std::vector<double> &planned, double c);

} // namespace hsbem
