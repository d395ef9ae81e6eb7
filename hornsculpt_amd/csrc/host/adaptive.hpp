// Near-pair geometry and the adaptive patch rule for disjoint panel pairs.
//
// Restates modules/bem/galerkin/pair_quadrature.cpp:43-683:
//   triangle-triangle distance and near test        :43-101, :786-807
//   subdivision, 36 P1/P1 moments of G and grad G   :103-197
//   projection partition / use / moments            :156-327
//   the largest-error state machine (AdaptiveRule)  :329-575
//   estimatePairPatchCpu (degree 6 vs 8, 12 vs 14)  :633-683
// The request/commit interface lets a batched estimator (host threads or a
// device kernel) drive many pairs; the decisions depend only on the estimates.
#pragma once

#include "rules.hpp"
#include "space.hpp"

#include <array>
#include <complex>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace hsbem {

double panel_distance(const Panel &test, const Panel &trial);
bool is_near_pair(const Panel &test, const Panel &trial, double distance_ratio);

struct PatchEstimate {
    std::array<std::complex<double>, 36> integral{};
    double absolute_error{}, inner_error{}, outer_error{};
};

PatchEstimate estimate_patch(const Panel &test, const Panel &trial, const PairPatch &patch, double k);

// One patch estimate of a batch: the panels (indices into the Space) and the patch.
struct PatchRequest {
    std::uint32_t test{}, trial{};
    PairPatch patch;
};

// A batched estimator (AdaptivePairBatch's role, adaptive_patch_estimator.hpp:62-86):
// estimates[i] receives the estimate of requests[i] at the wavenumber it was built
// for, with estimate_patch's meaning and refusals (std::runtime_error).
using PatchEstimator =
    std::function<void(const std::vector<PatchRequest> &requests, std::vector<PatchEstimate> &estimates)>;

// A refusal of one request of a batch (its index and estimate_patch's message).
struct PatchEstimateError : std::runtime_error {
    PatchEstimateError(std::size_t index, const std::string &what) : std::runtime_error(what), request(index) {}
    std::size_t request;
};

// The host estimator as a batch: estimate_patch on up to `threads` threads (0: all).
PatchEstimator host_patch_estimator(const Space &space, double k, unsigned threads = 0);

struct AdaptiveResult {
    std::vector<PairPatch> patches;
    double estimated_relative_error{-1};
    unsigned maximum_depth{};
};

class AdaptivePair {
  public:
    AdaptivePair(const Panel &test, const Panel &trial, double tolerance, unsigned max_depth, double k);
    // Patches whose estimates the next commit must supply, in order (at most 4).
    const std::vector<PairPatch> &requests() const { return requested_patches_; }
    void commit(const std::vector<PatchEstimate> &estimates);
    bool done() const { return finished_; }
    AdaptiveResult take() { return std::move(result_); }

  private:
    struct Sub {
        std::array<std::array<double, 3>, 3> v{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
        double fraction{1};
    };
    struct Leaf {
        Sub test, trial;
        std::array<std::complex<double>, 36> integral{};
        double error{}, test_diameter{}, trial_diameter{}, inner_error{}, outer_error{};
        unsigned depth{}, degree{8}, intervals{1};
    };
    Leaf make_leaf(const Sub &a, const Sub &b, unsigned depth, unsigned degree, unsigned intervals = 1) const;
    PairPatch to_patch(const Leaf &leaf) const;
    void request_initial();
    void include(Leaf leaf);
    void advance();

    const Panel &test_, &trial_;
    double tolerance_, k_;
    unsigned max_depth_;
    bool projection_{}, promote_far_root_{}, finished_{};
    std::array<std::complex<double>, 36> integral_{};
    double absolute_error_{};
    std::vector<Leaf> pending_, at_limit_; // pending_ is a max-heap on error
    std::vector<Leaf> requested_leaves_;
    std::vector<PairPatch> requested_patches_;
    std::vector<Sub> initial_;
    unsigned next_initial_{};
    AdaptiveResult result_;
};

// Runs an AdaptivePair to completion with host estimates.
AdaptiveResult adaptive_pair_rule(const Panel &test, const Panel &trial, double tolerance, unsigned max_depth,
                                  double k);

} // namespace hsbem
