// Per-level expansion plan of the wideband FMM path.
//
// Restates modules/bem/fmm/wideband_transfer.hpp:57-72 / wideband_transfer.cpp:266-310
// (makeWidebandLevelPlan) and support_level_plan.hpp:8-44 / support_level_plan.cpp:15-180
// (makeSupportExpansionPlan, enforceSupportDiagonalStability):
//   * order of a level: Gumerov & Duraiswami, J. Acoust. Soc. Am. 125 (2009),
//     Eq. 18, blending the geometric bound low = 1 - (ln eps + 1.5 ln(1 - 1/delta))/ln delta
//     and the oscillatory bound high = kr + 0.5 (-3 ln eps)^(2/3) (kr)^(1/3) as
//     p = max * ((low/max)^4 + (high/max)^4)^(1/4); signature order ceil(p),
//     translation and quadrature order 2 ceil(p); plane waves (HF) when p - high <= 2,
//     regular spherical (LF) otherwise;
//   * convergence ratios of far pairs delta_s = (D - r_t)/r_s, delta_t = (D - r_s)/r_t,
//     inherited downward (cumulative minimum) and a coarseward monotone radius
//     envelope;
//   * the diagonal translation is replaced by the stable spherical one when
//     eps_machine * kD sum_l (2l+1) |h_l^(2)(kD)| exceeds 1 % of the seed.
#pragma once

#include "support_tree.hpp"

#include <cstdint>
#include <vector>

namespace hsbem {

enum class LevelRepresentation : std::uint32_t { inactive = 0, regular_spherical = 1, plane_wave = 2 };

inline constexpr unsigned kMaximumSphericalOrder = 256;
inline constexpr double kMaximumSphericalElectricalRadius = 128;

struct WidebandLevelPlan {
    LevelRepresentation representation{LevelRepresentation::inactive};
    double active_radius{}, electrical_radius{}, accuracy_seed{};
    unsigned signature_order{}, translation_order{}, quadrature_order{};
    std::uint64_t spherical_coefficients{}, angular_samples{};
};

// Throws std::invalid_argument on non-physical inputs, std::runtime_error when
// the order leaves the supported angular range.
WidebandLevelPlan make_wideband_level_plan(double active_radius, double k, double separation, double accuracy);

// Returns the conditioning indicator; switches a plane-wave level to the
// spherical form when the indicator exceeds 1 % of the accuracy seed.
double enforce_diagonal_stability(WidebandLevelPlan &plan, double k, double minimum_far_distance);

struct SupportLevelExpansion {
    WidebandLevelPlan expansion;
    double source_separation{}, target_separation{}, policy_separation{};
    double max_active_radius{}, minimum_far_distance{}, diagonal_condition{};
};

struct SupportExpansionPlan {
    std::vector<SupportLevelExpansion> levels;
    std::vector<unsigned> active; // per node: kPressureOutgoing | kFluxOutgoing | kPressureIncoming
};

// panel_face_column: per panel, the active face column or UINT32_MAX.
SupportExpansionPlan make_support_expansion_plan(const SupportTree &tree,
                                                 const std::vector<std::uint32_t> &panel_face_column, double k,
                                                 double accuracy_seed);

} // namespace hsbem
