// Dyadic support octree and near/far partition of the wideband FMM path.
//
// Restates modules/bem/fmm/support_tree.hpp:11-104 and support_tree.cpp:10-414
// (makeSupportTree): P1 pressure supports (a vertex anchor with the bounds of
// its whole incident panel star) and DP0 flux supports (a panel centroid
// anchor with the panel's bounds) are sorted into one octree over the root
// cube of all panel bounds; anchors choose octants (>= center), complete
// support bounds and vertex balls (outward-rounded) decide the multipole
// acceptance. The dual-tree partition marks (target, source) node pairs far
// when they sit on one level with a positive bounds gap above
// nearRatio * h and a centre distance above separation * (r_t + r_s); leaf
// pairs that are not far are near. Interaction lists, node numbering (preorder)
// and level statistics match the reference plan exactly, which the FMM
// translation plans key on.
#pragma once

#include "space.hpp"

#include <array>
#include <cstdint>
#include <limits>
#include <vector>

namespace hsbem {

inline constexpr std::uint32_t kNoSupportNode = UINT32_MAX;
inline constexpr unsigned kPressureOutgoing = 1, kFluxOutgoing = 2, kPressureIncoming = 4;

struct SupportBox {
    Vec3 lower{}, upper{};
    double max_edge{};
};

struct SupportAnchor {
    Vec3 anchor{};
    SupportBox box{};
};

struct SupportNode {
    Vec3 center{};
    double half_width{};
    std::array<std::uint32_t, 3> coords{};
    std::uint32_t level{}, parent{kNoSupportNode};
    std::array<std::uint32_t, 8> children{kNoSupportNode, kNoSupportNode, kNoSupportNode, kNoSupportNode,
                                          kNoSupportNode, kNoSupportNode, kNoSupportNode, kNoSupportNode};
    std::uint32_t first_pressure{}, pressure_count{}, first_flux{}, flux_count{};
    SupportBox pressure_box{}, flux_box{};
    double pressure_radius{}, flux_radius{}; // expansion balls about center (nested upward)
    unsigned active{};                       // kPressureOutgoing | kFluxOutgoing | kPressureIncoming
    bool leaf() const {
        for (auto c : children)
            if (c != kNoSupportNode)
                return false;
        return true;
    }
};

struct SupportPair {
    std::uint32_t target{}, source{};
};

struct SupportLevelPair {
    std::uint32_t target_level{}, source_level{};
    std::uint64_t interactions{}, represented_pairs{};
};

struct SupportLists {
    std::vector<SupportPair> near, far;         // near: leaf pairs; far: equal levels
    std::vector<SupportLevelPair> far_levels;   // ordered by (target level, source level)
    std::uint64_t represented_near{}, represented_far{};
};

struct SupportLevelStats {
    std::vector<std::uint32_t> nodes;
    double max_radius{}, max_active_radius{};
    std::uint64_t pressure_outgoing{}, flux_outgoing{}, pressure_incoming{};
    std::uint64_t pp_far{}, pf_far{};
    std::uint64_t pp_offsets{}, pf_offsets{}, union_offsets{};
    std::uint64_t pp_distance_classes{}, pf_distance_classes{};
    std::uint64_t pp_polar_classes{}, pf_polar_classes{}, union_polar_classes{};
    std::vector<std::uint64_t> squared_offsets; // sorted exact |offset|^2 classes (union)
};

struct SupportTree {
    double separation{}, near_ratio{};
    std::vector<SupportAnchor> pressure_supports, flux_supports;
    std::vector<std::uint32_t> pressure_order, flux_order, pressure_leaf, flux_leaf;
    std::vector<SupportNode> nodes;
    std::vector<SupportLevelStats> levels;
    SupportLists pressure_pressure, pressure_flux;
};

struct SupportTreeSettings {
    unsigned leaf_capacity{32}, maximum_depth{20};
    double separation{1.5}, near_ratio{2};
    std::uint64_t maximum_interactions{std::numeric_limits<std::uint64_t>::max()};
};

// Throws std::invalid_argument on bad settings or geometry, std::runtime_error
// when the interaction budget is exceeded.
SupportTree make_support_tree(const Space &space, const SupportTreeSettings &settings = {});

} // namespace hsbem
