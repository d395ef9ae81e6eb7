// Near-field DOF plan of the wideband FMM path over a support tree.
//
// Restates modules/bem/fmm/support_near_plan.hpp:8-72 and
// support_near_plan.cpp:30-387 (makeSupportNearPlan, SupportNearJobIterator):
//   pressure CSR  rows = P1 nodes, columns = sorted P1 nodes of every source
//                 leaf in a near leaf pair of the pressure-pressure partition;
//   flux CSR      rows = P1 nodes, columns = sorted active-face columns met in
//                 the near flux leaves of the row's leaf;
//   leaf panels   per node, the sorted, deduplicated panels of its P1 supports;
//   flux blocks   per target leaf, its near pressure-flux interactions;
//   jobs          per test panel, every source panel touching a near entry,
//                 with the 3x3 local P1/P1 mask and the 3-bit P1/DP0 mask, in
//                 first-touch order (the iterator's order).
// The device near operator assembles exactly these entries.
#pragma once

#include "support_tree.hpp"

#include <cstdint>
#include <vector>

namespace hsbem {

struct SupportCsr {
    std::vector<std::uint32_t> row_offsets, columns;
};

struct SupportNearJob {
    std::uint32_t test{}, source{};
    std::uint16_t pressure_mask{}; // bit 3*test_local + source_local
    std::uint8_t flux_mask{};      // bit test_local
};

struct SupportNearPlan {
    SupportCsr pressure, flux;
    std::vector<std::uint32_t> active_faces;
    std::vector<std::uint32_t> panel_face_column; // UINT32_MAX: not an active driven panel
    std::vector<std::uint32_t> node_panel_offsets, support_panels;
    std::vector<std::uint32_t> flux_block_offsets, flux_block_indices;
};

// active_faces: face ids whose panels carry flux columns, in column order. Throws
// std::invalid_argument on duplicated or wall faces, faces without panels, or a
// tree that does not belong to the space.
SupportNearPlan make_support_near_plan(const Space &space, const SupportTree &tree,
                                       const std::vector<std::uint32_t> &active_faces);

// All near panel-pair jobs of the plan in iterator order.
std::vector<SupportNearJob> support_near_jobs(const Space &space, const SupportTree &tree, const SupportNearPlan &plan);

} // namespace hsbem
