// Matrix-free Burton-Miller pressure operator of the wideband FMM path on one GPU.
//
// Replaces SupportFmmOperator::applyPressureDevice and preconditionDevice
// (modules/bem/cuda/support_fmm_operator.hpp:84-90, support_fmm_operator.cu:1220-1296,
// support_preconditioner.cu:33-190): A p = A_near p + far field, where
//   A_near   exact near CSR of the support near plan (the near operator's entries);
//   far      leaf emission E p -> upward M2M -> M2L over the tree's far pairs and
//            L2L downward -> leaf testing T L, all in the low-frequency scaled
//            spherical basis of the level plan (orders from Eq. 18 at the accuracy
//            seed, basis radius = the level's active radius), every translation as
//            rotate / coaxial m-block plan / rotate back;
//   M^{-1}   block-Jacobi: explicit inverses of the 64 x 64 diagonal blocks of the
//            near operator in tree order.
// Numbering: original mesh vertices (global P1 numbering), complex FP64 vectors.
#pragma once

#include "hip/fmm_kernels.hpp"
#include "host/support_near_plan.hpp"
#include "host/support_tree.hpp"
#include "solver.hpp"

#include <complex>
#include <cstdint>
#include <vector>

namespace hsbem {

inline constexpr std::size_t kMaximumDirectAuditPanels = 20000;

struct FmmSettings {
    SupportTreeSettings tree{64, 20, 1.5, 2.0};
    double accuracy_seed{5e-7}; // Eq. 18 accuracy of the level orders
    unsigned endpoint_degree{};  // leaf panel rule degree; 0: the electrical minimum
    unsigned block_width{32};    // block-Jacobi width (preconditionerBlockWidth), power of two 32..2048
};

struct FmmStatistics {
    std::uint32_t nodes{}, levels{}, leaves{}, top_level{}, rule_points{}, endpoint_degree{};
    std::uint64_t near_nnz{}, far_pairs{}, m2m_links{}, m2l_links{}, l2l_links{}, rotations{}, axial_plans{};
    std::uint64_t expansion_values{}, moment_values{}, device_bytes{};
    std::uint64_t shared_bytes{}; // of device_bytes: near entries and preconditioner shared with later operators
    std::uint64_t plane_wave_levels{}, plane_wave_leaves{}, angular_links{}, diagonal_classes{};
    std::uint64_t plan_cache_hits{}, plans_built{}; // translation plans reused / built by this operator
    std::vector<std::uint32_t> level_orders, level_representation, level_rule_orders; // representation 1 LF, 2 HF
    double tree_ms{}, near_ms{}, plans_ms{}, moments_ms{}, near_kernel_ms{};
};

// Frequency-independent part of the near operator: support tree, near plan,
// the near jobs split by integration route and the CSR pattern, device resident.
struct NearStructure {
    SupportTree tree;
    SupportNearPlan plan;
    std::shared_ptr<DevicePlan> interaction; // Galerkin plan that classified the jobs
    DeviceBuffer<std::uint32_t> row_offsets, columns, flux_offsets, flux_columns, face_column;
    DeviceBuffer<std::uint32_t> regular, regular_masks, touching_masks, fan_masks, patch_masks;
    DeviceBuffer<TouchingJob> touching;
    DeviceBuffer<NearJob> fans;
    DeviceBuffer<dev::NearPatchItem> items;
    std::size_t nnz{}, fnnz{}, jobs{}, regular_count{}, touching_count{}, patch_count{}, fan_count{};
    std::uint64_t device_bytes() const {
        return row_offsets.capacity_bytes() + columns.capacity_bytes() + flux_offsets.capacity_bytes() +
               flux_columns.capacity_bytes() + face_column.capacity_bytes() + regular.capacity_bytes() +
               regular_masks.capacity_bytes() + touching_masks.capacity_bytes() + fan_masks.capacity_bytes() +
               patch_masks.capacity_bytes() + touching.capacity_bytes() + fans.capacity_bytes() +
               items.capacity_bytes();
    }
};

// Near entries at one wavenumber (values in the structure's CSR pattern).
struct DeviceNear {
    std::shared_ptr<const NearStructure> structure;
    DeviceBuffer<double> values, flux_values;
    std::size_t nnz{}, fnnz{};
    float kernel_ms{};
};

class FmmOperator {
  public:
    // y = A x (device pointers, nv complex each, y may not alias x).
    void apply(const double *x, double *y, hipStream_t stream) const;
    // The FP32-bulk form of the apply: expansions and translation workspaces in FP32
    // (the rotations, coaxial products and gathers run in FP32, the Wigner and coaxial
    // tables rounded when staged); the near part, the emission and testing sums and the
    // vectors stay FP64. Available for operators without plane-wave levels and with
    // orders up to 48 (fp32_available); its FP32 buffers are allocated on first use.
    bool fp32_available() const;
    void apply_fp32(const double *x, double *y, hipStream_t stream) const;
    // y = A x through a captured graph of the whole apply (M2L graph replay): the
    // graph runs on operator-owned input/output buffers, captured on first use on
    // the given (non-default) stream; x and y are copied in and out on the device.
    void apply_graph(const double *x, double *y, hipStream_t stream) const;
    // z = M^{-1} v (block-Jacobi of the near operator, blocks of the configured width).
    void precondition(const double *v, double *z, hipStream_t stream) const;
    // y = A x with the far part integrated directly over every far DOF pair (the
    // direct Galerkin audit, SupportDirectAudit::applyPressureDevice,
    // support_direct_audit.hpp:11-22): meshes of at most kMaximumDirectAuditPanels panels.
    void apply_direct(const double *x, double *y, hipStream_t stream) const;
    bool direct_available() const { return direct_leaf_count_ > 0; }
    // preferSupportDirectAudit (support_fmm_solve.cu:79-115): the direct audit when its
    // pair work is at most twice the multipole work of this operator's levels.
    bool prefer_direct() const;
    ~FmmOperator();
    std::uint32_t size() const { return n_; }
    const DeviceNear &near() const { return *near_; }
    const FmmStatistics &statistics() const { return stats_; }

  private:
    friend class MeshSolver;
    struct Launch {
        std::uint32_t first_destination{}, destination_count{}, first_link{}, link_count{}, first_chunk{},
            chunk_count{}, first_axial_chunk{}, axial_chunk_count{}, max_order{};
    };
    std::uint32_t n_{};
    FmmStatistics stats_;
    std::shared_ptr<const DeviceNear> near_;
    std::shared_ptr<const DeviceBuffer<double>> jacobi_;
    DeviceBuffer<dev::FmmLeaf> leaves_;
    DeviceBuffer<std::uint32_t> leaf_dofs_, order_;
    DeviceBuffer<double> emission_, testing_, pool_;
    DeviceBuffer<dev::FmmDestination> destinations_;
    DeviceBuffer<dev::FmmLink> links_;
    DeviceBuffer<dev::FmmChunk> chunks_;
    DeviceBuffer<std::uint32_t> link_order_, axial_order_;
    DeviceBuffer<dev::FmmChunk> axial_chunks_;
    mutable DeviceBuffer<double> work_y_, work_z_;
    DeviceBuffer<dev::FmmRotation> rotations_;
    DeviceBuffer<double> small_d_, azimuth_, axial_;
    DeviceBuffer<std::uint64_t> axial_offsets_;
    struct HfLaunch {
        dev::HfTransferKind kind{};
        std::uint32_t first_link{}, link_count{}, first_destination{}, destination_count{}, in_rings{}, out_rings{},
            out_samples{};
    };
    struct DiagonalLaunch {
        std::uint32_t first_destination{}, destination_count{}, first_link{};
        std::uint64_t samples{};
    };
    enum class StageKind : std::uint32_t { translate, angular, diagonal };
    struct Stage {
        StageKind kind;
        std::uint32_t index;
    };
    std::vector<Launch> translate_;
    std::vector<HfLaunch> angular_;
    std::vector<DiagonalLaunch> diagonal_;
    std::vector<Stage> stages_; // upward levels coarseward, then downward levels leafward
    DeviceBuffer<dev::HfRule> hf_rules_;
    DeviceBuffer<double> hf_polar_, hf_legendre_, diag_table_;
    DeviceBuffer<dev::HfLink> hf_links_;
    DeviceBuffer<dev::FmmDestination> hf_destinations_, diag_destinations_;
    DeviceBuffer<dev::HfDiagonalLink> diag_links_;
    mutable DeviceBuffer<double> hf_fourier_, hf_harmonic_, hf_field_;
    DeviceBuffer<dev::HfLeaf> hf_leaves_;
    DeviceBuffer<dev::HfLeafItem> hf_leaf_items_;
    DeviceBuffer<double> leaf_scale_, leaf_rule_;
    std::uint32_t hf_leaf_count_{}, hf_leaf_item_count_{}, hf_max_samples_{}, block_width_{64};
    std::uint32_t split_target_{2048}; // translations' stage-split target (autotuned per mesh and level orders)
    // direct audit: solver-owned geometry (alive while the operator is), leaf map, near matrix
    const double *geo_{}, *row_scale_{};
    const std::uint32_t *nodes_{};
    std::uint32_t nt_{}, direct_leaf_count_{};
    DeviceBuffer<std::uint32_t> vertex_leaf_;
    DeviceBuffer<std::uint8_t> near_leaves_;
    double direct_work_{}, multipole_work_{};
    double k_{};
    std::uint32_t leaf_count_{};
    mutable DeviceBuffer<float> pool32_, work_y32_, work_z32_; // the FP32-bulk form's buffers
    void apply_in(const double *x, double *y, hipStream_t stream, bool fp32) const;
    mutable DeviceBuffer<double> graph_x_, graph_y_;
    mutable hipGraph_t graph_{};
    mutable hipGraphExec_t graph_exec_{};
};

// One plane-wave transfer on the device through the operator's angular kernels
// (AngularTransferOperator::outgoingToParentDevice / incomingToChildDevice /
// scaledOutgoingToParentDevice / incomingToScaledChildDevice,
// modules/bem/cuda/wideband_operator.hpp:39-51). Orders: rule orders of the source
// and target signatures, or the spherical order (= bandlimit) of a hand-off's
// scaled expansion; radius: its basis radius.
std::vector<std::complex<double>> device_angular_transfer(dev::HfTransferKind kind, unsigned source_order,
                                                          unsigned target_order, unsigned bandlimit, double k,
                                                          double radius, Vec3 child_minus_parent,
                                                          const std::vector<std::complex<double>> &input,
                                                          hipStream_t stream);

} // namespace hsbem
