// Device-resident Burton-Miller solver state behind the C ABI.
//
// One MeshSolver per closed mesh (reference: CudaBemSolver::Implementation,
// modules/bem/cuda/cuda_solver.cu:742-987): recentred mesh, Galerkin space,
// device geometry in a spatial (Morton) cluster order, interaction plans cached
// per (numerics, planning k), and a solve workspace reused across calls.
#pragma once

#include "host/plan.hpp"
#include "host/space.hpp"
#include "host/support_tree.hpp"
#include "host/translation_plan_cache.hpp"
#include "hip/kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <complex>
#include <functional>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace hsbem {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Cancelled : std::runtime_error {
    Cancelled() : std::runtime_error("BEM calculation cancelled.") {}
};

void hip_check(hipError_t status, const char *what);

template <class T> class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t count) { resize(count); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    // Sets the element count; storage only grows (no device free on shrinking).
    void resize(std::size_t count) {
        if (count > capacity_) {
            release();
            void *p = nullptr;
            hip_check(hipMalloc(&p, count * sizeof(T)), "device allocation");
            data_ = static_cast<T *>(p);
            capacity_ = count;
        }
        count_ = count;
    }
    void release() {
        if (data_)
            (void)hipFree(data_);
        data_ = nullptr;
        count_ = capacity_ = 0;
    }
    void upload(const std::vector<T> &v, hipStream_t s) {
        resize(v.size());
        if (count_)
            hip_check(hipMemcpyAsync(data_, v.data(), bytes(), hipMemcpyHostToDevice, s), "upload");
    }
    T *data() const { return data_; }
    std::size_t size() const { return count_; }
    std::size_t bytes() const { return count_ * sizeof(T); }
    std::size_t capacity_bytes() const { return capacity_ * sizeof(T); }

  private:
    T *data_{};
    std::size_t count_{}, capacity_{};
};

struct Context {
    int device{};
    hipStream_t stream{};
    std::string description;
};

struct Numerics {
    unsigned quadrature_order{4}, singular_order{4};
    double near_ratio{1.0}, near_tolerance{1e-4};
    unsigned near_depth{6};
    double residual_tolerance{5e-5};
    unsigned refinement_steps{3};
    unsigned solver{0}; // 0 LU, 1 GMRES
    unsigned gmres_restart{50}, gmres_max_iterations{1000};
    unsigned precision{0};
    unsigned plan_threads{0};
    unsigned preconditioner{1}; // GMRES: 0 none, 1 block-Jacobi (64 x 64 diagonal blocks)
    unsigned acceleration{0};   // 0 dense, 1 fast multipole
    unsigned fmm_leaf_capacity{16};
    double fmm_separation{1.5};
    unsigned fmm_maximum_iterations{300}, fmm_restart{40};
    unsigned fmm_block_width{32}; // preconditionerBlockWidth: power of two 32..2048
};

struct Excitation {
    std::vector<std::pair<std::uint32_t, std::complex<double>>> faces;
};

struct Request {
    double density{1.21}, sound_speed{343};
    std::vector<double> frequencies, planning_frequencies;
    std::vector<Excitation> excitations;
    std::vector<Vec3> observers;
    bool far_field{};
    Vec3 far_origin{};
    double far_reference_distance{1};
    Numerics numerics;
};

struct Impedance {
    std::uint32_t face_id{};
    double area{};
    std::complex<double> mean_pressure, volume_velocity, impedance;
};

struct ExcitationResult {
    std::vector<std::complex<double>> surface_pressure, pressure_trace, normal_derivative, observer_pressure;
    std::vector<Impedance> impedances;
    double residual{};
    std::uint32_t iterations{}, refinements{};
};

struct FrequencyResult {
    double frequency{};
    double served_ms{}; // serving loop: submission to completion of this frequency point
    std::vector<ExcitationResult> excitations;
    double ms[6]{}; // assembly, solve, factorization, triangular, field, plan
    // device-event kernel times of this frequency (ms) and algorithmic counts
    double regular_kernel_ms{}, touching_kernel_ms{}, near_kernel_ms{};
    double regular_samples{}, touching_samples{}, near_samples{};
    // FMM form: seed-audit discrepancy ||(A_seed - A_seed/10) x|| / ||b|| (largest excitation),
    // whether the defect correction with the finer operator ran, and the operator's apply time
    double fmm_audit{}, fmm_apply_ms{};
    bool fmm_corrected{}, fmm_direct_audit{};
    bool fmm_fp32{}; // the Krylov iterations applied the operator with FP32 translations
    double fmm_quadrature_discrepancy{}, fmm_verification_seed{};
    unsigned fmm_endpoint_degree{}, fmm_promotions{};
    std::uint64_t fmm_plane_wave_levels{}, fmm_m2l_links{}, fmm_near_nnz{}, fmm_device_bytes{};
};

struct Result {
    std::vector<FrequencyResult> frequencies;
    std::uint32_t nt{}, nv{}, nobs{}, nexc{};
    std::uint64_t peak_bytes{};
};

struct GmresWorkspace {
    GmresWorkspace() = default;
    GmresWorkspace(const GmresWorkspace &) = delete;
    GmresWorkspace &operator=(const GmresWorkspace &) = delete;
    ~GmresWorkspace();
    void release();
    DeviceBuffer<double> basis, r, z, scratch, small;
    // pinned per-step copies of the Arnoldi column and its events (pipelined host Givens)
    double *pinned{};
    std::size_t pinned_doubles{};
    std::vector<hipEvent_t> events;
    // several right-hand sides in lockstep: one basis per system, packed operator
    // inputs / outputs, per-system Arnoldi columns
    DeviceBuffer<double> multi_basis, multi_in, multi_out, multi_small;
    double *multi_pinned{};
    std::size_t multi_pinned_doubles{};
    // device-resident cycles (dense_gmres): v_j and the operator output, the cycle
    // state, the captured Arnoldi step and what it was captured for, pinned state reads
    DeviceBuffer<double> cycle_current, cycle_w, cycle_z, cycle_small;
    DeviceBuffer<int> cycle_state;
    hipGraphExec_t step_graph{};
    std::tuple<const void *, bool, std::uint32_t, std::uint32_t, const double *, std::uint32_t, const double *,
               const double *>
        step_key{};
    int *cycle_pinned{};
    std::size_t cycle_pinned_ints{};
    std::vector<hipEvent_t> cycle_events;
};

// Right-hand sides solved together by dense_gmres_multi (one multi-vector
// product per Arnoldi step reads the matrix once for all of them).
inline constexpr std::uint32_t kGmresBatch = 8;

// Device buffers and stream of one solve in flight. The direct solve path owns
// one per mesh (on the context stream); the serving loop owns one per slot,
// each on its own stream, so solves of different requests overlap.
struct Workspace {
    explicit Workspace(hipStream_t stream, bool owns_stream = false);
    ~Workspace();
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    void release();                       // frees every buffer (device-synchronising)
    std::uint64_t capacity_bytes() const; // device memory held
    hipStream_t stream{};
    bool owns_stream{};
    DeviceBuffer<double> A, LU, BD, q, rhs, x, r, flux, scratch, obs, field, partials;
    DeviceBuffer<double> fmm_b, fmm_x, fmm_r, fmm_q; // FMM form, mesh vertex order
    DeviceBuffer<std::int32_t> flux_column, pivots;
    DeviceBuffer<double> jacobi, jacobi_blocks; // inverses; gathered blocks of a wide block-Jacobi
    DeviceBuffer<std::int32_t> jacobi_pivots;
    DeviceBuffer<int> error;
    DeviceBuffer<unsigned long long> samples;
    GmresWorkspace gmres;
    hipEvent_t events[4]{};
    std::uint64_t bound{}; // id of the PreparedRequest whose per-request data is uploaded
};

// Block-Jacobi inverses of a dense GMRES: 64 x 64 blocks (jacobi_factor layout) or
// wider ones (jacobi_factor_wide layout); no preconditioner when inverses is null.
struct BlockJacobi {
    const double *inverses = nullptr;
    std::uint32_t width = 64;
    explicit operator bool() const { return inverses != nullptr; }
};
void apply_block_jacobi(const BlockJacobi &jacobi, std::uint32_t n, const double *v, double *out, hipStream_t stream);
// Width of the dense forms' block-Jacobi: HSBEM_DENSE_JACOBI_WIDTH (a power of two in
// 64 ... 2048), else 64.
std::uint32_t dense_jacobi_width();

// Restarted GMRES(restart) with CGS2 Arnoldi on a dense complex system (A n x n,
// column-major, FP64 or FP32), right-preconditioned by the block-Jacobi inverses
// when given; converged when ||b - A x|| <= tol ||b||.
void dense_gmres(hipStream_t stream, const void *A, bool fp32, std::uint32_t n, const double *b, double *x,
                 std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved, std::uint32_t *iterations,
                 const volatile std::int32_t *cancel, GmresWorkspace &ws, const BlockJacobi &jacobi);

// dense_gmres for nrhs systems A x_e = b_e (b, x: nrhs vectors of n complex,
// consecutive) in lockstep: every Arnoldi step applies A to the active systems'
// preconditioned vectors in one multi-vector product; each system keeps its own
// basis, Hessenberg and rotations, leaves the cycle when its residual estimate
// reaches tol ||b_e||, and is finished when its true residual does.
void dense_gmres_multi(hipStream_t stream, const void *A, bool fp32, std::uint32_t n, std::uint32_t nrhs,
                       const double *b, double *x, std::uint32_t restart, std::uint32_t max_it, double tol,
                       double *achieved, std::uint32_t *iterations, const volatile std::int32_t *cancel,
                       GmresWorkspace &ws, const BlockJacobi &jacobi);

// A matrix-free operator for gmres(): y = A x and z = M^{-1} v (optional), both
// enqueued on the given stream; residual (optional) r = b - A x in one pass.
struct KrylovOperator {
    std::function<void(const double *x, double *y, hipStream_t)> apply;
    std::function<void(const double *v, double *z, hipStream_t)> precondition;
    std::function<void(const double *b, const double *x, double *r, hipStream_t)> residual;
};

// Restarted, right-preconditioned GMRES(restart) with CGS2 Arnoldi over any
// device operator; converged when ||b - A x|| <= tol ||b||.
void gmres(hipStream_t stream, const KrylovOperator &op, std::uint32_t n, const double *b, double *x,
           std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved, std::uint32_t *iterations,
           const volatile std::int32_t *cancel, GmresWorkspace &ws);

using ProgressFn = void (*)(void *, std::uint32_t, std::uint32_t, double, std::int32_t);

struct DevicePlan {
    InteractionPlan plan;
    DeviceBuffer<TouchingJob> touching;
    DeviceBuffer<NearJob> near;                // jobs with a projection-fan patch
    DeviceBuffer<dev::NearPatchItem> near_items; // tensor patches of all other jobs
    std::uint32_t near_job_count{}, near_item_count{};
    DeviceBuffer<PairPatch> patches;
    DeviceBuffer<std::uint32_t> near_offsets, near_trials;
    DeviceBuffer<std::uint8_t> whole;
    DeviceBuffer<double> gauss;
    double plan_ms{};
    double touching_samples{}, near_samples{}; // algorithmic sample counts of one assembly
    // Tile list: far tiles (no touching or near pair, degree 4 whenever
    // k * tile_edge <= 0.5) sorted by tile_edge, then all other cluster pairs.
    DeviceBuffer<std::uint32_t> tiles;
    std::vector<double> far_tile_edge; // ascending, one per far tile
    std::uint32_t tile_count{};
    std::uint64_t last_use{}; // plan cache recency (MeshSolver::plans_)
    std::uint64_t device_bytes() const {
        return touching.capacity_bytes() + near.capacity_bytes() + near_items.capacity_bytes() +
               patches.capacity_bytes() + near_offsets.capacity_bytes() + near_trials.capacity_bytes() +
               whole.capacity_bytes() + gauss.capacity_bytes() + tiles.capacity_bytes();
    }
};

// Plans kept per mesh (least recently used evicted beyond this, unless a solve holds it).
inline constexpr std::size_t kPlanCacheEntries = 4;

// Bounds of one spatial cluster of panels (for conservative tile tests).
struct ClusterBounds {
    Vec3 center{};
    double spread{};     // max distance of a panel centroid from center
    double ball{};       // max panel ball radius
    double max_edge{};
    bool shapes_ok{};
};

// A validated request with its host-side derived data (observers, flux columns,
// per-face velocities), shared by all of its frequency points.
struct PreparedRequest {
    std::uint64_t id{}; // unique per preparation (workspace binding key)
    Request request;
    std::vector<Vec3> observers; // recentred points, or unit far-field directions
    Vec3 far_origin{};
    std::vector<std::map<std::uint32_t, std::complex<double>>> face_velocity;
    std::vector<std::int32_t> flux_column;
    std::uint32_t nd{}, ne{}, nv{}, nt{};
    double k_max{};
    bool has_nonzero{};
    std::uint64_t footprint{};     // device bytes a workspace holds for one frequency point
    std::uint64_t plan_estimate{}; // device bytes of this request's interaction plan when first built
};

// Exact near part of the FMM operator split (pressure CSR over P1 x P1 and
// flux CSR over P1 x active faces), in mesh vertex order, rows scaled by
// 1/lumped mass; `applied` = A_near x when x is given.
struct DeviceNear;
struct NearStructure;
struct FmmSettings;
class FmmOperator;
struct SupportNearPlan;

struct NearOperator {
    std::vector<std::uint32_t> row_offsets, columns, flux_offsets, flux_columns;
    std::vector<double> values, flux_values, applied; // complex
    std::size_t jobs{}, regular{}, touching{}, patches{}, fan_jobs{};
    double kernel_ms{};
};

class MeshSolver {
  public:
    MeshSolver(Context *ctx, SurfaceMesh mesh);
    Result solve(const Request &request, ProgressFn progress, void *user, const volatile std::int32_t *cancel);
    // Serving seams: validation and host preparation (no device work), then one
    // frequency point on a caller-owned workspace. Thread safe across workspaces.
    std::shared_ptr<PreparedRequest> prepare(const Request &request) const;
    FrequencyResult solve_frequency(const PreparedRequest &prepared, std::size_t index, Workspace &ws,
                                    const volatile std::int32_t *cancel, bool *plan_built = nullptr,
                                    const std::function<void(std::int32_t)> &report = {});
    // Component seams.
    // The Galerkin interaction plan; device: its near-pair patch estimates come from the
    // device (estimate_patches, on `stream` or the mesh's stream), else from host threads.
    InteractionPlan plan(const Numerics &numerics, double k_max, hipStream_t stream = nullptr, bool device = false);
    // Whether the fast multipole form estimates its near pairs on the device (default;
    // HSBEM_PLAN_ESTIMATOR=host at mesh construction takes host threads).
    bool device_estimates() const { return device_estimates_; }
    NearOperator near_operator(const Numerics &numerics, double k_max, double k, const SupportTreeSettings &tree,
                               const std::vector<std::uint32_t> &active_faces, const double *x);
    void assemble_dense(const Numerics &numerics, double k_max, double k, double *A, double *B);
    // Matrix-free FMM pressure operator at wavenumber k (near entries planned for k_max).
    std::unique_ptr<FmmOperator> fmm_operator(const Numerics &numerics, double k_max, double k,
                                              const FmmSettings &settings, const std::vector<std::uint32_t> &active_faces);
    std::uint64_t hash() const { return hash_; }
    // Device bytes this mesh keeps resident between solves: geometry, clusters,
    // cached interaction plans and the cached FMM near structure.
    std::uint64_t resident_bytes();
    // Device bytes the plan of a request will add when it is not cached yet (0 when cached).
    std::uint64_t pending_plan_bytes(const PreparedRequest &prepared);
    std::size_t cached_plans();
    // Drops cached interaction plans no solve holds, least recently used first, until at
    // most `keep` remain; returns the device bytes released.
    std::uint64_t trim_plans(std::size_t keep);
    // Largest device footprint measured for one FMM point of this mesh with these settings (0: none yet).
    std::uint64_t fmm_point_bytes(const Numerics &numerics);
    void note_fmm_point_bytes(const Numerics &numerics, std::uint64_t bytes);
    const Space &space() const { return space_; }
    std::uint32_t nv() const { return static_cast<std::uint32_t>(space_.vertices.size()); }
    std::uint32_t nt() const { return static_cast<std::uint32_t>(space_.panels.size()); }
    std::mutex &mutex() { return mutex_; }
    Context *context() const { return ctx_; }

  private:
    // device: near-pair patch estimates on the device (the fast multipole form's near
    // operator, as the reference builds it), else on host threads (the dense form's plan,
    // the reference's CPU plan)
    std::shared_ptr<DevicePlan> device_plan(const Numerics &n, double k_max, hipStream_t stream, bool *built,
                                            bool device = false);
    dev::AssemblyLaunch assembly_launch(const DevicePlan &p, const Numerics &n, double k, double *A, double *BD,
                                        std::uint32_t nd, const std::int32_t *flux_column, int *error);
    void tune_regular_tiles(dev::AssemblyLaunch &launch, bool mixed, hipStream_t stream) const;
    void bind(const PreparedRequest &prepared, Workspace &ws);
    void build_clusters();
    // Near structure (tree, near plan, split jobs, CSR pattern; cached per settings)
    // and its entries at wavenumber k (original numbering, rows scaled by 1/mass).
    std::shared_ptr<const NearStructure> near_structure(const Numerics &n, double k_max, const SupportTreeSettings &ts,
                                                        const std::vector<std::uint32_t> &faces, hipStream_t stream);
    void integrate_near(const Numerics &n, double k, const std::shared_ptr<const NearStructure> &structure,
                        hipStream_t stream, DeviceNear &out);
    // FMM operator on a caller stream; share: reuse its near entries and preconditioner.
    std::unique_ptr<FmmOperator> build_fmm(const Numerics &n, double k_max, double k, const FmmSettings &settings,
                                           const std::vector<std::uint32_t> &faces, hipStream_t stream,
                                           const FmmOperator *share, unsigned endpoint_degree = 0);
    // One frequency point of the FMM form: right-hand sides, GMRES, residuals and the
    // seed audit; leaves the pressures in ws.x (internal order).
    void fmm_point(const PreparedRequest &P, Workspace &ws, double k, double omega, std::vector<double> &relative,
                   std::vector<std::uint32_t> &iterations, FrequencyResult &fr, const volatile std::int32_t *cancel);

    Context *ctx_;
    SurfaceMesh mesh_;
    Vec3 origin_;
    Space space_;
    std::uint64_t hash_{};
    std::vector<std::uint32_t> internal_of_; // original vertex -> internal node
    std::vector<std::uint32_t> vertex_of_;   // internal node -> original vertex
    std::vector<double> lumped_internal_;
    // device geometry
    DeviceBuffer<double> geo_, bounds_, row_scale_, lumped_, rules_;
    // observer / far-field panel rules of the endpoint degrees (4, 6, 8, 12, 14), resident
    // so that a served point allocates (and frees, which synchronises the device) nothing
    DeviceBuffer<double> endpoint_rules_, unit_;
    std::map<unsigned, std::pair<std::uint32_t, std::uint32_t>> endpoint_rule_slots_; // degree -> (first point, count)
    DeviceBuffer<std::uint32_t> nodes_;
    DeviceBuffer<std::uint8_t> shape_, local_;
    DeviceBuffer<std::uint32_t> cluster_panel_offsets_, cluster_panels_, cluster_node_offsets_, cluster_nodes_;
    DeviceBuffer<double> cluster_far_; // per cluster: the bounds of the far tiles' fixed-point scale
    std::uint32_t cluster_count_{};
    std::vector<ClusterBounds> cluster_bounds_;
    std::map<std::tuple<unsigned, unsigned, double, double, unsigned, double, bool>, std::shared_ptr<DevicePlan>> plans_;
    std::mutex plan_mutex_;       // serialises plan and near-structure builds
    bool device_estimates_{true}; // the FMM near operator's patch estimates on the device
    std::mutex estimator_mutex_;
    DeviceBuffer<dev::EstimatorPanel> estimator_panels_; // built with the first device-estimated plan
    DeviceBuffer<TrianglePoint> estimator_rules_;
    std::mutex plan_index_mutex_; // guards plans_ lookups, insertion and eviction only
    std::uint64_t plan_clock_{};
    std::atomic<std::uint64_t> plan_bytes_{0}, near_bytes_{0};
    std::uint64_t geometry_bytes_{};
    std::map<std::tuple<unsigned, double, unsigned, unsigned, unsigned>, std::uint64_t> fmm_point_bytes_;
    std::unique_ptr<Workspace> workspace_; // direct solve path (context stream), reused across calls
    DeviceBuffer<std::uint32_t> original_nodes_; // 3 x nt, mesh vertex ids (near operator)
    DeviceBuffer<double> original_row_scale_, original_lumped_;
    DeviceBuffer<std::uint32_t> internal_of_device_; // mesh vertex -> internal node
    using NearKey = std::tuple<unsigned, unsigned, double, double, std::vector<std::uint32_t>, unsigned, unsigned,
                               double, double, unsigned, double>;
    std::shared_ptr<const NearStructure> near_cache_;
    TranslationPlanCache translation_plans_; // FMM rotation and coaxial plans of this mesh (thread safe)
    NearKey near_cache_key_;
    std::mutex mutex_;                     // one direct solve per mesh at a time (as the reference)
};

} // namespace hsbem
