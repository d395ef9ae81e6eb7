// Device-resident Burton-Miller solver state behind the C ABI.
//
// One MeshSolver per closed mesh (reference: CudaBemSolver::Implementation,
// modules/bem/cuda/cuda_solver.cu:742-987): recentred mesh, Galerkin space,
// device geometry in a spatial (Morton) cluster order, interaction plans cached
// per (numerics, planning k), and a solve workspace reused across calls.
#pragma once

#include "host/plan.hpp"
#include "host/space.hpp"
#include "hip/kernels.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <complex>
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
    void resize(std::size_t count) {
        if (count == count_)
            return;
        release();
        if (count) {
            void *p = nullptr;
            hip_check(hipMalloc(&p, count * sizeof(T)), "device allocation");
            data_ = static_cast<T *>(p);
        }
        count_ = count;
    }
    void release() {
        if (data_)
            (void)hipFree(data_);
        data_ = nullptr;
        count_ = 0;
    }
    void upload(const std::vector<T> &v, hipStream_t s) {
        resize(v.size());
        if (count_)
            hip_check(hipMemcpyAsync(data_, v.data(), bytes(), hipMemcpyHostToDevice, s), "upload");
    }
    T *data() const { return data_; }
    std::size_t size() const { return count_; }
    std::size_t bytes() const { return count_ * sizeof(T); }

  private:
    T *data_{};
    std::size_t count_{};
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
    std::vector<ExcitationResult> excitations;
    double ms[6]{}; // assembly, solve, factorization, triangular, field, plan
    // device-event kernel times of this frequency (ms) and algorithmic counts
    double regular_kernel_ms{}, touching_kernel_ms{}, near_kernel_ms{};
    double regular_samples{}, touching_samples{}, near_samples{};
};

struct Result {
    std::vector<FrequencyResult> frequencies;
    std::uint32_t nt{}, nv{}, nobs{}, nexc{};
    std::uint64_t peak_bytes{};
};

struct GmresWorkspace {
    DeviceBuffer<double> basis, r, scratch, small;
};

// Restarted GMRES(restart) with CGS2 Arnoldi on a dense row-major-free complex
// system (A n x n, column-major); converged when ||b - A x|| <= tol ||b||.
void dense_gmres(hipStream_t stream, const void *A, bool fp32, std::uint32_t n, const double *b, double *x,
                 std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved, std::uint32_t *iterations,
                 const volatile std::int32_t *cancel, GmresWorkspace &ws);

using ProgressFn = void (*)(void *, std::uint32_t, std::uint32_t, double, std::int32_t);

struct DevicePlan {
    InteractionPlan plan;
    DeviceBuffer<TouchingJob> touching;
    DeviceBuffer<NearJob> near;
    DeviceBuffer<PairPatch> patches;
    DeviceBuffer<std::uint32_t> near_offsets, near_trials;
    DeviceBuffer<std::uint8_t> whole;
    DeviceBuffer<double> gauss;
    double plan_ms{};
    double touching_samples{}, near_samples{}; // algorithmic sample counts of one assembly
};

class MeshSolver {
  public:
    MeshSolver(Context *ctx, SurfaceMesh mesh);
    Result solve(const Request &request, ProgressFn progress, void *user, const volatile std::int32_t *cancel);
    // Component seams.
    InteractionPlan plan(const Numerics &numerics, double k_max);
    void assemble_dense(const Numerics &numerics, double k_max, double k, double *A, double *B);
    std::uint64_t hash() const { return hash_; }
    const Space &space() const { return space_; }
    std::uint32_t nv() const { return static_cast<std::uint32_t>(space_.vertices.size()); }
    std::uint32_t nt() const { return static_cast<std::uint32_t>(space_.panels.size()); }
    std::mutex &mutex() { return mutex_; }

  private:
    DevicePlan &device_plan(const Numerics &n, double k_max, bool *built);
    dev::AssemblyLaunch assembly_launch(const DevicePlan &p, const Numerics &n, double k, double *A, double *BD,
                                        std::uint32_t nd, const std::int32_t *flux_column);
    void build_clusters();

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
    DeviceBuffer<std::uint32_t> nodes_;
    DeviceBuffer<std::uint8_t> shape_, local_;
    DeviceBuffer<std::uint32_t> cluster_panel_offsets_, cluster_panels_, cluster_node_offsets_, cluster_nodes_;
    std::uint32_t cluster_count_{};
    DeviceBuffer<int> error_;
    DeviceBuffer<unsigned long long> samples_;
    std::map<std::tuple<unsigned, unsigned, double, double, unsigned, double>, std::unique_ptr<DevicePlan>> plans_;
    // workspace (reused across calls)
    DeviceBuffer<double> A_, LU_, BD_, q_, rhs_, x_, r_, flux_, scratch_, obs_, field_, partials_;
    GmresWorkspace gmres_;
    DeviceBuffer<std::int32_t> flux_column_, pivots_;
    std::mutex mutex_;
};

} // namespace hsbem
