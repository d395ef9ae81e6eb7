// C++ drop-in for hornsculpt::CudaBemSolver on a closed exterior mesh (reference:
// modules/bem/include/hornsculpt/bem.hpp:27-388, modules/core/include/hornsculpt/mesh.hpp:10-33)
// over the C ABI in hsbem.h: dense and fast multipole accelerations, the reference's
// request / result / progress types with their names and meanings. Header-only;
// link libhsbem.so. Branches outside this library's scope (subdomains, symmetry
// planes, exterior and infinite baffles, axisymmetric, LEM coupling) are not
// declared, so a caller that uses them does not compile against this header.
//
// Result fields the library does not measure keep the reference's "not recorded"
// value (zero, empty or nullopt): graph-execution evidence, historical Chebyshev
// evidence, per-level comparison/verification orders, field-quadrature checks and the
// pressure-flux pair counts of FastMultipoleStatistics.
#pragma once

#include "../hsbem.h"

#include <algorithm>
#include <array>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <stop_token>
#include <string>
#include <vector>

namespace hornsculpt {

// ---- mesh (mesh.hpp:10-33) ----
struct Vec3 {
    double x{};
    double y{};
    double z{};
    bool operator==(const Vec3 &) const = default;
};
enum class BoundaryCondition : std::uint8_t { wall, driven };
struct Triangle {
    std::array<std::uint32_t, 3> vertices{};
    std::uint32_t faceId{};
    BoundaryCondition boundary{BoundaryCondition::wall};
    bool operator==(const Triangle &) const = default;
};
struct SurfaceMesh {
    std::vector<Vec3> vertices;
    std::vector<Triangle> triangles;
    // Remeshing hint of the reference's mesher; the solver does not read it.
    std::vector<std::uint32_t> fixedVertices;
    bool operator==(const SurfaceMesh &) const = default;
};

// ---- request (bem.hpp:27-191) ----
struct AcousticMedium {
    double density{1.21};
    double soundSpeed{343.0};
};
struct FaceVelocity {
    std::uint32_t faceId{};
    std::complex<double> normalVelocity{1.0, 0.0}; // RMS m/s
};
struct BemExcitation {
    std::string name;
    std::vector<FaceVelocity> faces;
};
enum class BemAcceleration { dense, fastMultipole };
enum class BemResidualEvaluation { directGalerkin, higherOrderFastMultipole };
enum class BemFastMultipoleRepresentation { regularSpherical, planeWave };
struct BemFastMultipoleLevelEvidence {
    unsigned level{};
    BemFastMultipoleRepresentation representation{BemFastMultipoleRepresentation::regularSpherical};
    unsigned momentBandlimit{}, translationOrder{}, quadratureOrder{};
    bool operator==(const BemFastMultipoleLevelEvidence &) const = default;
};
struct BemHistoricalChebyshevEvidence {
    unsigned comparisonOrder{}, auditOrder{};
    bool operator==(const BemHistoricalChebyshevEvidence &) const = default;
};
struct FastMultipoleSettings {
    unsigned leafSupportCapacity{16};
    double separationRatio{1.5};
    unsigned maximumIterations{300};
    unsigned restartLength{40};
    unsigned preconditionerBlockWidth{32}; // powers of two 32..2048
    bool operator==(const FastMultipoleSettings &) const = default;
};
struct BemHistoricalChebyshevStatistics {
    std::size_t directPanelPairCount{}, representedFarPanelPairs{};
    bool operator==(const BemHistoricalChebyshevStatistics &) const = default;
};
struct BemFastMultipoleGraphEvidence {
    std::uint64_t captureAttempts{}, instantiations{}, launches{}, reuses{}, executedBatches{}, canonicalChunks{},
        canonicalBatches{}, invalidatedExecutables{}, budgetBypasses{}, admissionClosures{},
        maximumCachedExecutables{}, maximumCachedNodes{}, peakOwnedHostMetadataBytes{};
    double captureMilliseconds{}, instantiationMilliseconds{};
    bool memoryObservationAvailable{};
    std::uint64_t minimumObservedFreeDeviceBytes{}, totalDeviceBytes{}, requiredFreeReserveBytes{};
    bool operator==(const BemFastMultipoleGraphEvidence &) const = default;
};
struct FastMultipoleStatistics {
    std::size_t nodeCount{}, leafCount{}, farInteractionCount{};
    std::uint64_t pressurePressureNearPairs{}, pressurePressureFarPairs{};
    std::uint64_t pressureFluxNearPairs{}, pressureFluxFarPairs{};
    std::optional<BemHistoricalChebyshevStatistics> historicalChebyshevStatistics;
    std::optional<BemFastMultipoleGraphEvidence> graphExecution;
};
struct BemNumerics {
    unsigned quadratureOrder{4};
    unsigned singularOrder{4};
    double nearDistanceRatio{1.0};
    double nearRelativeTolerance{1e-4};
    unsigned nearMaximumDepth{6};
    double relativeResidualTolerance{5e-5};
    unsigned maximumRefinementSteps{3};
    BemAcceleration acceleration{BemAcceleration::dense};
    FastMultipoleSettings fastMultipole;
};
struct FarFieldObservation {
    Vec3 origin{};
    double referenceDistance{1.0};
    bool operator==(const FarFieldObservation &) const = default;
};
struct BemRequest {
    AcousticMedium medium;
    std::vector<double> frequencies;
    std::vector<BemExcitation> excitations;
    std::vector<Vec3> observers;
    // Empty: every observer is valid. A closed exterior mesh has no masked half-space,
    // so a nonempty mask must mark every observer valid.
    std::vector<std::uint8_t> observerValidity;
    BemNumerics numerics;
    std::optional<FarFieldObservation> farField;
    std::vector<double> interactionPlanFrequencies;
};

// ---- results (bem.hpp:193-355) ----
struct FaceImpedance {
    std::uint32_t faceId{};
    double area{};
    std::complex<double> meanPressure;
    std::complex<double> volumeVelocity;
    std::complex<double> volumeImpedance;
};
struct BemBoundaryData {
    std::vector<std::complex<double>> surfacePressure;
    std::vector<std::complex<double>> pressureTrace;
    std::vector<std::complex<double>> normalDerivative;
    std::vector<std::complex<double>> exteriorSurfacePressure; // empty: no exterior driven mesh here
};
struct BemExcitationResult {
    std::string name;
    std::vector<std::complex<double>> observerPressure;
    std::vector<std::uint8_t> observerValidity;
    std::vector<FaceImpedance> impedances;
    unsigned refinementSteps{};
    double boundaryResidual{};
    BemResidualEvaluation residualEvaluation{BemResidualEvaluation::directGalerkin};
    double fastMultipoleKrylovResidual{};
    double fastMultipoleKrylovTarget{};
    double fastMultipoleOperatorDiscrepancy{};
    double fastMultipoleAccuracySeed{};
    double fastMultipoleComparisonAccuracySeed{};
    double fastMultipoleVerificationAccuracySeed{};
    unsigned fastMultipoleAccuracyRefinements{};
    std::vector<BemFastMultipoleLevelEvidence> fastMultipoleComparisonLevels;
    std::vector<BemFastMultipoleLevelEvidence> fastMultipoleVerificationLevels;
    std::optional<BemHistoricalChebyshevEvidence> historicalChebyshevEvidence;
    unsigned farQuadratureComparisonDegree{}, farQuadratureVerificationDegree{};
    unsigned farQuadratureRefinements{};
    double farQuadratureDiscrepancy{};
    double farQuadraturePreparationMilliseconds{}, farQuadraturePressureMilliseconds{},
        farQuadratureFluxMilliseconds{};
    unsigned fieldQuadratureComparisonDegree{}, fieldQuadratureVerificationDegree{};
    double fieldQuadratureDiscrepancy{};
    std::vector<FaceVelocity> solvedVelocities;
    unsigned iterations{}; // Krylov steps; zero for the dense LU
    void setBoundary(BemBoundaryData value) { boundary_ = std::make_shared<const BemBoundaryData>(std::move(value)); }
    const std::shared_ptr<const BemBoundaryData> &boundary() const noexcept { return boundary_; }

  private:
    std::shared_ptr<const BemBoundaryData> boundary_;
};
struct BemFrequencyResult {
    double frequency{};
    std::vector<BemExcitationResult> excitations;
    double assemblyMilliseconds{};
    double solveMilliseconds{};
    double factorizationMilliseconds{};
    double triangularSolveMilliseconds{};
    double fieldMilliseconds{};
    double residualCheckMilliseconds{};
    FastMultipoleStatistics fastMultipole;
    double interactionPlanMilliseconds{};
    bool interactionPlanReused{};
};
enum class BemGeometryMode { threeDimensional, axisymmetric };
struct BemResult {
    std::string device;
    std::size_t triangleCount{};
    std::size_t pressureDofCount{};
    std::size_t peakDeviceBytes{};
    std::vector<BemFrequencyResult> frequencies;
    BemAcceleration acceleration{BemAcceleration::dense};
    BemGeometryMode geometryMode{BemGeometryMode::threeDimensional};
    std::size_t liveDeviceBytes{};
};
enum class BemSolvePhase { preparing, nearConstruction, farPreparation, krylovSolve, residualAudit, fieldEvaluation,
                           frequencyComplete };
struct BemProgress {
    std::size_t completedFrequencies{};
    std::size_t totalFrequencies{};
    double frequency{};
    BemSolvePhase phase{BemSolvePhase::preparing};
    // The library reports phases per frequency; Krylov steps and operator
    // applications are not streamed (zero), the right-hand sides are the request's.
    unsigned iterations{};
    std::size_t pressureOperatorApplications{}, faceFluxOperatorApplications{};
    std::size_t rightHandSideIndex{}, rightHandSideCount{};
};

// Storage choice of the reference's quadrature patches (bem.hpp:362-366). This
// library keeps them in HBM in every case; the option is accepted for source
// compatibility and changes nothing numerically, as in the reference.
enum class CudaCacheResidency { automatic, device, mappedHost };
struct CudaStorageOptions {
    CudaCacheResidency quadraturePatches{CudaCacheResidency::automatic};
};

namespace hsbem_detail {
inline void check(int status) {
    if (status == HSBEM_OK)
        return;
    const std::string message = hsbem_last_error();
    if (status == HSBEM_INVALID_ARGUMENT)
        throw std::invalid_argument(message);
    throw std::runtime_error(message);
}
inline std::complex<double> at(const std::vector<double> &v, std::size_t i) { return {v[2 * i], v[2 * i + 1]}; }

// Field layout of hsbem_result_fmm.
enum FmmEvidence : int {
    kAuditDiscrepancy = 0, kApplyMs = 1, kCorrected = 2, kPlaneWaveLevels = 3, kM2lLinks = 4, kNearEntries = 5,
    kDeviceBytes = 6, kDirectAudit = 7, kQuadratureDiscrepancy = 8, kEndpointDegree = 9, kPromotions = 10,
    kVerificationSeed = 11, kFp32Translations = 12, kFmmEvidenceCount = 13
};
} // namespace hsbem_detail

class CudaBemSolver {
  public:
    explicit CudaBemSolver(SurfaceMesh mesh) : CudaBemSolver(std::move(mesh), CudaStorageOptions{}) {}
    CudaBemSolver(SurfaceMesh mesh, CudaStorageOptions storage) {
        if (storage.quadraturePatches != CudaCacheResidency::automatic &&
            storage.quadraturePatches != CudaCacheResidency::device &&
            storage.quadraturePatches != CudaCacheResidency::mappedHost)
            throw std::invalid_argument("unknown quadrature patch residency");
        hsbem_detail::check(hsbem_context_create(0, &context_));
        std::vector<double> xyz;
        std::vector<std::uint32_t> tri, faces;
        std::vector<std::uint8_t> bc;
        xyz.reserve(3 * mesh.vertices.size());
        for (const auto &v : mesh.vertices)
            xyz.insert(xyz.end(), {v.x, v.y, v.z});
        for (const auto &t : mesh.triangles) {
            tri.insert(tri.end(), t.vertices.begin(), t.vertices.end());
            faces.push_back(t.faceId);
            bc.push_back(t.boundary == BoundaryCondition::driven ? HSBEM_BOUNDARY_DRIVEN : HSBEM_BOUNDARY_WALL);
        }
        const int status = hsbem_mesh_create(context_, xyz.data(), static_cast<std::uint32_t>(mesh.vertices.size()),
                                             tri.data(), faces.data(), bc.data(),
                                             static_cast<std::uint32_t>(mesh.triangles.size()), &mesh_);
        if (status != HSBEM_OK) {
            hsbem_context_destroy(context_);
            hsbem_detail::check(status);
        }
    }
    ~CudaBemSolver() {
        hsbem_mesh_destroy(mesh_);
        hsbem_context_destroy(context_);
    }
    CudaBemSolver(const CudaBemSolver &) = delete;
    CudaBemSolver &operator=(const CudaBemSolver &) = delete;

    static std::string deviceDescription() {
        char buffer[256];
        hsbem_detail::check(hsbem_device_description(buffer, sizeof buffer));
        return buffer;
    }

    BemResult solve(const BemRequest &request, std::stop_token stop = {},
                    std::function<void(const BemProgress &)> progress = {}) {
        if (!request.observerValidity.empty()) {
            if (request.observerValidity.size() != request.observers.size())
                throw std::invalid_argument("observer validity mask must match the observers");
            for (std::uint8_t v : request.observerValidity)
                if (v != 1)
                    throw std::invalid_argument("a closed exterior mesh has no masked observers");
        }
        std::vector<std::uint32_t> offsets{0}, faces;
        std::vector<double> velocities, observers;
        for (const auto &e : request.excitations) {
            for (const auto &f : e.faces) {
                faces.push_back(f.faceId);
                velocities.insert(velocities.end(), {f.normalVelocity.real(), f.normalVelocity.imag()});
            }
            offsets.push_back(static_cast<std::uint32_t>(faces.size()));
        }
        for (const auto &o : request.observers)
            observers.insert(observers.end(), {o.x, o.y, o.z});
        hsbem_request rq{};
        hsbem_numerics_default(&rq.numerics);
        const BemNumerics &n = request.numerics;
        rq.numerics.quadrature_order = n.quadratureOrder;
        rq.numerics.singular_order = n.singularOrder;
        rq.numerics.near_distance_ratio = n.nearDistanceRatio;
        rq.numerics.near_relative_tolerance = n.nearRelativeTolerance;
        rq.numerics.near_maximum_depth = n.nearMaximumDepth;
        rq.numerics.relative_residual_tolerance = n.relativeResidualTolerance;
        rq.numerics.maximum_refinement_steps = n.maximumRefinementSteps;
        const bool fmm = n.acceleration == BemAcceleration::fastMultipole;
        rq.numerics.acceleration = fmm ? HSBEM_ACCELERATION_FAST_MULTIPOLE : HSBEM_ACCELERATION_DENSE;
        if (fmm)
            rq.numerics.solver = HSBEM_SOLVER_GMRES;
        rq.numerics.fmm_leaf_capacity = n.fastMultipole.leafSupportCapacity;
        rq.numerics.fmm_separation_ratio = n.fastMultipole.separationRatio;
        rq.numerics.fmm_maximum_iterations = n.fastMultipole.maximumIterations;
        rq.numerics.fmm_restart_length = n.fastMultipole.restartLength;
        rq.numerics.fmm_block_width = n.fastMultipole.preconditionerBlockWidth;
        rq.density = request.medium.density;
        rq.sound_speed = request.medium.soundSpeed;
        rq.frequencies = request.frequencies.data();
        rq.frequency_count = static_cast<std::uint32_t>(request.frequencies.size());
        rq.excitation_count = static_cast<std::uint32_t>(request.excitations.size());
        rq.excitation_offsets = offsets.data();
        rq.face_ids = faces.data();
        rq.velocities = velocities.data();
        rq.observers = observers.data();
        rq.observer_count = static_cast<std::uint32_t>(request.observers.size());
        rq.far_field = request.farField ? 1 : 0;
        if (request.farField) {
            rq.far_origin[0] = request.farField->origin.x;
            rq.far_origin[1] = request.farField->origin.y;
            rq.far_origin[2] = request.farField->origin.z;
            rq.far_reference_distance = request.farField->referenceDistance;
        }
        rq.planning_frequencies = request.interactionPlanFrequencies.data();
        rq.planning_frequency_count = static_cast<std::uint32_t>(request.interactionPlanFrequencies.size());
        // The stop token is polled by the library through this flag.
        volatile std::int32_t cancel = 0;
        std::stop_callback on_stop(stop, [&cancel] { cancel = 1; });
        struct Bridge {
            std::function<void(const BemProgress &)> *fn;
            std::size_t rhs;
        } bridge{&progress, request.excitations.size()};
        auto trampoline = [](void *user, std::uint32_t done, std::uint32_t total, double f, std::int32_t phase) {
            auto *b = static_cast<Bridge *>(user);
            if (*b->fn) {
                BemProgress p{done, total, f, static_cast<BemSolvePhase>(phase)};
                p.rightHandSideCount = b->rhs;
                (*b->fn)(p);
            }
        };
        hsbem_result *raw = nullptr;
        hsbem_detail::check(hsbem_solve(mesh_, &rq, trampoline, &bridge, &cancel, &raw));
        std::unique_ptr<hsbem_result, int (*)(hsbem_result *)> res(raw, hsbem_result_destroy);
        std::uint32_t nf{}, ne{}, nt{}, nv{}, no{};
        hsbem_detail::check(hsbem_result_sizes(raw, &nf, &ne, &nt, &nv, &no));
        BemResult out;
        out.device = deviceDescription();
        out.triangleCount = nt;
        out.pressureDofCount = nv;
        out.acceleration = n.acceleration;
        std::uint64_t peak{};
        hsbem_detail::check(hsbem_result_peak_device_bytes(raw, &peak));
        out.peakDeviceBytes = peak;
        for (std::uint32_t f = 0; f < nf; ++f) {
            BemFrequencyResult fr;
            fr.frequency = request.frequencies[f];
            double ms[12]{};
            hsbem_detail::check(hsbem_result_timings(raw, f, ms));
            fr.assemblyMilliseconds = ms[0];
            fr.solveMilliseconds = ms[1];
            fr.factorizationMilliseconds = ms[2];
            fr.triangularSolveMilliseconds = ms[3];
            fr.fieldMilliseconds = ms[4];
            fr.interactionPlanMilliseconds = ms[5];
            fr.interactionPlanReused = f == 0 && ms[5] == 0.0;
            double ev[hsbem_detail::kFmmEvidenceCount]{};
            if (fmm) {
                hsbem_detail::check(hsbem_result_fmm(raw, f, ev));
                out.liveDeviceBytes = std::max<std::size_t>(out.liveDeviceBytes,
                                                            static_cast<std::size_t>(ev[hsbem_detail::kDeviceBytes]));
                const auto near = static_cast<std::uint64_t>(ev[hsbem_detail::kNearEntries]);
                fr.fastMultipole.farInteractionCount = static_cast<std::size_t>(ev[hsbem_detail::kM2lLinks]);
                fr.fastMultipole.pressurePressureNearPairs = near;
                fr.fastMultipole.pressurePressureFarPairs = static_cast<std::uint64_t>(nv) * nv - near;
            }
            for (std::uint32_t e = 0; e < ne; ++e) {
                BemExcitationResult er;
                er.name = request.excitations[e].name;
                er.solvedVelocities = request.excitations[e].faces;
                std::vector<double> sp(2 * nt), tr(2 * nv), dq(2 * nt), ob(2 * no);
                hsbem_detail::check(hsbem_result_surface_pressure(raw, f, e, sp.data()));
                hsbem_detail::check(hsbem_result_pressure_trace(raw, f, e, tr.data()));
                hsbem_detail::check(hsbem_result_normal_derivative(raw, f, e, dq.data()));
                hsbem_detail::check(hsbem_result_observer_pressure(raw, f, e, ob.data()));
                BemBoundaryData b;
                for (std::uint32_t i = 0; i < nt; ++i) {
                    b.surfacePressure.push_back(hsbem_detail::at(sp, i));
                    b.normalDerivative.push_back(hsbem_detail::at(dq, i));
                }
                for (std::uint32_t i = 0; i < nv; ++i)
                    b.pressureTrace.push_back(hsbem_detail::at(tr, i));
                for (std::uint32_t i = 0; i < no; ++i)
                    er.observerPressure.push_back(hsbem_detail::at(ob, i));
                er.setBoundary(std::move(b));
                std::uint32_t count{};
                hsbem_detail::check(hsbem_result_impedances(raw, f, e, &count, nullptr, nullptr, nullptr, nullptr, nullptr));
                std::vector<std::uint32_t> ids(count);
                std::vector<double> area(count), mp(2 * count), vv(2 * count), z(2 * count);
                hsbem_detail::check(
                    hsbem_result_impedances(raw, f, e, &count, ids.data(), area.data(), mp.data(), vv.data(), z.data()));
                for (std::uint32_t i = 0; i < count; ++i)
                    er.impedances.push_back({ids[i], area[i], hsbem_detail::at(mp, i), hsbem_detail::at(vv, i),
                                             hsbem_detail::at(z, i)});
                hsbem_detail::check(hsbem_result_residual(raw, f, e, &er.boundaryResidual, &er.iterations,
                                                          &er.refinementSteps));
                if (fmm) {
                    er.residualEvaluation = ev[hsbem_detail::kDirectAudit] != 0.0
                                                ? BemResidualEvaluation::directGalerkin
                                                : BemResidualEvaluation::higherOrderFastMultipole;
                    er.fastMultipoleOperatorDiscrepancy = ev[hsbem_detail::kAuditDiscrepancy];
                    er.fastMultipoleVerificationAccuracySeed = ev[hsbem_detail::kVerificationSeed];
                    er.farQuadratureDiscrepancy = ev[hsbem_detail::kQuadratureDiscrepancy];
                    er.farQuadratureComparisonDegree = static_cast<unsigned>(ev[hsbem_detail::kEndpointDegree]);
                    er.farQuadratureRefinements = static_cast<unsigned>(ev[hsbem_detail::kPromotions]);
                }
                fr.excitations.push_back(std::move(er));
            }
            out.frequencies.push_back(std::move(fr));
        }
        return out;
    }

  private:
    hsbem_context *context_{};
    hsbem_mesh *mesh_{};
};

} // namespace hornsculpt
