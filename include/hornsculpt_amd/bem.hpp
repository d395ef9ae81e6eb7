// C++ drop-in for the dense closed-exterior path of hornsculpt::CudaBemSolver
// (reference: modules/bem/include/hornsculpt/bem.hpp:27-388) over the C ABI in
// hsbem.h. Header-only; link libhsbem.so. Types keep the reference names and
// meanings; branches outside this library's scope (subdomains, baffles,
// axisymmetric, LEM coupling) are not declared.
#pragma once

#include "../hsbem.h"

#include <complex>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <stop_token>
#include <string>
#include <vector>

namespace hornsculpt {

struct Vec3 {
    double x{}, y{}, z{};
};
enum class BoundaryCondition : std::uint8_t { wall, driven };
struct Triangle {
    std::uint32_t vertices[3]{};
    std::uint32_t faceId{};
    BoundaryCondition boundary{BoundaryCondition::wall};
};
struct SurfaceMesh {
    std::vector<Vec3> vertices;
    std::vector<Triangle> triangles;
};

struct AcousticMedium {
    double density{1.21};
    double soundSpeed{343.0};
};
struct FaceVelocity {
    std::uint32_t faceId{};
    std::complex<double> normalVelocity{1.0, 0.0};
};
struct BemExcitation {
    std::string name;
    std::vector<FaceVelocity> faces;
};
enum class BemAcceleration { dense };
struct BemNumerics {
    unsigned quadratureOrder{4};
    unsigned singularOrder{4};
    double nearDistanceRatio{1.0};
    double nearRelativeTolerance{1e-4};
    unsigned nearMaximumDepth{6};
    double relativeResidualTolerance{5e-5};
    unsigned maximumRefinementSteps{3};
    BemAcceleration acceleration{BemAcceleration::dense};
};
struct FarFieldObservation {
    Vec3 origin{};
    double referenceDistance{1.0};
};
struct BemRequest {
    AcousticMedium medium;
    std::vector<double> frequencies;
    std::vector<BemExcitation> excitations;
    std::vector<Vec3> observers;
    BemNumerics numerics;
    std::optional<FarFieldObservation> farField;
    std::vector<double> interactionPlanFrequencies;
};
struct FaceImpedance {
    std::uint32_t faceId{};
    double area{};
    std::complex<double> meanPressure, volumeVelocity, volumeImpedance;
};
struct BemBoundaryData {
    std::vector<std::complex<double>> surfacePressure, pressureTrace, normalDerivative;
};
struct BemExcitationResult {
    std::string name;
    std::vector<std::complex<double>> observerPressure;
    std::vector<FaceImpedance> impedances;
    unsigned refinementSteps{};
    double boundaryResidual{};
    unsigned iterations{};
    std::vector<FaceVelocity> solvedVelocities;
    void setBoundary(BemBoundaryData value) { boundary_ = std::make_shared<const BemBoundaryData>(std::move(value)); }
    const std::shared_ptr<const BemBoundaryData> &boundary() const noexcept { return boundary_; }

  private:
    std::shared_ptr<const BemBoundaryData> boundary_;
};
struct BemFrequencyResult {
    double frequency{};
    std::vector<BemExcitationResult> excitations;
    double assemblyMilliseconds{}, solveMilliseconds{}, factorizationMilliseconds{}, triangularSolveMilliseconds{},
        fieldMilliseconds{}, residualCheckMilliseconds{}, interactionPlanMilliseconds{};
};
struct BemResult {
    std::string device;
    std::size_t triangleCount{}, pressureDofCount{}, peakDeviceBytes{};
    std::vector<BemFrequencyResult> frequencies;
    BemAcceleration acceleration{BemAcceleration::dense};
};
enum class BemSolvePhase { preparing, nearConstruction, farPreparation, krylovSolve, residualAudit, fieldEvaluation,
                           frequencyComplete };
struct BemProgress {
    std::size_t completedFrequencies{}, totalFrequencies{};
    double frequency{};
    BemSolvePhase phase{BemSolvePhase::preparing};
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
} // namespace hsbem_detail

class CudaBemSolver {
  public:
    explicit CudaBemSolver(SurfaceMesh mesh) {
        hsbem_detail::check(hsbem_context_create(0, &context_));
        std::vector<double> xyz;
        std::vector<std::uint32_t> tri, faces;
        std::vector<std::uint8_t> bc;
        for (const auto &v : mesh.vertices)
            xyz.insert(xyz.end(), {v.x, v.y, v.z});
        for (const auto &t : mesh.triangles) {
            tri.insert(tri.end(), {t.vertices[0], t.vertices[1], t.vertices[2]});
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
        rq.numerics.quadrature_order = request.numerics.quadratureOrder;
        rq.numerics.singular_order = request.numerics.singularOrder;
        rq.numerics.near_distance_ratio = request.numerics.nearDistanceRatio;
        rq.numerics.near_relative_tolerance = request.numerics.nearRelativeTolerance;
        rq.numerics.near_maximum_depth = request.numerics.nearMaximumDepth;
        rq.numerics.relative_residual_tolerance = request.numerics.relativeResidualTolerance;
        rq.numerics.maximum_refinement_steps = request.numerics.maximumRefinementSteps;
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
        } bridge{&progress};
        auto trampoline = [](void *user, std::uint32_t done, std::uint32_t total, double f, std::int32_t phase) {
            auto *b = static_cast<Bridge *>(user);
            if (*b->fn)
                (*b->fn)(BemProgress{done, total, f, static_cast<BemSolvePhase>(phase)});
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
