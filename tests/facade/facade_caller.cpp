// A C++ caller of the drop-in header (include/hornsculpt_amd/bem.hpp), written the
// way a user of the reference's CudaBemSolver would (modules/bem/include/hornsculpt/
// bem.hpp:367-388): mesh in, one dense and one fast multipole request, results out.
//
// usage: facade_caller MESH.bin OUT.bin
//   MESH.bin: u32 nv, u32 nt, f64 xyz[3 nv], u32 tri[3 nt], u32 face[nt], u8 driven[nt]
//   OUT.bin:  per request (dense, then fast multipole): f64 re/im pressure trace [nv],
//             f64 re/im observer pressure [nobs], f64 boundary residual, u32 iterations,
//             u32 progress callbacks, u64 m2l links, u64 near pairs, u8 residual evaluation
#include <hornsculpt_amd/bem.hpp>

#include <cstdio>
#include <fstream>
#include <iostream>

namespace hs = hornsculpt;

template <class T> static void read(std::ifstream &in, T *p, std::size_t n) {
    in.read(reinterpret_cast<char *>(p), static_cast<std::streamsize>(n * sizeof(T)));
}
template <class T> static void write(std::ofstream &out, const T *p, std::size_t n) {
    out.write(reinterpret_cast<const char *>(p), static_cast<std::streamsize>(n * sizeof(T)));
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: facade_caller MESH.bin OUT.bin\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    std::uint32_t nv = 0, nt = 0;
    read(in, &nv, 1);
    read(in, &nt, 1);
    std::vector<double> xyz(3 * nv);
    std::vector<std::uint32_t> tri(3 * nt), face(nt);
    std::vector<std::uint8_t> driven(nt);
    read(in, xyz.data(), xyz.size());
    read(in, tri.data(), tri.size());
    read(in, face.data(), face.size());
    read(in, driven.data(), driven.size());
    if (!in) {
        std::cerr << "short mesh file\n";
        return 2;
    }
    hs::SurfaceMesh mesh;
    for (std::uint32_t i = 0; i < nv; ++i)
        mesh.vertices.push_back({xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});
    for (std::uint32_t t = 0; t < nt; ++t)
        mesh.triangles.push_back({{tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]}, face[t],
                                  driven[t] ? hs::BoundaryCondition::driven : hs::BoundaryCondition::wall});
    try {
        hs::CudaBemSolver solver(mesh, hs::CudaStorageOptions{hs::CudaCacheResidency::device});
        std::ofstream out(argv[2], std::ios::binary);
        for (const bool fmm : {false, true}) {
            hs::BemRequest request;
            request.medium = {1.205, 343.0};
            request.frequencies = {1000.0};
            request.excitations = {{"unit", {{1, {1.0, 0.0}}}}};
            request.observers = {{0.0, 0.0, 1.0}, {1.0, 0.0, 0.0}, {0.0, -1.0, 0.0}};
            request.farField = hs::FarFieldObservation{{0.0, 0.0, 0.0}, 1.0};
            request.numerics.relativeResidualTolerance = 1e-10;
            if (fmm) {
                request.numerics.acceleration = hs::BemAcceleration::fastMultipole;
                request.numerics.relativeResidualTolerance = 1e-6;
                request.numerics.fastMultipole.leafSupportCapacity = 32;
                request.numerics.fastMultipole.restartLength = 50;
            }
            std::uint32_t callbacks = 0;
            const hs::BemResult result =
                solver.solve(request, {}, [&callbacks](const hs::BemProgress &p) { callbacks += p.rightHandSideCount; });
            const hs::BemFrequencyResult &point = result.frequencies.at(0);
            const hs::BemExcitationResult &ex = point.excitations.at(0);
            write(out, reinterpret_cast<const double *>(ex.boundary()->pressureTrace.data()), 2 * nv);
            write(out, reinterpret_cast<const double *>(ex.observerPressure.data()), 2 * ex.observerPressure.size());
            write(out, &ex.boundaryResidual, 1);
            const std::uint32_t its = ex.iterations;
            write(out, &its, 1);
            write(out, &callbacks, 1);
            const std::uint64_t links = point.fastMultipole.farInteractionCount,
                                near = point.fastMultipole.pressurePressureNearPairs;
            write(out, &links, 1);
            write(out, &near, 1);
            const std::uint8_t eval = ex.residualEvaluation == hs::BemResidualEvaluation::directGalerkin ? 0 : 1;
            write(out, &eval, 1);
        }
    } catch (const std::exception &e) {
        std::cerr << "facade_caller: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
