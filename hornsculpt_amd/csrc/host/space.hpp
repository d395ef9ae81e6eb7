// P1 pressure / DP0 flux Galerkin space on a closed triangle mesh.
//
// Mirrors hornsculpt::galerkin::Space (modules/bem/galerkin/space.hpp:11-46):
// pressure DOF i is mesh vertex i, flux DOF t is triangle t; per panel the
// area, unit normal, centroid, longest edge, P1 gradients and surface curls.
#pragma once

#include "vec.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace hsbem {

enum class Boundary : std::uint8_t { wall = 0, driven = 1 };

struct MeshTriangle {
    std::array<std::uint32_t, 3> vertices{};
    std::uint32_t face_id{};
    Boundary boundary{Boundary::wall};
};

struct SurfaceMesh {
    std::vector<Vec3> vertices;
    std::vector<MeshTriangle> triangles;
};

struct Panel {
    std::array<std::uint32_t, 3> nodes{};
    std::array<Vec3, 3> vertices{};
    Vec3 centroid, normal;
    double area{}, max_edge{};
    std::array<Vec3, 3> gradients{}, curls{};
    std::uint32_t face_id{};
    Boundary boundary{Boundary::wall};
};

struct Incidence {
    std::uint32_t panel{};
    std::uint8_t local{};
};

struct Space {
    std::vector<Vec3> vertices;
    std::vector<Panel> panels;
    std::vector<std::size_t> incidence_offsets; // vertex -> [first, last) into incidences
    std::vector<Incidence> incidences;
    std::vector<double> lumped_mass;            // sum of A/3 over incident panels
    double surface_area{};
};

// Throws std::invalid_argument on non-finite coordinates, bad indices,
// degenerate panels or orphan vertices (space.cpp:27-107 of the reference).
Space build_space(const SurfaceMesh &mesh);

} // namespace hsbem
