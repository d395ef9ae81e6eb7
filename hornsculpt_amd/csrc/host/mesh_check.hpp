// Closed exterior mesh validation (the role of inspectMesh + validateExteriorComponents,
// modules/meshing/include/hornsculpt/meshing.hpp:63 and galerkin/exterior_domain.hpp:13):
// every undirected edge is shared by exactly two triangles with opposite
// orientation, each connected component encloses a positive signed volume
// (outward normals, compensated sum about a local origin), no two triangles cut
// through each other (the role of selfIntersectionPairs, meshing.hpp:18), and no
// component lies inside another (solid-angle test of one vertex).
#pragma once

#include "space.hpp"

#include <cstddef>

namespace hsbem {

void validate_closed_exterior(const SurfaceMesh &mesh);

// Number of triangle pairs that cut through each other, counting stops at `limit`.
std::size_t self_intersecting_pairs(const SurfaceMesh &mesh, std::size_t limit);

} // namespace hsbem
