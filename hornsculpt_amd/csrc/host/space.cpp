// Galerkin space construction. Follows modules/bem/galerkin/space.cpp:27-107:
// the same degeneracy threshold (128 eps h^2 on twice the area), the same
// centroid form a+(ab+ac)/3, gradients n x (opposite edge)/(2A), curls n x grad,
// lumped P1 mass A/3 per incident panel and a compensated surface-area sum.
#include "space.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>

namespace hsbem {
namespace {
double hypot3(Vec3 v) { return std::hypot(v.x, v.y, v.z); }
Vec3 over(Vec3 v, double d) { return {v.x / d, v.y / d, v.z / d}; }
bool finite3(Vec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }
} // namespace

Space build_space(const SurfaceMesh &mesh) {
    if (mesh.vertices.empty() || mesh.triangles.empty())
        throw std::invalid_argument("the surface mesh is empty");
    if (mesh.vertices.size() > std::numeric_limits<std::uint32_t>::max() ||
        mesh.triangles.size() > std::numeric_limits<std::uint32_t>::max())
        throw std::invalid_argument("the mesh has too many triangles for 32-bit indices");
    for (const auto &v : mesh.vertices)
        if (!finite3(v))
            throw std::invalid_argument("mesh vertex coordinates must be finite");

    Space space;
    space.vertices = mesh.vertices;
    const std::size_t nv = mesh.vertices.size();
    space.panels.reserve(mesh.triangles.size());
    space.incidence_offsets.assign(nv + 1, 0);
    space.lumped_mass.assign(nv, 0.0);
    double carry = 0.0; // Kahan compensation of the surface-area sum
    for (const auto &tri : mesh.triangles) {
        Panel p;
        p.nodes = tri.vertices;
        p.face_id = tri.face_id;
        p.boundary = tri.boundary;
        for (int l = 0; l < 3; ++l) {
            if (p.nodes[l] >= nv)
                throw std::invalid_argument("a triangle refers to a vertex that does not exist");
            p.vertices[l] = space.vertices[p.nodes[l]];
        }
        const Vec3 a = p.vertices[0], b = p.vertices[1], c = p.vertices[2];
        const Vec3 ab = vsub(b, a), ac = vsub(c, a), bc = vsub(c, b);
        const Vec3 twice_area_vector = vcross(ab, ac);
        const double twice_area = hypot3(twice_area_vector);
        p.max_edge = std::max({hypot3(ab), hypot3(ac), hypot3(bc)});
        const double threshold =
            128.0 * std::numeric_limits<double>::epsilon() * p.max_edge * p.max_edge;
        if (!std::isfinite(twice_area) || !std::isfinite(p.max_edge) || twice_area <= threshold)
            throw std::invalid_argument("a triangle has (numerically) zero area");
        p.area = 0.5 * twice_area;
        p.normal = over(twice_area_vector, twice_area);
        p.centroid = {a.x + (ab.x + ac.x) / 3.0, a.y + (ab.y + ac.y) / 3.0,
                      a.z + (ab.z + ac.z) / 3.0};
        const Vec3 opposite[3] = {bc, vsub(a, c), ab};
        for (int l = 0; l < 3; ++l) {
            p.gradients[l] = over(vcross(p.normal, opposite[l]), twice_area);
            p.curls[l] = vcross(p.normal, p.gradients[l]);
            if (!finite3(p.gradients[l]) || !finite3(p.curls[l]))
                throw std::invalid_argument("a triangle's basis gradient is not finite");
            ++space.incidence_offsets[static_cast<std::size_t>(p.nodes[l]) + 1];
            space.lumped_mass[p.nodes[l]] += p.area / 3.0;
        }
        const double y = p.area - carry;
        const double t = space.surface_area + y;
        carry = (t - space.surface_area) - y;
        space.surface_area = t;
        space.panels.push_back(p);
    }
    for (std::size_t v = 0; v < nv; ++v) {
        if (space.incidence_offsets[v + 1] == 0)
            throw std::invalid_argument("a vertex belongs to no triangle");
        space.incidence_offsets[v + 1] += space.incidence_offsets[v];
    }
    space.incidences.resize(space.incidence_offsets.back());
    std::vector<std::size_t> cursor(space.incidence_offsets.begin(), space.incidence_offsets.end() - 1);
    for (std::size_t t = 0; t < space.panels.size(); ++t)
        for (int l = 0; l < 3; ++l)
            space.incidences[cursor[space.panels[t].nodes[l]]++] = {static_cast<std::uint32_t>(t),
                                                                   static_cast<std::uint8_t>(l)};
    return space;
}

} // namespace hsbem
