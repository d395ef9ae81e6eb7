#include "mesh_check.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <numbers>
#include <numeric>
#include <stdexcept>
#include <utility>
#include <vector>

namespace hsbem {
namespace {

std::uint32_t find_root(std::vector<std::uint32_t> &parent, std::uint32_t x) {
    while (parent[x] != x) {
        parent[x] = parent[parent[x]];
        x = parent[x];
    }
    return x;
}

// Solid angle subtended by triangle (a,b,c) at the origin (Van Oosterom & Strackee 1983).
double solid_angle(Vec3 a, Vec3 b, Vec3 c) {
    const double la = std::sqrt(vnorm2(a)), lb = std::sqrt(vnorm2(b)), lc = std::sqrt(vnorm2(c));
    const double den = la * lb * lc + vdot(a, b) * lc + vdot(b, c) * la + vdot(c, a) * lb;
    return 2 * std::atan2(vdot(a, vcross(b, c)), den);
}

} // namespace

void validate_closed_exterior(const SurfaceMesh &mesh) {
    const std::size_t nt = mesh.triangles.size();
    std::map<std::pair<std::uint32_t, std::uint32_t>, int> directed;
    for (const auto &t : mesh.triangles)
        for (int l = 0; l < 3; ++l) {
            const std::uint32_t a = t.vertices[l], b = t.vertices[(l + 1) % 3];
            if (a >= mesh.vertices.size() || b >= mesh.vertices.size() || a == b)
                throw std::invalid_argument("BEM requires a closed, outward-oriented, non-intersecting mesh.");
            if (++directed[{a, b}] > 1)
                throw std::invalid_argument("BEM requires a closed, outward-oriented, non-intersecting mesh.");
        }
    for (const auto &[edge, count] : directed)
        if (!directed.count({edge.second, edge.first}))
            throw std::invalid_argument("BEM requires a closed, outward-oriented, non-intersecting mesh.");
    // connected components over shared vertices
    std::vector<std::uint32_t> parent(mesh.vertices.size());
    std::iota(parent.begin(), parent.end(), 0u);
    for (const auto &t : mesh.triangles)
        for (int l = 1; l < 3; ++l) {
            const auto a = find_root(parent, t.vertices[0]), b = find_root(parent, t.vertices[l]);
            if (a != b)
                parent[a] = b;
        }
    std::map<std::uint32_t, std::vector<std::size_t>> components;
    for (std::size_t i = 0; i < nt; ++i)
        components[find_root(parent, mesh.triangles[i].vertices[0])].push_back(i);
    std::vector<std::uint32_t> probe;
    for (const auto &[root, tris] : components) {
        const Vec3 origin = mesh.vertices[mesh.triangles[tris.front()].vertices[0]];
        double volume = 0, carry = 0;
        for (auto i : tris) {
            const auto &t = mesh.triangles[i];
            const Vec3 a = vsub(mesh.vertices[t.vertices[0]], origin), b = vsub(mesh.vertices[t.vertices[1]], origin),
                       c = vsub(mesh.vertices[t.vertices[2]], origin);
            const double y = vdot(a, vcross(b, c)) / 6 - carry, s = volume + y;
            carry = (s - volume) - y;
            volume = s;
        }
        if (!(volume > 0))
            throw std::invalid_argument("BEM mesh component has non-positive enclosed volume (inward orientation).");
        probe.push_back(mesh.triangles[tris.front()].vertices[0]);
    }
    // nesting: a probe vertex of one component inside another component
    std::size_t ci = 0;
    for (const auto &[root_i, tris_i] : components) {
        std::size_t cj = 0;
        for (const auto &[root_j, tris_j] : components) {
            if (ci != cj) {
                const Vec3 p = mesh.vertices[probe[ci]];
                double omega = 0;
                for (auto t : tris_j) {
                    const auto &tr = mesh.triangles[t];
                    omega += solid_angle(vsub(mesh.vertices[tr.vertices[0]], p), vsub(mesh.vertices[tr.vertices[1]], p),
                                         vsub(mesh.vertices[tr.vertices[2]], p));
                }
                if (std::abs(omega) > 2 * std::numbers::pi)
                    throw std::invalid_argument("BEM mesh has nested surfaces; enclosed air cavities are unsupported.");
            }
            ++cj;
        }
        ++ci;
    }
}

} // namespace hsbem
