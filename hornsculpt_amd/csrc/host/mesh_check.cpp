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

// Sign of the volume of tetrahedron (a, b, c, d): > 0 when d lies on the side of
// plane (a, b, c) that its counter-clockwise normal points to.
double orient(const Vec3 &a, const Vec3 &b, const Vec3 &c, const Vec3 &d) {
    return vdot(vsub(d, a), vcross(vsub(b, a), vsub(c, a)));
}

// Coplanar segment-triangle overlap, in the coordinate plane that drops the
// normal's largest component.
bool coplanar_overlap(const Vec3 &p, const Vec3 &q, const Vec3 tri[3], const Vec3 &n) {
    const double ax = std::abs(n.x), ay = std::abs(n.y), az = std::abs(n.z);
    auto flat = [&](const Vec3 &v) {
        return ax >= ay && ax >= az ? std::pair{v.y, v.z} : ay >= az ? std::pair{v.x, v.z} : std::pair{v.x, v.y};
    };
    using P2 = std::pair<double, double>;
    auto cross2 = [](P2 o, P2 a, P2 b) { return (a.first - o.first) * (b.second - o.second) - (a.second - o.second) * (b.first - o.first); };
    const P2 P = flat(p), Q = flat(q), T[3] = {flat(tri[0]), flat(tri[1]), flat(tri[2])};
    const double area = cross2(T[0], T[1], T[2]);
    auto inside = [&](P2 x) {
        const double a = cross2(T[0], T[1], x), b = cross2(T[1], T[2], x), c = cross2(T[2], T[0], x);
        return area > 0 ? (a > 0 && b > 0 && c > 0) : (a < 0 && b < 0 && c < 0);
    };
    if (inside(P) || inside(Q))
        return true;
    for (int e = 0; e < 3; ++e) { // proper crossing of PQ with an edge
        const P2 A = T[e], B = T[(e + 1) % 3];
        const double d1 = cross2(A, B, P), d2 = cross2(A, B, Q), d3 = cross2(P, Q, A), d4 = cross2(P, Q, B);
        if (((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0)) && ((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0)))
            return true;
    }
    return false;
}

// Segment pq passes through the plane of triangle tri strictly between its end
// points at a point of the closed triangle (a crossing through an edge of tri
// counts: it would otherwise fall between two neighbouring triangles).
bool segment_crosses(const Vec3 &p, const Vec3 &q, const Vec3 tri[3]) {
    const double sp = orient(tri[0], tri[1], tri[2], p), sq = orient(tri[0], tri[1], tri[2], q);
    if ((sp > 0 && sq > 0) || (sp < 0 && sq < 0))
        return false;
    if (sp == 0 && sq == 0)
        return coplanar_overlap(p, q, tri, vcross(vsub(tri[1], tri[0]), vsub(tri[2], tri[0])));
    if (sp == 0 || sq == 0)
        return false; // an endpoint on the plane: a contact, not a crossing
    const double o0 = orient(p, q, tri[0], tri[1]), o1 = orient(p, q, tri[1], tri[2]), o2 = orient(p, q, tri[2], tri[0]);
    return (o0 >= 0 && o1 >= 0 && o2 >= 0) || (o0 <= 0 && o1 <= 0 && o2 <= 0);
}

// Pairs of triangles that cut through each other. Triangles with no common vertex
// intersect when an edge of one crosses the other; triangles with one common
// vertex when the opposite edge of either crosses the other; triangles with a
// common edge when they fold onto each other (coplanar, opposite vertices on the
// same side). Candidate pairs come from a uniform grid over the bounding boxes
// (a pair is tested in the one cell holding the low corner of their box overlap).
std::size_t count_self_intersections(const SurfaceMesh &mesh, std::size_t limit) {
    const std::size_t nt = mesh.triangles.size();
    if (nt < 2)
        return 0;
    std::vector<Vec3> lo(nt), hi(nt);
    double edge_sum = 0;
    Vec3 glo{INFINITY, INFINITY, INFINITY}, ghi{-INFINITY, -INFINITY, -INFINITY};
    for (std::size_t i = 0; i < nt; ++i) {
        const auto &t = mesh.triangles[i];
        const Vec3 a = mesh.vertices[t.vertices[0]], b = mesh.vertices[t.vertices[1]], c = mesh.vertices[t.vertices[2]];
        lo[i] = {std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
        hi[i] = {std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
        glo = {std::min(glo.x, lo[i].x), std::min(glo.y, lo[i].y), std::min(glo.z, lo[i].z)};
        ghi = {std::max(ghi.x, hi[i].x), std::max(ghi.y, hi[i].y), std::max(ghi.z, hi[i].z)};
        edge_sum += std::sqrt(vnorm2(vsub(b, a))) + std::sqrt(vnorm2(vsub(c, b))) + std::sqrt(vnorm2(vsub(a, c)));
    }
    const double extent = std::max({ghi.x - glo.x, ghi.y - glo.y, ghi.z - glo.z});
    double cell = std::max(2 * edge_sum / (3.0 * static_cast<double>(nt)), extent / 512);
    if (!(cell > 0) || !std::isfinite(cell))
        return 0;
    auto cell_of = [&](double v, double origin) { return static_cast<std::int64_t>(std::floor((v - origin) / cell)); };
    const std::int64_t nx = cell_of(ghi.x, glo.x) + 1, ny = cell_of(ghi.y, glo.y) + 1;
    std::vector<std::pair<std::uint64_t, std::uint32_t>> entries;
    for (std::size_t i = 0; i < nt; ++i)
        for (std::int64_t z = cell_of(lo[i].z, glo.z); z <= cell_of(hi[i].z, glo.z); ++z)
            for (std::int64_t y = cell_of(lo[i].y, glo.y); y <= cell_of(hi[i].y, glo.y); ++y)
                for (std::int64_t x = cell_of(lo[i].x, glo.x); x <= cell_of(hi[i].x, glo.x); ++x)
                    entries.push_back({static_cast<std::uint64_t>((z * ny + y) * nx + x), static_cast<std::uint32_t>(i)});
    std::sort(entries.begin(), entries.end());
    std::size_t found = 0;
    for (std::size_t b = 0; b < entries.size();) {
        std::size_t e = b;
        while (e < entries.size() && entries[e].first == entries[b].first)
            ++e;
        const std::uint64_t key = entries[b].first;
        for (std::size_t u = b; u < e; ++u)
            for (std::size_t w = u + 1; w < e; ++w) {
                const std::uint32_t i = entries[u].second, j = entries[w].second;
                const Vec3 olo{std::max(lo[i].x, lo[j].x), std::max(lo[i].y, lo[j].y), std::max(lo[i].z, lo[j].z)};
                if (olo.x > std::min(hi[i].x, hi[j].x) || olo.y > std::min(hi[i].y, hi[j].y) ||
                    olo.z > std::min(hi[i].z, hi[j].z))
                    continue;
                const std::uint64_t home = static_cast<std::uint64_t>(
                    (cell_of(olo.z, glo.z) * ny + cell_of(olo.y, glo.y)) * nx + cell_of(olo.x, glo.x));
                if (home != key)
                    continue;
                const auto &ti = mesh.triangles[i].vertices, &tj = mesh.triangles[j].vertices;
                const Vec3 A[3] = {mesh.vertices[ti[0]], mesh.vertices[ti[1]], mesh.vertices[ti[2]]};
                const Vec3 B[3] = {mesh.vertices[tj[0]], mesh.vertices[tj[1]], mesh.vertices[tj[2]]};
                int shared = 0, si = -1, sj = -1;
                for (int a = 0; a < 3; ++a)
                    for (int c = 0; c < 3; ++c)
                        if (ti[a] == tj[c]) {
                            ++shared;
                            si = a;
                            sj = c;
                        }
                bool cut = false;
                if (shared == 0) {
                    for (int a = 0; a < 3 && !cut; ++a)
                        cut = segment_crosses(A[a], A[(a + 1) % 3], B) || segment_crosses(B[a], B[(a + 1) % 3], A);
                } else if (shared == 1) {
                    cut = segment_crosses(A[(si + 1) % 3], A[(si + 2) % 3], B) ||
                          segment_crosses(B[(sj + 1) % 3], B[(sj + 2) % 3], A);
                } else if (shared == 2) {
                    int oi = 0, oj = 0; // the vertices off the common edge
                    for (int a = 0; a < 3; ++a) {
                        if (ti[a] != tj[0] && ti[a] != tj[1] && ti[a] != tj[2])
                            oi = a;
                        if (tj[a] != ti[0] && tj[a] != ti[1] && tj[a] != ti[2])
                            oj = a;
                    }
                    if (orient(A[0], A[1], A[2], B[oj]) == 0) { // coplanar: folded when both lie on one side
                        const Vec3 n = vcross(vsub(A[1], A[0]), vsub(A[2], A[0]));
                        const int e0 = (oi + 1) % 3, e1 = (oi + 2) % 3;
                        const double s1 = vdot(vcross(vsub(A[e1], A[e0]), vsub(A[oi], A[e0])), n),
                                     s2 = vdot(vcross(vsub(A[e1], A[e0]), vsub(B[oj], A[e0])), n);
                        cut = (s1 > 0 && s2 > 0) || (s1 < 0 && s2 < 0);
                    }
                }
                if (cut && ++found >= limit)
                    return found;
            }
        b = e;
    }
    return found;
}

} // namespace

std::size_t self_intersecting_pairs(const SurfaceMesh &mesh, std::size_t limit) {
    return count_self_intersections(mesh, limit);
}

void validate_closed_exterior(const SurfaceMesh &mesh) {
    const std::size_t nt = mesh.triangles.size();
    std::map<std::pair<std::uint32_t, std::uint32_t>, int> directed;
    for (const auto &t : mesh.triangles)
        for (int l = 0; l < 3; ++l) {
            const std::uint32_t a = t.vertices[l], b = t.vertices[(l + 1) % 3];
            if (a >= mesh.vertices.size() || b >= mesh.vertices.size() || a == b)
                throw std::invalid_argument("the mesh must be closed, outward-oriented and free of self-intersections");
            if (++directed[{a, b}] > 1)
                throw std::invalid_argument("the mesh must be closed, outward-oriented and free of self-intersections");
        }
    for (const auto &[edge, count] : directed)
        if (!directed.count({edge.second, edge.first}))
            throw std::invalid_argument("the mesh must be closed, outward-oriented and free of self-intersections");
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
    if (count_self_intersections(mesh, 1))
        throw std::invalid_argument("BEM mesh intersects itself: two of its triangles cut through each other.");
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
