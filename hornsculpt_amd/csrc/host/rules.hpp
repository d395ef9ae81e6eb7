// Quadrature rules and triangle-pair point maps, shared by host planning and
// device kernels.
//
// Reference interfaces restated here (modules/bem/galerkin/pair_quadrature.hpp):
//   PairTopology / PairPermutation / classifyPairIds          :22-97
//   mapPairPatch                                               :99-111
//   pairSectorCount / canonicalPairTransform (Sauter-Schwab)   :112-195
//   permutePairPoint                                           :197-207
//   gaussLegendre01 / triangleRule                             pair_quadrature.cpp:685-769
// and the near-projection maps of modules/bem/galerkin/near_projection.hpp:14-255.
#pragma once

#include "vec.hpp"

#include <cstdint>
#include <vector>

namespace hsbem {

enum class Topology : std::uint8_t { disjoint = 0, shared_vertex = 1, shared_edge = 2, coincident = 3 };

// Local vertex permutations: shared vertices first, in increasing global id.
struct Permutation {
    Topology topology{Topology::disjoint};
    std::uint8_t test[3]{0, 1, 2};
    std::uint8_t trial[3]{0, 1, 2};
};

struct PairPoint {
    double test[3]{};
    double trial[3]{};
    double weight{}; // reference measure: the weights of a pair rule sum to 1/4
};

struct TrianglePoint {
    double bary[3]{};
    double weight{}; // weights of a triangle rule sum to 1/2
};

struct GaussPoint {
    double position{}, weight{};
};

// Near-pair patch: three child vertices of each panel in parent barycentrics.
struct PairPatch {
    double test[9]{};
    double trial[9]{};
    double area_fraction{};
    unsigned degree{8};          // 8 (16-point rule) or 14 (64-point tensor rule)
    unsigned projection{};       // 1: inner integral over a projected fan with sinh maps
    unsigned projection_intervals{1};
};

HS_HD inline Permutation classify_pair_ids(const std::uint32_t *test, const std::uint32_t *trial) {
    Permutation out;
    std::uint32_t shared[3]{};
    unsigned count = 0;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j)
            if (test[i] == trial[j] && count < 3)
                shared[count++] = test[i];
    // insertion order does not matter: sort the (at most three) shared ids
    for (unsigned i = 0; i < count; ++i)
        for (unsigned j = i + 1; j < count; ++j)
            if (shared[j] < shared[i]) {
                const auto t = shared[i];
                shared[i] = shared[j];
                shared[j] = t;
            }
    out.topology = count == 3 ? Topology::coincident
                 : count == 2 ? Topology::shared_edge
                 : count == 1 ? Topology::shared_vertex
                              : Topology::disjoint;
    unsigned nx = 0, ny = 0;
    for (unsigned s = 0; s < count; ++s)
        for (unsigned i = 0; i < 3; ++i) {
            if (test[i] == shared[s])
                out.test[nx++] = static_cast<std::uint8_t>(i);
            if (trial[i] == shared[s])
                out.trial[ny++] = static_cast<std::uint8_t>(i);
        }
    for (unsigned i = 0; i < 3; ++i) {
        bool in_test = false, in_trial = false;
        for (unsigned s = 0; s < count; ++s) {
            in_test = in_test || test[i] == shared[s];
            in_trial = in_trial || trial[i] == shared[s];
        }
        if (!in_test)
            out.test[nx++] = static_cast<std::uint8_t>(i);
        if (!in_trial)
            out.trial[ny++] = static_cast<std::uint8_t>(i);
    }
    return out;
}

HS_HD inline unsigned sector_count(Topology t) {
    return t == Topology::coincident ? 6 : t == Topology::shared_edge ? 5 : t == Topology::shared_vertex ? 2 : 1;
}

// Sauter-Schwab / Duffy sectors on {0<=t<=s<=1}^2 (Sauter & Schwab, "Boundary
// Element Methods", ch. 5; the sector table of pair_quadrature.hpp:114-195).
// Returns barycentrics of the canonical (permuted) test and trial points and the
// absolute 4-D Jacobian; multiply by the four Gauss weights. Every coordinate is
// linear in r, which the device kernels exploit for the radial loop.
HS_HD inline PairPoint sauter_schwab_point(Topology topology, unsigned sector, double r, double u,
                                           double v, double w) {
    double xs{}, xt{}, ys{}, yt{}, jac{};
    if (topology == Topology::disjoint) {
        xs = r; xt = r * u; ys = v; yt = v * w; jac = r * v;
    } else if (topology == Topology::shared_vertex) {
        xs = r; xt = r * u; ys = r * v; yt = r * v * w; jac = r * r * r * v;
        if (sector == 1) {
            const double s0 = xs, t0 = xt;
            xs = ys; xt = yt; ys = s0; yt = t0;
        }
    } else if (topology == Topology::shared_edge) {
        jac = r * r * r * u * u;
        switch (sector) {
        case 0: xs = r; xt = r * u * w; ys = r * (1 - u * v); yt = r * u * (1 - v); break;
        case 1: xs = r; xt = r * u; ys = r * (1 - u * v * w); yt = r * u * v * (1 - w); jac *= v; break;
        case 2: xs = r * (1 - u * v); xt = r * u * (1 - v); ys = r; yt = r * u * v * w; jac *= v; break;
        case 3: xs = r * (1 - u * v * w); xt = r * u * v * (1 - w); ys = r; yt = r * u; jac *= v; break;
        default: xs = r * (1 - u * v * w); xt = r * u * (1 - v * w); ys = r; yt = r * u * v; jac *= v; break;
        }
    } else {
        jac = r * r * r * u * u * v;
        switch (sector % 3) {
        case 0: xs = r; xt = r * (1 - u + u * v); ys = r * (1 - u * v * w); yt = r * (1 - u); break;
        case 1: xs = r; xt = r * u * (1 - v + v * w); ys = r * (1 - u * v); yt = r * u * (1 - v); break;
        default: xs = r * (1 - u * v * w); xt = r * u * (1 - v * w); ys = r; yt = r * u * (1 - v); break;
        }
        if (sector >= 3) {
            const double s0 = xs, t0 = xt;
            xs = ys; xt = yt; ys = s0; yt = t0;
        }
    }
    PairPoint p;
    p.test[0] = 1 - xs; p.test[1] = xs - xt; p.test[2] = xt;
    p.trial[0] = 1 - ys; p.trial[1] = ys - yt; p.trial[2] = yt;
    p.weight = jac;
    return p;
}

// Map canonical barycentrics back to each panel's own local vertex order.
HS_HD inline PairPoint unpermute(const PairPoint &p, const Permutation &perm) {
    PairPoint out;
    out.weight = p.weight;
    for (unsigned i = 0; i < 3; ++i) {
        out.test[perm.test[i]] = p.test[i];
        out.trial[perm.trial[i]] = p.trial[i];
    }
    return out;
}

HS_HD inline PairPoint map_patch_point(const TrianglePoint &x, const TrianglePoint &y, const PairPatch &patch) {
    PairPoint p;
    p.weight = x.weight * y.weight * patch.area_fraction;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j) {
            p.test[i] += x.bary[j] * patch.test[3 * j + i];
            p.trial[i] += y.bary[j] * patch.trial[3 * j + i];
        }
    return p;
}

// --------------------------------------------------------------------------
// Projection fan (near_projection.hpp:120-255): the trial triangle is split
// into three signed fans about the orthogonal projection of the test point;
// the angular and radial fan coordinates use sinh maps clustered at the
// projected point, so nearly touching parallel panels integrate accurately.
// Gauss-Legendre literals on [0,1] for orders 8, 14 and 20 (half rules).
HS_HD inline GaussPoint fixed_gauss(unsigned order, unsigned index) {
    const bool mirror = index >= order / 2;
    const unsigned i = mirror ? order - 1 - index : index;
    static constexpr double g8[4][2] = {{.019855071751231856, .05061426814518839},
                                        {.10166676129318658, .11119051722668723},
                                        {.2372337950418355, .1568533229389437},
                                        {.4082826787521751, .181341891689181}};
    static constexpr double g14[7][2] = {{.006858095651593843, .01755973016587453},
                                         {.035782558168213185, .04007904357988019},
                                         {.08639934246511749, .06075928534395156},
                                         {.15635354759415726, .0786015835790968},
                                         {.24237568182092295, .09276919873896892},
                                         {.3404438155360551, .10259923186064775},
                                         {.44597252564632817, .10763192673157888}};
    static constexpr double g20[10][2] = {{.0034357004074525577, .008807003569575289},
                                          {.018014036361043095, .020300714900193525},
                                          {.04388278587433703, .03133602416705452},
                                          {.08044151408889055, .04163837078835237},
                                          {.1268340467699246, .05096505990862024},
                                          {.1819731596367425, .05909726598075916},
                                          {.24456649902458644, .0658443192245883},
                                          {.3131469556422902, .07104805465919095},
                                          {.38610707442917747, .07458649323630191},
                                          {.46173673943325133, .07637669356536299}};
    GaussPoint g;
    if (order == 8) {
        g = {g8[i][0], g8[i][1]};
    } else if (order == 14) {
        g = {g14[i][0], g14[i][1]};
    } else {
        g = {g20[i][0], g20[i][1]};
    }
    if (mirror)
        g.position = 1 - g.position;
    return g;
}

// Minimal panel geometry needed by the projection maps.
struct FanPanel {
    Vec3 v[3];
    Vec3 normal;
    double area;
};

struct FanOuter {
    double test[3]{}, projection[3]{};
    Vec3 relative{};
    double height{}, h{}, weight{};
};
struct FanEdge {
    Vec3 offset{}, direction{};
    double center{}, width{}, lower{}, upper{};
    unsigned edge{}, next{}, opposite{};
};
struct FanRay {
    double boundary[3]{};
    double stretch{}, denominator{}, fan_weight{}, du{}, angular_weight{};
};

HS_HD inline FanOuter fan_outer(const TrianglePoint &outer, const PairPatch &patch, const FanPanel &test,
                                const FanPanel &trial) {
    FanOuter o;
    for (unsigned i = 0; i < 3; ++i)
        for (unsigned j = 0; j < 3; ++j)
            o.test[i] += outer.bary[j] * patch.test[3 * j + i];
    const Vec3 e1 = vsub(trial.v[1], trial.v[0]), e2 = vsub(trial.v[2], trial.v[0]);
    const Vec3 t1 = vsub(test.v[1], test.v[0]), t2 = vsub(test.v[2], test.v[0]);
    o.relative = vsub(test.v[0], trial.v[0]);
    o.relative.x += t1.x * o.test[1] + t2.x * o.test[2];
    o.relative.y += t1.y * o.test[1] + t2.y * o.test[2];
    o.relative.z += t1.z * o.test[1] + t2.z * o.test[2];
    const Vec3 g1 = vcross(e2, trial.normal), g2 = vcross(trial.normal, e1);
    o.projection[1] = vdot(o.relative, g1) / (2 * trial.area);
    o.projection[2] = vdot(o.relative, g2) / (2 * trial.area);
    o.projection[0] = 1 - o.projection[1] - o.projection[2];
    o.height = vdot(o.relative, trial.normal);
    o.h = fabs(o.height);
    o.weight = outer.weight * patch.area_fraction;
    return o;
}

HS_HD inline FanEdge fan_edge(const FanOuter &o, unsigned edge, const FanPanel &trial) {
    FanEdge e;
    e.edge = edge;
    e.next = (edge + 1) % 3;
    e.opposite = (edge + 2) % 3;
    const Vec3 a = vsub(trial.v[edge], trial.v[0]);
    e.direction = vsub(trial.v[e.next], trial.v[edge]);
    e.offset = {a.x - o.relative.x + o.height * trial.normal.x, a.y - o.relative.y + o.height * trial.normal.y,
                a.z - o.relative.z + o.height * trial.normal.z};
    const double len2 = vdot(e.direction, e.direction);
    e.center = -vdot(e.offset, e.direction) / len2;
    const Vec3 perp{e.offset.x + e.center * e.direction.x, e.offset.y + e.center * e.direction.y,
                    e.offset.z + e.center * e.direction.z};
    e.width = sqrt((vdot(perp, perp) + o.h * o.h) / len2);
    e.lower = asinh(-e.center / e.width);
    e.upper = asinh((1 - e.center) / e.width);
    return e;
}

HS_HD inline FanRay fan_ray(const FanOuter &o, const FanEdge &e, GaussPoint angular) {
    FanRay r;
    const double param = e.lower + (e.upper - e.lower) * angular.position;
    const double u = e.center + e.width * sinh(param);
    r.du = e.width * (e.upper - e.lower) * cosh(param);
    const Vec3 ray{e.offset.x + u * e.direction.x, e.offset.y + u * e.direction.y, e.offset.z + u * e.direction.z};
    r.stretch = asinh(sqrt(vdot(ray, ray)) / o.h);
    r.denominator = sinh(r.stretch);
    for (unsigned i = 0; i < 3; ++i)
        r.boundary[i] = i == e.edge ? 1 - u : (i == e.next ? u : 0);
    r.fan_weight = o.weight * o.projection[e.opposite];
    r.angular_weight = angular.weight;
    return r;
}

HS_HD inline PairPoint fan_point(const FanOuter &o, const FanRay &r, GaussPoint radial) {
    PairPoint p;
    const double s = r.stretch, den = r.denominator;
    const double t = s == 0 ? radial.position : sinh(s * radial.position) / den;
    const double dt = s == 0 ? 1 : s * cosh(s * radial.position) / den;
    for (unsigned i = 0; i < 3; ++i) {
        p.test[i] = o.test[i];
        p.trial[i] = o.projection[i] + t * (r.boundary[i] - o.projection[i]);
    }
    p.weight = r.fan_weight * t * dt * r.du * r.angular_weight * radial.weight;
    return p;
}

HS_HD inline unsigned patch_point_count(const PairPatch &patch, unsigned outer_count, unsigned inner_order = 20) {
    const unsigned order = inner_order * patch.projection_intervals;
    return patch.projection ? outer_count * 3 * order * order : outer_count * outer_count;
}

// Point `index` of a patch: tensor pairs for ordinary patches, or
// (outer, edge, angular, radial) for projection patches.
HS_HD inline PairPoint patch_point(const TrianglePoint *rule, unsigned count, const PairPatch &patch,
                                   unsigned index, const FanPanel &test, const FanPanel &trial,
                                   unsigned inner_order = 20) {
    if (!patch.projection)
        return map_patch_point(rule[index / count], rule[index % count], patch);
    const unsigned order = inner_order * patch.projection_intervals;
    const unsigned radial = index % order;
    index /= order;
    const unsigned angular = index % order;
    index /= order;
    const unsigned edge = index % 3, outer = index / 3;
    GaussPoint u = fixed_gauss(inner_order, angular % inner_order);
    GaussPoint v = fixed_gauss(inner_order, radial % inner_order);
    u.position = (angular / inner_order + u.position) / patch.projection_intervals;
    v.position = (radial / inner_order + v.position) / patch.projection_intervals;
    u.weight /= patch.projection_intervals;
    v.weight /= patch.projection_intervals;
    const FanOuter o = fan_outer(rule[outer], patch, test, trial);
    return fan_point(o, fan_ray(o, fan_edge(o, edge, trial), u), v);
}

// Host-only rule construction (rules.cpp).
std::vector<GaussPoint> gauss_legendre01(unsigned order);
std::vector<TrianglePoint> triangle_rule(unsigned degree);

} // namespace hsbem
