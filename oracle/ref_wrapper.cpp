// Test-infrastructure only: a flat C interface over the reference planner
// (modules/bem/galerkin/{space,pair_quadrature,interaction_plan}.cpp), compiled
// from the reference sources by oracle/Makefile into oracle/_ref/. The product
// never links or loads it; tests use it to pin the product planner and the
// oracle's rules, Sauter-Schwab points and near-patch moments.
#include "galerkin/interaction_plan.hpp"
#include "galerkin/near_projection.hpp"
#include "galerkin/pair_quadrature.hpp"
#include "galerkin/regular_qualification.hpp"
#include "galerkin/space.hpp"
#include "fmm/rcr_work.hpp"
#include "fmm/spherical_rotation.hpp"
#include "fmm/support_level_plan.hpp"
#include "fmm/support_near_plan.hpp"
#include "fmm/support_tree.hpp"
#include "fmm/wideband_transfer.hpp"
#include "fmm/scaled_translation.hpp"
#include <hornsculpt/acoustic_analysis.hpp>
#include <hornsculpt/acoustic_frame.hpp>

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

using namespace hornsculpt;
using namespace hornsculpt::galerkin;

namespace {
thread_local std::string last_error;

SurfaceMesh make_mesh(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri) {
    SurfaceMesh m;
    m.vertices.resize(nv);
    for (std::uint32_t i = 0; i < nv; ++i)
        m.vertices[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    m.triangles.resize(nt);
    for (std::uint32_t t = 0; t < nt; ++t)
        m.triangles[t] = {{tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]}, 1, BoundaryCondition::driven};
    return m;
}

template <class F> int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 1;
    }
}

void pack_patch(const PairQuadraturePatch &p, double *out) {
    for (int i = 0; i < 9; ++i) {
        out[i] = p.test[i];
        out[9 + i] = p.trial[i];
    }
    out[18] = p.areaFraction;
    out[19] = p.degree;
    out[20] = p.projection;
    out[21] = p.projectionIntervals;
}
PairQuadraturePatch unpack_patch(const double *in) {
    PairQuadraturePatch p{};
    for (int i = 0; i < 9; ++i) {
        p.test[i] = in[i];
        p.trial[i] = in[9 + i];
    }
    p.areaFraction = in[18];
    p.degree = static_cast<unsigned>(in[19]);
    p.projection = static_cast<unsigned>(in[20]);
    p.projectionIntervals = static_cast<unsigned>(in[21]);
    return p;
}

struct PlanHandle {
    InteractionPlan plan;
};

// Same flat layout as the product's hsbem_support_tree_build (see include/hsbem.h).
std::vector<double> tree_blob(const wideband::SupportTree &t, std::uint32_t nv, std::uint32_t nt) {
    std::vector<double> b;
    auto box = [&](const wideband::SupportBounds &x) {
        for (double v : {x.lower.x, x.lower.y, x.lower.z, x.upper.x, x.upper.y, x.upper.z, x.maximumPanelEdge})
            b.push_back(v);
    };
    const auto &pp = t.pressurePressure, &pf = t.pressureFlux;
    for (double v : {double(t.nodes.size()), double(t.levels.size()), double(pp.nearInteractions.size()),
                     double(pp.farInteractions.size()), double(pf.nearInteractions.size()),
                     double(pf.farInteractions.size()), double(pp.farLevelPairs.size()), double(pf.farLevelPairs.size()),
                     double(nv), double(nt), double(pp.representedNearPairs), double(pp.representedFarPairs),
                     double(pf.representedNearPairs), double(pf.representedFarPairs)})
        b.push_back(v);
    for (const auto &n : t.nodes) {
        for (double v : {n.center.x, n.center.y, n.center.z, n.halfWidth})
            b.push_back(v);
        for (auto c : n.coordinates)
            b.push_back(c);
        b.push_back(n.level);
        b.push_back(n.parent);
        for (auto c : n.children)
            b.push_back(c);
        for (auto v : {n.firstPressure, n.pressureCount, n.firstFlux, n.fluxCount})
            b.push_back(v);
        box(n.pressureBounds);
        box(n.fluxBounds);
        b.push_back(n.pressureRadius);
        b.push_back(n.fluxRadius);
        b.push_back(n.activeChannels);
    }
    for (const auto *v : {&t.pressureOrder, &t.fluxOrder, &t.pressureLeaf, &t.fluxLeaf})
        for (auto x : *v)
            b.push_back(x);
    for (const auto *l : {&pp.nearInteractions, &pp.farInteractions, &pf.nearInteractions, &pf.farInteractions})
        for (const auto &p : *l) {
            b.push_back(p.target);
            b.push_back(p.source);
        }
    for (const auto *l : {&pp.farLevelPairs, &pf.farLevelPairs})
        for (const auto &e : *l)
            for (double v : {double(e.targetLevel), double(e.sourceLevel), double(e.interactions),
                             double(e.representedPairs)})
                b.push_back(v);
    for (const auto &l : t.levels) {
        b.push_back(double(l.nodes.size()));
        for (auto id : l.nodes)
            b.push_back(id);
        b.push_back(l.maximumSupportRadius);
        b.push_back(l.maximumActiveSupportRadius);
        for (auto v : {l.pressureOutgoingNodes, l.fluxOutgoingNodes, l.pressureIncomingNodes,
                       l.pressurePressureFarInteractions, l.pressureFluxFarInteractions, l.pressurePressureOffsets,
                       l.pressureFluxOffsets, l.unionOffsets, l.pressurePressureDistanceClasses,
                       l.pressureFluxDistanceClasses, l.pressurePressurePolarClasses, l.pressureFluxPolarClasses,
                       l.unionPolarClasses})
            b.push_back(double(v));
        b.push_back(double(l.squaredOffsetClasses.size()));
        for (auto v : l.squaredOffsetClasses)
            b.push_back(double(v));
    }
    return b;
}
} // namespace

extern "C" {

const char *ref_last_error() { return last_error.c_str(); }

// Panel data per triangle: area, max edge, normal[3], centroid[3], curls[9], gradients[9] (26 doubles),
// and the lumped mass per vertex.
int ref_space(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, double *panel_out,
              double *lumped_out) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        for (std::uint32_t t = 0; t < nt; ++t) {
            const auto &p = s.panels[t];
            double *o = panel_out + 26 * t;
            o[0] = p.area;
            o[1] = p.maximumEdgeLength;
            o[2] = p.normal.x, o[3] = p.normal.y, o[4] = p.normal.z;
            o[5] = p.centroid.x, o[6] = p.centroid.y, o[7] = p.centroid.z;
            for (int l = 0; l < 3; ++l) {
                o[8 + 3 * l] = p.curls[l].x, o[9 + 3 * l] = p.curls[l].y, o[10 + 3 * l] = p.curls[l].z;
                o[17 + 3 * l] = p.gradients[l].x, o[18 + 3 * l] = p.gradients[l].y, o[19 + 3 * l] = p.gradients[l].z;
            }
        }
        for (std::uint32_t v = 0; v < nv; ++v)
            lumped_out[v] = s.lumpedMass[v];
    });
}

int ref_triangle_rule(unsigned degree, double *out, unsigned *count) {
    return guarded([&] {
        const auto r = triangleRule(degree);
        *count = static_cast<unsigned>(r.size());
        if (out)
            for (std::size_t i = 0; i < r.size(); ++i) {
                for (int j = 0; j < 3; ++j)
                    out[4 * i + j] = r[i].barycentric[j];
                out[4 * i + 3] = r[i].weight;
            }
    });
}

int ref_gauss(unsigned order, double *out) {
    return guarded([&] {
        const auto g = gaussLegendre01(order);
        for (std::size_t i = 0; i < g.size(); ++i) {
            out[2 * i] = g[i].position;
            out[2 * i + 1] = g[i].weight;
        }
    });
}

// Explicit pair rule of makePairQuadrature for the ordered pair (a, b) of a
// mesh: touching pairs give Sauter-Schwab points in original vertex order.
// Points: test[3], trial[3], weight (7 doubles each).
int ref_pair_points(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, std::uint32_t a,
                    std::uint32_t b, unsigned singular_order, double *out, unsigned *count, int *topology) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        PairQuadratureSettings q;
        q.singularOrder = singular_order;
        q.nearDistanceRatio = 0;
        const auto rule = makePairQuadrature(s.panels[a], s.panels[b], q, 0.0);
        *topology = static_cast<int>(rule.topology);
        *count = static_cast<unsigned>(rule.points.size());
        if (out)
            for (std::size_t i = 0; i < rule.points.size(); ++i) {
                for (int j = 0; j < 3; ++j) {
                    out[7 * i + j] = rule.points[i].test[j];
                    out[7 * i + 3 + j] = rule.points[i].trial[j];
                }
                out[7 * i + 6] = rule.points[i].weight;
            }
    });
}

// estimatePairPatchCpu: 36 complex moments (re, im interleaved) and errors.
int ref_estimate_patch(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                       std::uint32_t a, std::uint32_t b, const double *patch, double k, double *moments,
                       double *errors) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        const auto e = estimatePairPatchCpu(s.panels[a], s.panels[b], unpack_patch(patch), k);
        for (int i = 0; i < 36; ++i) {
            moments[2 * i] = e.integral[i].real();
            moments[2 * i + 1] = e.integral[i].imag();
        }
        errors[0] = e.absoluteError;
        errors[1] = e.innerError;
        errors[2] = e.outerError;
    });
}

// One projection/tensor patch point as the device evaluation maps it.
int ref_patch_point(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri, std::uint32_t a,
                    std::uint32_t b, const double *patch, unsigned index, double *out) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        const auto p = unpack_patch(patch);
        const auto rule = triangleRule(p.degree == 14 ? 14 : 8);
        const auto pt = mapNearPatchPoint(rule.data(), static_cast<unsigned>(rule.size()), p, index, s.panels[a],
                                          s.panels[b]);
        for (int j = 0; j < 3; ++j) {
            out[j] = pt.test[j];
            out[3 + j] = pt.trial[j];
        }
        out[6] = pt.weight;
    });
}

int ref_qualified_degree(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                         std::uint32_t a, std::uint32_t b, double k, double tolerance, unsigned minimum,
                         unsigned *degree) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        *degree = qualifiedRegularDegree(makeRegularPanelBounds(s.panels[a]), makeRegularPanelBounds(s.panels[b]), k,
                                         tolerance, minimum);
    });
}

void *ref_plan_create(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                      unsigned regular_order, unsigned singular_order, double near_ratio, double tolerance,
                      unsigned depth, double k_max, unsigned threads) {
    PlanHandle *h = nullptr;
    const int rc = guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        PairQuadratureSettings q;
        q.regularOrder = regular_order;
        q.singularOrder = singular_order;
        q.nearDistanceRatio = near_ratio;
        q.relativeTolerance = tolerance;
        q.maximumSubdivisionDepth = depth;
        h = new PlanHandle{makeInteractionPlan(s, q, k_max, {}, threads, std::size_t{16} << 30)};
    });
    return rc ? nullptr : h;
}

void ref_plan_sizes(void *handle, std::uint64_t *sizes) {
    const auto &p = static_cast<PlanHandle *>(handle)->plan;
    sizes[0] = p.touching.size();
    sizes[1] = p.near.size();
    sizes[2] = p.patches.size();
    sizes[3] = p.wholePanelDegrees.size();
    sizes[4] = p.candidatePairs;
    sizes[5] = p.maximumDepth;
}

// touching: 8 uint32 per job (test, trial, topology, perm test[3], perm trial[3] packed as 3+3 -> 8 total)
void ref_plan_copy(void *handle, std::uint32_t *touching, std::uint32_t *near, double *patches,
                   std::uint8_t *whole, std::uint32_t *offsets, std::uint32_t *trials, double *max_error) {
    const auto &p = static_cast<PlanHandle *>(handle)->plan;
    for (std::size_t i = 0; i < p.touching.size(); ++i) {
        const auto &j = p.touching[i];
        std::uint32_t *o = touching + 9 * i;
        o[0] = j.test;
        o[1] = j.trial;
        o[2] = static_cast<std::uint32_t>(j.permutation.topology);
        for (int l = 0; l < 3; ++l) {
            o[3 + l] = j.permutation.test[l];
            o[6 + l] = j.permutation.trial[l];
        }
    }
    for (std::size_t i = 0; i < p.near.size(); ++i) {
        const auto &j = p.near[i];
        std::uint32_t *o = near + 5 * i;
        o[0] = j.test, o[1] = j.trial, o[2] = j.firstPatch, o[3] = j.patchCount, o[4] = j.swapBarycentric;
    }
    for (std::size_t i = 0; i < p.patches.size(); ++i)
        pack_patch(p.patches[i], patches + 22 * i);
    if (!p.wholePanelDegrees.empty())
        std::memcpy(whole, p.wholePanelDegrees.data(), p.wholePanelDegrees.size());
    for (std::size_t i = 0; i < p.nearOffsets.size(); ++i)
        offsets[i] = p.nearOffsets[i];
    for (std::size_t i = 0; i < p.nearTrials.size(); ++i)
        trials[i] = p.nearTrials[i];
    *max_error = p.maximumEstimatedRelativeError;
}

void ref_plan_destroy(void *handle) { delete static_cast<PlanHandle *>(handle); }

// makeSupportTree (fmm/support_tree.cpp) on the raw coordinates, serialised.
int ref_support_tree(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                     unsigned leaf_capacity, unsigned maximum_depth, double separation, double near_ratio,
                     double *blob, std::uint64_t *size) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        wideband::SupportTreeSettings settings;
        settings.leafCapacity = leaf_capacity;
        settings.maximumDepth = maximum_depth;
        settings.separationRatio = separation;
        settings.nearDistanceRatio = near_ratio;
        const auto v = tree_blob(wideband::makeSupportTree(s, settings), nv, nt);
        if (blob) {
            if (*size < v.size())
                throw std::runtime_error("support tree buffer too small");
            std::memcpy(blob, v.data(), v.size() * sizeof(double));
        }
        *size = v.size();
    });
}

// makeSupportNearPlan + SupportNearJobIterator (fmm/support_near_plan.cpp) with
// face ids and driven flags, serialised like hsbem_support_near_plan_build.
int ref_support_near_plan(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                          const std::uint32_t *face_ids, const std::uint8_t *driven, const std::uint32_t *faces,
                          std::uint32_t nfaces, unsigned leaf_capacity, unsigned maximum_depth, double separation,
                          double near_ratio, double *blob, std::uint64_t *size) {
    return guarded([&] {
        auto mesh = make_mesh(nv, xyz, nt, tri);
        for (std::uint32_t t = 0; t < nt; ++t) {
            mesh.triangles[t].faceId = face_ids[t];
            mesh.triangles[t].boundary = driven[t] ? BoundaryCondition::driven : BoundaryCondition::wall;
        }
        const auto s = buildSpace(mesh);
        wideband::SupportTreeSettings settings;
        settings.leafCapacity = leaf_capacity;
        settings.maximumDepth = maximum_depth;
        settings.separationRatio = separation;
        settings.nearDistanceRatio = near_ratio;
        const auto tree = wideband::makeSupportTree(s, settings);
        const auto plan = wideband::makeSupportNearPlan(s, tree, std::span<const std::uint32_t>(faces, nfaces),
                                                        std::size_t{8} << 30);
        std::vector<double> b;
        for (double v : {double(plan.pressure.rowOffsets.size()), double(plan.pressure.columns.size()),
                         double(plan.flux.columns.size()), double(plan.nodePanelOffsets.size()),
                         double(plan.supportPanels.size()), double(plan.fluxBlockIndices.size())})
            b.push_back(v);
        const std::size_t job_slot = b.size();
        b.push_back(0);
        for (const auto *v : {&plan.pressure.rowOffsets, &plan.pressure.columns, &plan.flux.rowOffsets,
                              &plan.flux.columns, &plan.panelToActiveFaceColumn, &plan.nodePanelOffsets,
                              &plan.supportPanels, &plan.fluxBlockOffsets, &plan.fluxBlockIndices})
            b.insert(b.end(), v->begin(), v->end());
        wideband::SupportNearJobIterator it(s, tree, plan);
        std::vector<wideband::SupportNearJob> batch(4096);
        std::size_t jobs = 0;
        while (!it.done()) {
            const auto n = it.next(batch, batch.size() * sizeof(wideband::SupportNearJob));
            for (std::size_t i = 0; i < n; ++i)
                for (double v : {double(batch[i].testPanel), double(batch[i].sourcePanel), double(batch[i].pressureMask),
                                 double(batch[i].fluxMask)})
                    b.push_back(v);
            jobs += n;
        }
        b[job_slot] = double(jobs);
        if (blob) {
            if (*size < b.size())
                throw std::runtime_error("near plan buffer too small");
            std::memcpy(blob, b.data(), b.size() * sizeof(double));
        }
        *size = b.size();
    });
}

// makeSupportExpansionPlan (fmm/support_level_plan.cpp, with makeWidebandLevelPlan
// of fmm/wideband_transfer.cpp), serialised like hsbem_support_expansion_plan_build.
int ref_support_expansion_plan(std::uint32_t nv, const double *xyz, std::uint32_t nt, const std::uint32_t *tri,
                               const std::uint32_t *face_ids, const std::uint8_t *driven, const std::uint32_t *faces,
                               std::uint32_t nfaces, unsigned leaf_capacity, unsigned maximum_depth,
                               double separation, double near_ratio, double k, double seed, double *blob,
                               std::uint64_t *size) {
    return guarded([&] {
        const auto s = buildSpace(make_mesh(nv, xyz, nt, tri));
        wideband::SupportTreeSettings settings;
        settings.leafCapacity = leaf_capacity;
        settings.maximumDepth = maximum_depth;
        settings.separationRatio = separation;
        settings.nearDistanceRatio = near_ratio;
        const auto tree = wideband::makeSupportTree(s, settings);
        std::vector<std::uint32_t> column(nt, UINT32_MAX);
        for (std::uint32_t t = 0; t < nt; ++t)
            for (std::uint32_t f = 0; f < nfaces; ++f)
                if (face_ids[t] == faces[f] && driven[t]) {
                    column[t] = f;
                    break;
                }
        const auto plan = wideband::makeSupportExpansionPlan(tree, column, k, seed);
        std::vector<double> b{double(plan.levels.size()), double(plan.activeChannels.size())};
        for (const auto &l : plan.levels) {
            const auto &e = l.expansion;
            for (double v : {double(static_cast<int>(e.representation)), e.activeRadius, e.electricalRadius,
                             e.accuracySeed, double(e.signatureOrder), double(e.translationOrder),
                             double(e.quadratureOrder), double(e.sphericalCoefficients), double(e.angularSamples),
                             l.sourceSeparation, l.targetSeparation, l.policySeparation, l.maximumActiveRadius,
                             l.minimumFarDistance, l.diagonalConditionIndicator})
                b.push_back(v);
        }
        for (auto a : plan.activeChannels)
            b.push_back(a);
        if (blob) {
            if (*size < b.size())
                throw std::runtime_error("expansion plan buffer too small");
            std::memcpy(blob, b.data(), b.size() * sizeof(double));
        }
        *size = b.size();
    });
}

// makeSphericalRotationPlan / rotateToAxis / rotateFromAxis (fmm/spherical_rotation.cpp).
int ref_rotation_plan(unsigned order, const double *axis, double *small_d, double *azimuth, const double *coefficients,
                      int direction, double *rotated) {
    return guarded([&] {
        const auto p = wideband::makeSphericalRotationPlan(order, {axis[0], axis[1], axis[2]});
        std::memcpy(small_d, p.smallD.data(), p.smallD.size() * sizeof(double));
        for (std::size_t i = 0; i < p.azimuthPhases.size(); ++i) {
            azimuth[2 * i] = p.azimuthPhases[i].real();
            azimuth[2 * i + 1] = p.azimuthPhases[i].imag();
        }
        if (coefficients) {
            const std::size_t nc = static_cast<std::size_t>(order + 1) * (order + 1);
            std::vector<wideband::Complex> c(nc);
            for (std::size_t i = 0; i < nc; ++i)
                c[i] = {coefficients[2 * i], coefficients[2 * i + 1]};
            const auto r = direction > 0 ? wideband::rotateToAxis(p, c) : wideband::rotateFromAxis(p, c);
            for (std::size_t i = 0; i < nc; ++i) {
                rotated[2 * i] = r[i].real();
                rotated[2 * i + 1] = r[i].imag();
            }
        }
    });
}

int ref_rcr_queue_policy(unsigned source_order, unsigned target_order, unsigned reserved, std::uint64_t *out) {
    return guarded([&] {
        const auto p = wideband::makeRcrQueuePolicy(source_order, target_order, reserved);
        out[0] = p.batchCapacity;
        out[1] = p.maximumChunkBatches;
        out[2] = p.workPerLink;
    });
}
}

namespace {
std::vector<wideband::Complex> ref_complex_in(const double *v, std::size_t n) {
    std::vector<wideband::Complex> out(n);
    for (std::size_t i = 0; i < n; ++i)
        out[i] = {v[2 * i], v[2 * i + 1]};
    return out;
}
void ref_complex_out(const std::vector<wideband::Complex> &v, double *out) {
    for (std::size_t i = 0; i < v.size(); ++i) {
        out[2 * i] = v[i].real();
        out[2 * i + 1] = v[i].imag();
    }
}
} // namespace

extern "C" {
int ref_sphere_rule(unsigned order, double *directions, double *weights) {
    return guarded([&] {
        const auto r = wideband::makeSphereRule(order);
        for (std::size_t q = 0; q < r.samples.size(); ++q) {
            directions[3 * q] = r.samples[q].direction.x;
            directions[3 * q + 1] = r.samples[q].direction.y;
            directions[3 * q + 2] = r.samples[q].direction.z;
            weights[q] = r.samples[q].weight;
        }
    });
}

int ref_spherical_harmonics(unsigned order, const double *d, double *values) {
    return guarded([&] { ref_complex_out(wideband::sphericalHarmonics(order, {d[0], d[1], d[2]}), values); });
}

int ref_regular_radial_table(unsigned order, double x, double *values, double *derivatives) {
    return guarded([&] {
        const auto t = wideband::regularRadialTable(order, x);
        std::memcpy(values, t.values.data(), t.values.size() * sizeof(double));
        std::memcpy(derivatives, t.derivatives.data(), t.derivatives.size() * sizeof(double));
    });
}

int ref_angular_transfer(unsigned source_order, unsigned target_order, unsigned bandlimit, int mode, double k,
                         const double *shift, const double *values, double *out) {
    return guarded([&] {
        const auto s = wideband::makeSphereRule(source_order), t = wideband::makeSphereRule(target_order);
        const auto v = ref_complex_in(values, s.samples.size());
        const Vec3 d = shift ? Vec3{shift[0], shift[1], shift[2]} : Vec3{};
        if (mode == 0)
            ref_complex_out(wideband::resampleSpectrum(v, s, t, bandlimit), out);
        else if (mode == 1)
            ref_complex_out(wideband::outgoingToParent(v, s, t, bandlimit, k, d), out);
        else
            ref_complex_out(wideband::incomingToChild(v, s, t, bandlimit, k, d), out);
    });
}

int ref_harmonic_spectrum(unsigned order, unsigned rule_order, int direction, const double *in, double *out) {
    return guarded([&] {
        const auto r = wideband::makeSphereRule(rule_order);
        if (direction > 0)
            ref_complex_out(wideband::multipoleToSpectrum(
                                ref_complex_in(in, static_cast<std::size_t>(order + 1) * (order + 1)), order, r),
                            out);
        else
            ref_complex_out(wideband::spectrumToLocal(ref_complex_in(in, r.samples.size()), r, order), out);
    });
}

int ref_diagonal_translation(unsigned rule_order, unsigned expansion_order, double k, const double *r, double *out) {
    return guarded([&] {
        ref_complex_out(
            wideband::diagonalTranslation(wideband::makeSphereRule(rule_order), expansion_order, k, {r[0], r[1], r[2]}),
            out);
    });
}
}

namespace {
void ref_wave_out(const wideband::ScalarWave &w, double *out) {
    out[0] = w.value.real();
    out[1] = w.value.imag();
    for (int i = 0; i < 3; ++i) {
        out[2 + 2 * i] = w.gradient[i].real();
        out[3 + 2 * i] = w.gradient[i].imag();
    }
}
} // namespace

extern "C" {
int ref_scaled_waves(unsigned order, double k, double radius, const double *p, int outgoing, double *waves) {
    return guarded([&] {
        const auto b = outgoing ? wideband::scaledOutgoingWaves(order, k, radius, {p[0], p[1], p[2]})
                                : wideband::scaledRegularWaves(order, k, radius, {p[0], p[1], p[2]});
        for (std::size_t j = 0; j < b.waves.size(); ++j)
            ref_wave_out(b.waves[j], waves + 8 * j);
    });
}

int ref_scaled_expansion_evaluate(unsigned order, double radius, double k, const double *c, int local,
                                  const double *p, double *wave) {
    return guarded([&] {
        const auto coefficients = ref_complex_in(c, static_cast<std::size_t>(order + 1) * (order + 1));
        if (local)
            ref_wave_out(wideband::evaluateScaledLocal({order, radius, coefficients}, k, {p[0], p[1], p[2]}), wave);
        else
            ref_wave_out(wideband::evaluateScaledMultipole({order, radius, coefficients}, k, {p[0], p[1], p[2]}),
                         wave);
    });
}

int ref_scaled_spectrum(unsigned order, double radius, double k, unsigned rule_order, int direction, const double *in,
                        double *out) {
    return guarded([&] {
        const auto rule = wideband::makeSphereRule(rule_order);
        if (direction > 0)
            ref_complex_out(wideband::scaledMultipoleToSpectrum(
                                {order, radius, ref_complex_in(in, static_cast<std::size_t>(order + 1) * (order + 1))},
                                k, rule),
                            out);
        else
            ref_complex_out(
                wideband::spectrumToScaledLocal(ref_complex_in(in, rule.samples.size()), rule, order, k, radius)
                    .coefficients,
                out);
    });
}

int ref_axial_plan(int kind, unsigned ps, unsigned pt, double k, double as, double at, double d, unsigned polar,
                   double projection, std::uint64_t *count, double *coefficients, double *probe, const double *source,
                   double *applied) {
    return guarded([&] {
        const auto plan = wideband::makeScaledAxialPlan(static_cast<wideband::ScaledAxialKind>(kind), ps, pt, k, as,
                                                        at, d, polar, projection);
        *count = plan.coefficients.size();
        if (coefficients)
            ref_complex_out(plan.coefficients, coefficients);
        if (probe)
            *probe = plan.projectionRadius;
        if (source)
            ref_complex_out(wideband::applyScaledAxialPlan(
                                plan, ref_complex_in(source, static_cast<std::size_t>(ps + 1) * (ps + 1))),
                            applied);
    });
}
}

namespace {
std::vector<AcousticsAngularSample> ref_samples(const double *a, const double *l, unsigned n) {
    std::vector<AcousticsAngularSample> s(n);
    for (unsigned i = 0; i < n; ++i)
        s[i] = {a[i], l[i]};
    return s;
}
} // namespace

extern "C" {
int ref_spl_db(double re, double im, double reference, double *out) {
    return guarded([&] { *out = soundPressureLevelDecibels({re, im}, reference); });
}

int ref_normalized_volume_impedance(double re, double im, double area, double rho, double c, double *out) {
    return guarded([&] {
        const auto z = normalizedVolumeImpedance({re, im}, area, rho, c);
        out[0] = z.real();
        out[1] = z.imag();
    });
}

int ref_angular_level_db(const double *a, const double *l, unsigned n, double angle, double *level, int *present) {
    return guarded([&] {
        const auto s = ref_samples(a, l, n);
        const auto v = angularLevelDecibels(s, angle);
        *present = v.has_value();
        *level = v.value_or(0.0);
    });
}

int ref_beamwidth(const double *a, const double *l, unsigned n, double attenuation, double *out, int *present) {
    return guarded([&] {
        const auto s = ref_samples(a, l, n);
        const auto b = beamwidthDegrees(s, attenuation);
        *present = b.has_value();
        out[0] = b ? b->leftDegrees : 0.0;
        out[1] = b ? b->rightDegrees : 0.0;
        out[2] = b ? b->widthDegrees : 0.0;
    });
}

int ref_cta2034_cuts(const double *ha, const double *hl, unsigned hn, const double *va, const double *vl, unsigned vn,
                     double *out, int *present) {
    return guarded([&] {
        const auto h = ref_samples(ha, hl, hn), v = ref_samples(va, vl, vn);
        const auto c = cta2034Cuts(h, v);
        const std::optional<double> x[7] = {c.onAxisDb,    c.listeningWindowDb, c.earlyReflectionsDb,
                                            c.earlyReflectionsDiDb, c.soundPowerDb, c.soundPowerDiDb,
                                            c.predictedInRoomDb};
        for (int i = 0; i < 7; ++i) {
            present[i] = x[i].has_value();
            out[i] = x[i].value_or(0.0);
        }
    });
}

int ref_angular_frequency_response(const double *f, const unsigned *offsets, const double *a, const double *l,
                                   unsigned rows, double angle, int relative, double reference, double *out,
                                   int *present) {
    return guarded([&] {
        std::vector<std::vector<AcousticsAngularSample>> data(rows);
        std::vector<AngularFrequencyRow> r(rows);
        for (unsigned i = 0; i < rows; ++i) {
            data[i] = ref_samples(a + offsets[i], l + offsets[i], offsets[i + 1] - offsets[i]);
            r[i] = {f[i], data[i]};
        }
        const auto v = angularFrequencyResponse(
            r, angle, relative ? AngularResponseNormalization::referenceAngle : AngularResponseNormalization::absolute,
            reference);
        for (unsigned i = 0; i < rows; ++i) {
            present[i] = v[i].levelDecibels.has_value();
            out[i] = v[i].levelDecibels.value_or(0.0);
        }
    });
}

int ref_acoustic_frame(const double *o, const double *f, const double *h, double *out) {
    return guarded([&] {
        const AcousticFrame fr = h ? normalizedAcousticFrame({{o[0], o[1], o[2]}, {f[0], f[1], f[2]}, {h[0], h[1], h[2]}})
                                   : defaultAcousticFrame({o[0], o[1], o[2]}, {f[0], f[1], f[2]});
        const Vec3 v = verticalDirection(fr);
        const double w[12] = {fr.origin.x,     fr.origin.y,     fr.origin.z,     fr.forward.x,
                              fr.forward.y,    fr.forward.z,    fr.horizontal.x, fr.horizontal.y,
                              fr.horizontal.z, v.x,             v.y,             v.z};
        std::memcpy(out, w, sizeof(w));
    });
}
}
