// Solve orchestration. Reference behaviour restated (cuda_solver.cu):
//   recentring on vertex 0 and observer transform ........ :771-781, :1107-1137
//   request validation ................................... :1054-1080, :1165-1193
//   velocities per panel / active face columns .......... :1165-1212
//   per-frequency q = -i rho omega v, assembly, solve .... :1292-1518
//   FP64 residual relative to the mass-weighted rhs ...... :555-585, :1472-1481, :1519-1581
//   observer/far-field endpoint-degree convergence loop .. :1717-1799
//   surface outputs and port impedances ................. :1800-1846
#include "solver.hpp"

#include "autotune.hpp"
#include "host/mesh_check.hpp"
#include "host/qualification.hpp"

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <numbers>
#include <numeric>
#include <set>
#include <sstream>

namespace hsbem {

void hip_check(hipError_t status, const char *what) {
    if (status != hipSuccess)
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(status));
}

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

std::uint64_t spread21(std::uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

void check_error_flag(const DeviceBuffer<int> &flag, hipStream_t s) {
    int code = 0;
    hip_check(hipMemcpyAsync(&code, flag.data(), sizeof(int), hipMemcpyDeviceToHost, s), "error flag download");
    hip_check(hipStreamSynchronize(s), "stream synchronize");
    if (code == dev::ERR_DISTANCE)
        throw std::runtime_error("a quadrature sample landed on (or too close to) its source point");
    if (code == dev::ERR_OVERFLOW)
        throw std::runtime_error("BEM calculation produced a non-finite value.");
    if (code == dev::ERR_FIELD)
        throw std::runtime_error("observer field: near-field refinement did not reach its tolerance");
    if (code == dev::ERR_ESCAPED)
        throw std::runtime_error("a panel pair needs a rule the interaction plan did not qualify");
    if (code == dev::ERR_RANGE)
        throw std::runtime_error("a far-tile entry exceeded its fixed-point bound (internal error)");
}

bool point_inside(const std::vector<Panel> &panels, Vec3 p) {
    // Solid-angle sum (cuda_solver.cu:679-704): inside if |Omega| > 2 pi; points on a panel count as inside.
    double omega = 0;
    for (const auto &pn : panels) {
        const Vec3 e0 = vsub(pn.vertices[1], pn.vertices[0]), e1 = vsub(pn.vertices[2], pn.vertices[0]);
        const Vec3 rel = vsub(p, pn.vertices[0]);
        const double scale = std::max(std::sqrt(vnorm2(p)), std::sqrt(vnorm2(pn.vertices[0])));
        const double tol = std::max(pn.max_edge * 1e-10, scale * 32 * std::numeric_limits<double>::epsilon());
        if (std::abs(vdot(rel, pn.normal)) <= tol) {
            const double aa = vdot(e0, e0), ab = vdot(e0, e1), bb = vdot(e1, e1), pa = vdot(rel, e0),
                         pb = vdot(rel, e1), det = aa * bb - ab * ab;
            const double u = (bb * pa - ab * pb) / det, v = (aa * pb - ab * pa) / det;
            if (u >= -1e-10 && v >= -1e-10 && u + v <= 1 + 1e-10)
                return true;
        }
        const Vec3 a = vsub(pn.vertices[0], p), b = vsub(pn.vertices[1], p), c = vsub(pn.vertices[2], p);
        const double la = std::sqrt(vnorm2(a)), lb = std::sqrt(vnorm2(b)), lc = std::sqrt(vnorm2(c));
        const double den = la * lb * lc + vdot(a, b) * lc + vdot(b, c) * la + vdot(c, a) * lb;
        omega += 2 * std::atan2(vdot(a, vcross(b, c)), den);
    }
    return std::abs(omega) > 2 * std::numbers::pi;
}

std::vector<double> rule_table() {
    std::vector<double> t;
    for (unsigned d : {4u, 6u, 8u, 14u})
        for (const auto &p : triangle_rule(d)) {
            t.push_back(p.bary[0]);
            t.push_back(p.bary[1]);
            t.push_back(p.bary[2]);
            t.push_back(p.weight);
        }
    return t;
}

} // namespace

MeshSolver::MeshSolver(Context *ctx, SurfaceMesh mesh) : ctx_(ctx) {
    if (mesh.vertices.empty())
        throw std::invalid_argument("the mesh has no vertices");
    origin_ = mesh.vertices.front();
    for (auto &v : mesh.vertices) {
        v = vsub(v, origin_);
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z))
            throw std::invalid_argument("mesh vertex coordinates must be finite");
    }
    validate_closed_exterior(mesh);
    space_ = build_space(mesh);
    mesh_ = std::move(mesh);
    // FNV-1a over the recentred coordinates and the triangles: the mesh's key for the
    // launch-shape autotuner
    if (const char *e = std::getenv("HSBEM_PLAN_ESTIMATOR"))
        device_estimates_ = std::string(e) != "host";
    hash_ = 14695981039346656037ull;
    auto mix = [this](const void *p, std::size_t n) {
        const auto *b = static_cast<const unsigned char *>(p);
        for (std::size_t i = 0; i < n; ++i)
            hash_ = (hash_ ^ b[i]) * 1099511628211ull;
    };
    for (const auto &v : mesh_.vertices) {
        const double c[3] = {v.x, v.y, v.z};
        mix(c, sizeof(c));
    }
    for (const auto &t : mesh_.triangles) {
        mix(t.vertices.data(), sizeof(std::uint32_t) * 3);
        mix(&t.face_id, sizeof(t.face_id));
    }
    build_clusters();
    rules_.upload(rule_table(), ctx_->stream);
    {
        std::vector<double> table;
        for (const unsigned degree : {4u, 6u, 8u, 12u, 14u}) {
            const auto rule = triangle_rule(degree);
            endpoint_rule_slots_[degree] = {static_cast<std::uint32_t>(table.size() / 4),
                                            static_cast<std::uint32_t>(rule.size())};
            for (const auto &p : rule)
                table.insert(table.end(), {p.bary[0], p.bary[1], p.bary[2], p.weight});
        }
        endpoint_rules_.upload(table, ctx_->stream);
        unit_.upload(std::vector<double>{1.0, 0.0}, ctx_->stream);
    }
    hip_check(hipStreamSynchronize(ctx_->stream), "mesh upload");
    geometry_bytes_ = geo_.capacity_bytes() + bounds_.capacity_bytes() + row_scale_.capacity_bytes() +
                      lumped_.capacity_bytes() + rules_.capacity_bytes() + nodes_.capacity_bytes() +
                      shape_.capacity_bytes() + local_.capacity_bytes() + cluster_panel_offsets_.capacity_bytes() +
                      cluster_panels_.capacity_bytes() + cluster_node_offsets_.capacity_bytes() +
                      cluster_nodes_.capacity_bytes() + cluster_far_.capacity_bytes() +
                      endpoint_rules_.capacity_bytes() + unit_.capacity_bytes();
    workspace_ = std::make_unique<Workspace>(ctx_->stream);
}

Workspace::Workspace(hipStream_t s, bool owns) : stream(s), owns_stream(owns) {
    error.resize(1);
    samples.resize(1);
    for (auto &e : events)
        hip_check(hipEventCreate(&e), "event create");
}

Workspace::~Workspace() {
    for (auto &e : events)
        if (e)
            (void)hipEventDestroy(e);
    if (owns_stream && stream) {
        (void)hipStreamSynchronize(stream);
        release();
        (void)hipStreamDestroy(stream);
    }
}

void Workspace::release() {
    for (auto *b : {&A, &LU, &BD, &q, &rhs, &x, &r, &flux, &scratch, &obs, &field, &partials, &fmm_b, &fmm_x, &fmm_r,
                     &fmm_q})
        b->release();
    flux_column.release();
    pivots.release();
    jacobi.release();
    jacobi_blocks.release();
    jacobi_pivots.release();
    gmres.release();
    bound = 0;
}

std::uint64_t Workspace::capacity_bytes() const {
    std::uint64_t total = 0;
    for (const auto *b : {&A, &LU, &BD, &q, &rhs, &x, &r, &flux, &scratch, &obs, &field, &partials, &fmm_b, &fmm_x,
                           &fmm_r, &fmm_q})
        total += b->capacity_bytes();
    return total + flux_column.capacity_bytes() + pivots.capacity_bytes() + jacobi.capacity_bytes() +
           jacobi_blocks.capacity_bytes() + jacobi_pivots.capacity_bytes() + gmres.z.capacity_bytes() + gmres.basis.capacity_bytes() +
           gmres.r.capacity_bytes() + gmres.scratch.capacity_bytes() + gmres.small.capacity_bytes() +
           gmres.multi_basis.capacity_bytes() + gmres.multi_in.capacity_bytes() + gmres.multi_out.capacity_bytes() +
           gmres.multi_small.capacity_bytes();
}

void MeshSolver::build_clusters() {
    const std::uint32_t nt = this->nt(), nv = this->nv();
    Vec3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (const auto &p : space_.panels) {
        lo = {std::min(lo.x, p.centroid.x), std::min(lo.y, p.centroid.y), std::min(lo.z, p.centroid.z)};
        hi = {std::max(hi.x, p.centroid.x), std::max(hi.y, p.centroid.y), std::max(hi.z, p.centroid.z)};
    }
    const double ext = std::max({hi.x - lo.x, hi.y - lo.y, hi.z - lo.z, 1e-300});
    std::vector<std::pair<std::uint64_t, std::uint32_t>> keyed(nt);
    for (std::uint32_t t = 0; t < nt; ++t) {
        const Vec3 c = space_.panels[t].centroid;
        auto q = [&](double v, double l) {
            return static_cast<std::uint64_t>(std::min(2097151.0, std::max(0.0, (v - l) / ext * 2097151.0)));
        };
        keyed[t] = {spread21(q(c.x, lo.x)) | spread21(q(c.y, lo.y)) << 1 | spread21(q(c.z, lo.z)) << 2, t};
    }
    std::sort(keyed.begin(), keyed.end());
    constexpr std::size_t kPanels = 64, kNodes = 64;
    std::vector<std::uint32_t> panel_offsets{0}, panels, node_offsets{0}, cluster_vertices;
    std::vector<std::uint32_t> cur_nodes;
    std::vector<std::uint32_t> cluster_of(nt);
    auto flush = [&] {
        cluster_vertices.insert(cluster_vertices.end(), cur_nodes.begin(), cur_nodes.end());
        node_offsets.push_back(static_cast<std::uint32_t>(cluster_vertices.size()));
        panel_offsets.push_back(static_cast<std::uint32_t>(panels.size()));
        cur_nodes.clear();
    };
    for (const auto &[key, t] : keyed) {
        std::size_t fresh = 0;
        for (auto v : space_.panels[t].nodes)
            fresh += std::find(cur_nodes.begin(), cur_nodes.end(), v) == cur_nodes.end();
        if (panels.size() - panel_offsets.back() == kPanels || cur_nodes.size() + fresh > kNodes)
            flush();
        for (auto v : space_.panels[t].nodes)
            if (std::find(cur_nodes.begin(), cur_nodes.end(), v) == cur_nodes.end())
                cur_nodes.push_back(v);
        cluster_of[t] = static_cast<std::uint32_t>(panel_offsets.size() - 1);
        panels.push_back(t);
    }
    flush();
    cluster_count_ = static_cast<std::uint32_t>(panel_offsets.size() - 1);
    // internal numbering: first appearance in cluster order
    internal_of_.assign(nv, UINT32_MAX);
    vertex_of_.clear();
    for (auto v : cluster_vertices)
        if (internal_of_[v] == UINT32_MAX) {
            internal_of_[v] = static_cast<std::uint32_t>(vertex_of_.size());
            vertex_of_.push_back(v);
        }
    std::vector<std::uint32_t> cluster_nodes(cluster_vertices.size());
    for (std::size_t i = 0; i < cluster_vertices.size(); ++i)
        cluster_nodes[i] = internal_of_[cluster_vertices[i]];
    std::vector<std::uint8_t> local(3 * static_cast<std::size_t>(nt));
    std::vector<std::uint32_t> nodes(3 * static_cast<std::size_t>(nt));
    std::vector<double> geo(dev::G_COUNT * static_cast<std::size_t>(nt)), bounds(dev::B_COUNT * static_cast<std::size_t>(nt));
    std::vector<std::uint8_t> shape(nt);
    for (std::uint32_t t = 0; t < nt; ++t) {
        const Panel &p = space_.panels[t];
        const std::uint32_t c = cluster_of[t];
        for (int l = 0; l < 3; ++l) {
            nodes[static_cast<std::size_t>(l) * nt + t] = internal_of_[p.nodes[l]];
            const auto first = cluster_vertices.begin() + node_offsets[c], last = cluster_vertices.begin() + node_offsets[c + 1];
            local[static_cast<std::size_t>(l) * nt + t] = static_cast<std::uint8_t>(std::find(first, last, p.nodes[l]) - first);
        }
        const Vec3 e1 = vsub(p.vertices[1], p.vertices[0]), e2 = vsub(p.vertices[2], p.vertices[0]);
        const double g[dev::G_COUNT] = {p.vertices[0].x, p.vertices[0].y, p.vertices[0].z, e1.x, e1.y, e1.z, e2.x, e2.y,
                                        e2.z, p.normal.x, p.normal.y, p.normal.z, p.curls[0].x, p.curls[0].y,
                                        p.curls[0].z, p.curls[1].x, p.curls[1].y, p.curls[1].z, p.curls[2].x,
                                        p.curls[2].y, p.curls[2].z, p.area};
        for (int f = 0; f < dev::G_COUNT; ++f)
            geo[static_cast<std::size_t>(f) * nt + t] = g[f];
        const PanelBounds b = make_panel_bounds(p);
        const double bb[dev::B_COUNT] = {b.center.x, b.center.y, b.center.z, b.radius, b.max_edge};
        for (int f = 0; f < dev::B_COUNT; ++f)
            bounds[static_cast<std::size_t>(f) * nt + t] = bb[f];
        shape[t] = static_cast<std::uint8_t>(b.qualified_shape);
    }
    lumped_internal_.resize(nv);
    std::vector<double> row_scale(nv);
    for (std::uint32_t i = 0; i < nv; ++i) {
        lumped_internal_[i] = space_.lumped_mass[vertex_of_[i]];
        row_scale[i] = 1.0 / lumped_internal_[i];
    }
    hipStream_t s = ctx_->stream;
    geo_.upload(geo, s);
    bounds_.upload(bounds, s);
    shape_.upload(shape, s);
    nodes_.upload(nodes, s);
    local_.upload(local, s);
    row_scale_.upload(row_scale, s);
    lumped_.upload(lumped_internal_, s);
    // cluster bounds for the far-tile test
    cluster_bounds_.assign(cluster_count_, {});
    for (std::uint32_t c = 0; c < cluster_count_; ++c) {
        ClusterBounds &cb = cluster_bounds_[c];
        Vec3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
        for (std::uint32_t q = panel_offsets[c]; q < panel_offsets[c + 1]; ++q) {
            const Vec3 p = space_.panels[panels[q]].centroid;
            lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
            hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
        }
        cb.center = {(lo.x + hi.x) / 2, (lo.y + hi.y) / 2, (lo.z + hi.z) / 2};
        cb.shapes_ok = true;
        for (std::uint32_t q = panel_offsets[c]; q < panel_offsets[c + 1]; ++q) {
            const std::uint32_t t = panels[q];
            cb.spread = std::max(cb.spread, std::sqrt(vnorm2(vsub(space_.panels[t].centroid, cb.center))));
            cb.ball = std::max(cb.ball, bounds[static_cast<std::size_t>(dev::B_R) * nt + t]);
            cb.max_edge = std::max(cb.max_edge, space_.panels[t].max_edge);
            cb.shapes_ok = cb.shapes_ok && shape[t];
        }
    }
    // Far-tile fixed-point bounds per cluster (dev::kClusterFar doubles): the centre,
    // the reach (spread + ball: every point of every panel lies within it), the
    // largest panel area, the largest |curl phi_i| and the largest number of the
    // cluster's panels sharing one of its nodes.
    std::vector<double> far(static_cast<std::size_t>(dev::kClusterFar) * cluster_count_, 0.0);
    for (std::uint32_t c = 0; c < cluster_count_; ++c) {
        const ClusterBounds &cb = cluster_bounds_[c];
        double *f = far.data() + static_cast<std::size_t>(dev::kClusterFar) * c;
        f[0] = cb.center.x;
        f[1] = cb.center.y;
        f[2] = cb.center.z;
        f[3] = cb.spread + cb.ball;
        std::vector<std::uint32_t> share(node_offsets[c + 1] - node_offsets[c], 0);
        for (std::uint32_t q = panel_offsets[c]; q < panel_offsets[c + 1]; ++q) {
            const Panel &pn = space_.panels[panels[q]];
            f[4] = std::max(f[4], pn.area);
            for (int l = 0; l < 3; ++l) {
                f[5] = std::max(f[5], std::sqrt(vnorm2(pn.curls[l])));
                ++share[local[static_cast<std::size_t>(l) * nt + panels[q]]];
            }
        }
        for (const auto v : share)
            f[6] = std::max(f[6], static_cast<double>(v));
    }
    cluster_far_.upload(far, s);
    cluster_panel_offsets_.upload(panel_offsets, s);
    cluster_panels_.upload(panels, s);
    cluster_node_offsets_.upload(node_offsets, s);
    cluster_nodes_.upload(cluster_nodes, s);
}

InteractionPlan MeshSolver::plan(const Numerics &n, double k_max, hipStream_t s, bool device) {
    QuadratureSettings q;
    q.regular_order = n.quadrature_order;
    q.singular_order = n.singular_order;
    q.near_distance_ratio = n.near_ratio;
    q.near_tolerance = n.near_tolerance;
    q.near_max_depth = n.near_depth;
    if (!device)
        return make_interaction_plan(space_, q, k_max, n.plan_threads);
    if (!s)
        s = ctx_->stream;
    {
        std::lock_guard lock(estimator_mutex_);
        if (!estimator_panels_.size()) {
            std::vector<dev::EstimatorPanel> panels(space_.panels.size());
            for (std::size_t t = 0; t < panels.size(); ++t) {
                const Panel &p = space_.panels[t];
                panels[t].fan = {{p.vertices[0], p.vertices[1], p.vertices[2]}, p.normal, p.area};
                panels[t].max_edge = p.max_edge;
            }
            std::vector<TrianglePoint> rules;
            for (const unsigned degree : {6u, 8u, 12u, 14u}) {
                const auto r = triangle_rule(degree);
                rules.insert(rules.end(), r.begin(), r.end());
            }
            if (rules.size() != 12 + 16 + 49 + 64)
                throw std::logic_error("patch estimator: unexpected triangle rule sizes");
            estimator_panels_.upload(panels, s);
            estimator_rules_.upload(rules, s);
            hip_check(hipStreamSynchronize(s), "patch estimator upload");
        }
    }
    // Device buffers of one call (tasks by kind, results); requests split into tensor
    // and projection launches, each result put back at its request's index.
    struct Buffers {
        DeviceBuffer<dev::EstimatorTask> tasks;
        DeviceBuffer<dev::EstimatorResult> out;
    };
    auto buffers = std::make_shared<Buffers>();
    const PatchEstimator estimator = [this, k_max, s, buffers](const std::vector<PatchRequest> &requests,
                                                               std::vector<PatchEstimate> &estimates) {
        const std::size_t n = requests.size();
        if (n > std::numeric_limits<std::uint32_t>::max())
            throw std::runtime_error("patch estimator: too many requests in one batch");
        std::vector<dev::EstimatorTask> tasks(n);
        std::vector<std::size_t> where(n);
        std::size_t tensor = 0;
        for (std::size_t i = 0; i < n; ++i)
            tensor += requests[i].patch.projection ? 0 : 1;
        for (std::size_t i = 0, a = 0, b = tensor; i < n; ++i) {
            const std::size_t at = requests[i].patch.projection ? b++ : a++;
            tasks[at] = {requests[i].test, requests[i].trial, requests[i].patch};
            where[i] = at;
        }
        buffers->tasks.upload(tasks, s);
        buffers->out.resize(n);
        dev::EstimatorLaunch L{estimator_panels_.data(), buffers->tasks.data(), static_cast<std::uint32_t>(tensor),
                               estimator_rules_.data(), k_max, buffers->out.data()};
        dev::launch_estimate_patches(L, false, s);
        L.tasks += tensor;
        L.out += tensor;
        L.count = static_cast<std::uint32_t>(n - tensor);
        dev::launch_estimate_patches(L, true, s);
        hip_check(hipGetLastError(), "patch estimator launch");
        std::vector<dev::EstimatorResult> out(n);
        hip_check(hipMemcpyAsync(out.data(), buffers->out.data(), n * sizeof(dev::EstimatorResult),
                                 hipMemcpyDeviceToHost, s),
                  "patch estimates download");
        hip_check(hipStreamSynchronize(s), "patch estimates");
        estimates.resize(n);
        for (std::size_t i = 0; i < n; ++i) {
            const dev::EstimatorResult &r = out[where[i]];
            if (r.status == 1)
                throw PatchEstimateError(i, "near pair: the two panels overlap or repeat a vertex position");
            if (r.status != 0)
                throw PatchEstimateError(i, "near pair: a quadrature moment is not finite");
            PatchEstimate &e = estimates[i];
            for (unsigned m = 0; m < 36; ++m)
                e.integral[m] = {r.integral[2 * m], r.integral[2 * m + 1]};
            e.outer_error = r.outer_error;
            e.inner_error = r.inner_error;
            e.absolute_error = r.outer_error + r.inner_error;
        }
    };
    return make_interaction_plan(space_, q, k_max, n.plan_threads, &estimator);
}

std::shared_ptr<DevicePlan> MeshSolver::device_plan(const Numerics &n, double k_max, hipStream_t s, bool *built,
                                                    bool device) {
    const auto key = std::make_tuple(n.quadrature_order, n.singular_order, n.near_ratio, n.near_tolerance, n.near_depth,
                                     k_max, device);
    // Held while building: other streams needing the same mesh's plan wait for it.
    std::lock_guard lock(plan_mutex_);
    {
        std::lock_guard index(plan_index_mutex_);
        auto it = plans_.find(key);
        if (it != plans_.end()) {
            it->second->last_use = ++plan_clock_;
            if (built)
                *built = false;
            return it->second;
        }
    }
    const auto t0 = Clock::now();
    auto dp = std::make_shared<DevicePlan>();
    dp->plan = plan(n, k_max, s, device);
    dp->touching.upload(dp->plan.touching, s);
    // Near jobs: tensor patches go to the per-patch kernel (degree 8 first, then
    // 14, for uniform trip counts); jobs holding a projection fan stay whole.
    std::vector<NearJob> fan_jobs;
    std::vector<dev::NearPatchItem> items8, items14;
    for (const auto &j : dp->plan.near) {
        bool fan = false;
        for (std::uint32_t q = 0; q < j.patch_count; ++q)
            fan |= dp->plan.patches[j.first_patch + q].projection != 0;
        if (fan) {
            fan_jobs.push_back(j);
            continue;
        }
        for (std::uint32_t q = 0; q < j.patch_count; ++q)
            (dp->plan.patches[j.first_patch + q].degree == 14 ? items14 : items8)
                .push_back({j.test, j.trial, j.first_patch + q, j.swap});
    }
    items8.insert(items8.end(), items14.begin(), items14.end());
    dp->near_job_count = static_cast<std::uint32_t>(fan_jobs.size());
    dp->near_item_count = static_cast<std::uint32_t>(items8.size());
    if (fan_jobs.empty())
        fan_jobs.push_back({});
    if (items8.empty())
        items8.push_back({});
    dp->near.upload(fan_jobs, s);
    dp->near_items.upload(items8, s);
    dp->patches.upload(dp->plan.patches, s);
    dp->near_offsets.upload(dp->plan.near_offsets, s);
    std::vector<std::uint32_t> trials = dp->plan.near_trials;
    if (trials.empty())
        trials.push_back(0);
    dp->near_trials.upload(trials, s);
    if (!dp->plan.whole_panel_degrees.empty())
        dp->whole.upload(dp->plan.whole_panel_degrees, s);
    std::vector<double> g;
    for (const auto &p : gauss_legendre01(n.singular_order)) {
        g.push_back(p.position);
        g.push_back(p.weight);
    }
    dp->gauss.upload(g, s);
    for (const auto &j : dp->plan.touching)
        dp->touching_samples += static_cast<double>(sector_count(j.permutation.topology)) * n.singular_order *
                                n.singular_order * n.singular_order * n.singular_order;
    for (const auto &j : dp->plan.near)
        for (std::uint32_t q = 0; q < j.patch_count; ++q) {
            const auto &patch = dp->plan.patches[j.first_patch + q];
            dp->near_samples += patch_point_count(patch, patch.degree == 14 ? 64 : 16);
        }
    // Far tiles: cluster-level bounds guarantee, with a relative margin far above
    // rounding, that every pair of the tile is disjoint, not in the near list
    // (ball gap above max(3, ratio) h and degree > 0 at k_max through the
    // "k h <= 2, one-edge separation" branch) and qualifies for degree 4 at any
    // solve k with k * max_edge <= 0.5 (the "k h <= .5, 3-edge separation" branch).
    {
        const std::uint32_t nc = cluster_count_;
        std::vector<std::pair<double, std::uint32_t>> far;  // (tile edge, tile index)
        std::vector<std::uint32_t> general;
        const double budget = .2 * n.near_tolerance, reserve = 1e-6;
        const bool budgets = reserve + 4 * 2.3e-6 <= budget && reserve + 4 * 6e-7 <= budget;
        // unordered cluster pairs a <= b: the kernels form both orientations of each
        if (!dp->plan.all_pairs && budgets && n.quadrature_order <= 4) {
            const double gap_edges = std::max(3.0, n.near_ratio);
            for (std::uint32_t a = 0; a < nc; ++a)
                for (std::uint32_t b = a; b < nc; ++b) {
                    const ClusterBounds &ca = cluster_bounds_[a], &cb = cluster_bounds_[b];
                    const double h = std::max(ca.max_edge, cb.max_edge);
                    const double d = std::sqrt(vnorm2(vsub(ca.center, cb.center)));
                    const double gap = d - ca.spread - cb.spread;
                    if (a != b && ca.shapes_ok && cb.shapes_ok &&
                        gap >= (ca.ball + cb.ball + gap_edges * h) * (1 + 1e-9) && k_max * h <= 2 * (1 - 1e-12))
                        far.push_back({h, a * nc + b});
                    else
                        general.push_back(a * nc + b);
                }
        } else {
            for (std::uint32_t a = 0; a < nc; ++a)
                for (std::uint32_t b = a; b < nc; ++b)
                    general.push_back(a * nc + b);
        }
        std::stable_sort(far.begin(), far.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        std::vector<std::uint32_t> tiles;
        tiles.reserve(static_cast<std::size_t>(nc) * (nc + 1));
        for (const auto &[h, t] : far) {
            dp->far_tile_edge.push_back(h);
            tiles.push_back(t / nc);
            tiles.push_back(t % nc);
        }
        for (std::uint32_t t : general) {
            tiles.push_back(t / nc);
            tiles.push_back(t % nc);
        }
        dp->tile_count = nc * (nc + 1) / 2;
        if (tiles.empty())
            tiles.push_back(0);
        dp->tiles.upload(tiles, s);
    }
    hip_check(hipStreamSynchronize(s), "plan upload");
    dp->plan_ms = ms_since(t0);
    if (built)
        *built = true;
    {
        std::lock_guard index(plan_index_mutex_);
        dp->last_use = ++plan_clock_;
        plans_.emplace(key, dp);
        plan_bytes_ += dp->device_bytes();
    }
    trim_plans(kPlanCacheEntries);
    return dp;
}

std::uint64_t MeshSolver::trim_plans(std::size_t keep) {
    std::lock_guard index(plan_index_mutex_);
    std::uint64_t released = 0;
    while (plans_.size() > keep) {
        auto victim = plans_.end();
        for (auto it = plans_.begin(); it != plans_.end(); ++it)
            if (it->second.use_count() == 1 && (victim == plans_.end() || it->second->last_use < victim->second->last_use))
                victim = it;
        if (victim == plans_.end())
            break; // every remaining plan is held by a solve or the near structure
        const std::uint64_t bytes = victim->second->device_bytes();
        plans_.erase(victim); // the last reference: device buffers are freed here
        plan_bytes_ -= bytes;
        released += bytes;
    }
    return released;
}

std::size_t MeshSolver::cached_plans() {
    std::lock_guard index(plan_index_mutex_);
    return plans_.size();
}

std::uint64_t MeshSolver::resident_bytes() { return geometry_bytes_ + plan_bytes_.load() + near_bytes_.load(); }

std::uint64_t MeshSolver::pending_plan_bytes(const PreparedRequest &P) {
    const Numerics &n = P.request.numerics;
    const auto key = std::make_tuple(n.quadrature_order, n.singular_order, n.near_ratio, n.near_tolerance, n.near_depth,
                                     P.k_max, n.acceleration == 1 && device_estimates_);
    std::lock_guard index(plan_index_mutex_);
    return plans_.count(key) && (n.acceleration == 0 || near_bytes_.load() > 0) ? 0 : P.plan_estimate;
}

namespace {
std::tuple<unsigned, double, unsigned, unsigned, unsigned> fmm_bytes_key(const Numerics &n) {
    return {n.fmm_leaf_capacity, n.fmm_separation, n.fmm_block_width, n.fmm_restart, n.quadrature_order};
}
} // namespace

std::uint64_t MeshSolver::fmm_point_bytes(const Numerics &n) {
    std::lock_guard index(plan_index_mutex_);
    const auto it = fmm_point_bytes_.find(fmm_bytes_key(n));
    return it == fmm_point_bytes_.end() ? 0 : it->second;
}

void MeshSolver::note_fmm_point_bytes(const Numerics &n, std::uint64_t bytes) {
    std::lock_guard index(plan_index_mutex_);
    auto &v = fmm_point_bytes_[fmm_bytes_key(n)];
    v = std::max(v, bytes);
}

dev::AssemblyLaunch MeshSolver::assembly_launch(const DevicePlan &p, const Numerics &n, double k, double *A, double *BD,
                                                std::uint32_t nd, const std::int32_t *flux_column, int *error) {
    dev::AssemblyLaunch L{};
    L.args.panels = {geo_.data(), nodes_.data(), bounds_.data(), shape_.data(), nt(), nv()};
    L.args.row_scale = row_scale_.data();
    L.args.flux_column = flux_column;
    if (n.precision == 1) {
        L.args.A = nullptr;
        L.args.A32 = reinterpret_cast<float *>(A);
    } else {
        L.args.A = A;
        L.args.A32 = nullptr;
    }
    L.args.BD = BD;
    L.args.nd = nd;
    L.args.k = k;
    L.args.tolerance = n.near_tolerance;
    L.args.regular_order = n.quadrature_order;
    L.args.error = error;
    L.clusters = {cluster_panel_offsets_.data(), cluster_panels_.data(), cluster_node_offsets_.data(),
                  cluster_nodes_.data(), local_.data(), cluster_far_.data(), cluster_count_};
    // far tiles on the uniform degree-4 path at this k: the prefix with k * edge <= 0.5
    // (the same product and comparison as the qualification of each pair)
    L.tiles = p.tiles.data();
    L.tile_count = p.tile_count;
    L.far_tiles = static_cast<std::uint32_t>(
        std::upper_bound(p.far_tile_edge.begin(), p.far_tile_edge.end(), 0.0,
                         [k](double, double h) { return !(k * h <= .5); }) -
        p.far_tile_edge.begin());
    L.near = {p.near_offsets.data(), p.near_trials.data(), p.whole.size() ? p.whole.data() : nullptr};
    L.rules = rules_.data();
    L.touching = p.touching.data();
    L.touching_count = static_cast<std::uint32_t>(p.plan.touching.size());
    L.gauss = p.gauss.data();
    L.singular_order = n.singular_order;
    L.near_jobs = p.near.data();
    L.near_count = p.near_job_count;
    L.near_items = p.near_items.data();
    L.near_item_count = p.near_item_count;
    L.patches = p.patches.data();
    return L;
}

// Trial block of the regular tiles (4, 8 or 16 trial panels per wavefront block),
// chosen per (mesh, precision, far or general tiles) by the autotuner on a prefix of
// that tile list (max(512, count / 32) tiles, all when fewer): one warm and one timed
// launch per candidate. The launches add into A, B, the error flag and the sample
// counter, which the caller clears afterwards. HSBEM_TILE_BLOCK = 4 / 8 / 16 pins it.
void MeshSolver::tune_regular_tiles(dev::AssemblyLaunch &L, bool mixed, hipStream_t s) const {
    const auto pinned = autotune::pinned_from_env("HSBEM_TILE_BLOCK", {4, 8, 16});
    std::ostringstream key;
    key << "mesh " << std::hex << hash_ << std::dec << " (" << nt() << " panels) " << (mixed ? "fp32" : "fp64");
    auto tune = [&](bool far) {
        const std::uint32_t count = far ? L.far_tiles : L.tile_count - L.far_tiles;
        if (!count)
            return 8u;
        dev::AssemblyLaunch T = L;
        const std::uint32_t prefix = std::min(count, std::max(512u, count / 32));
        if (far) {
            T.tile_count = T.far_tiles = prefix;
        } else {
            T.tiles = L.tiles + 2 * static_cast<std::size_t>(L.far_tiles);
            T.tile_count = prefix;
            T.far_tiles = 0;
        }
        return autotune::choose(
            "tile_trial_block", key.str() + (far ? " far" : " general"), {4, 8, 16}, 8, s,
            [&](std::uint32_t w) {
                T.trial_block = T.far_trial_block = w;
                dev::launch_regular(T, s);
            },
            1, pinned);
    };
    L.far_trial_block = tune(true);
    L.trial_block = tune(false);
}

void MeshSolver::assemble_dense(const Numerics &n, double k_max, double k, double *A_out, double *B_out) {
    std::lock_guard lock(mutex_);
    hipStream_t s = ctx_->stream;
    const std::shared_ptr<DevicePlan> plan = device_plan(n, k_max, s, nullptr);
    const DevicePlan &p = *plan;
    const std::uint32_t nv = this->nv(), nt = this->nt();
    DeviceBuffer<int> error(1);
    const bool fp32 = n.precision == 1;
    DeviceBuffer<double> A((fp32 ? 1 : 2) * static_cast<std::size_t>(nv) * nv), BD(2 * static_cast<std::size_t>(nv) * nt);
    std::vector<std::int32_t> cols(nt);
    std::iota(cols.begin(), cols.end(), 0);
    DeviceBuffer<std::int32_t> fc;
    fc.upload(cols, s);
    auto L = assembly_launch(p, n, k, A.data(), BD.data(), nt, fc.data(), error.data());
    tune_regular_tiles(L, fp32, s); // its launches add into A, B and the error flag: cleared next
    hip_check(hipMemsetAsync(A.data(), 0, A.bytes(), s), "clear A");
    hip_check(hipMemsetAsync(BD.data(), 0, BD.bytes(), s), "clear B");
    hip_check(hipMemsetAsync(error.data(), 0, sizeof(int), s), "clear error");
    dev::launch_assembly(L, s);
    hip_check(hipGetLastError(), "assembly launch");
    check_error_flag(error, s);
    std::vector<double> a(2 * static_cast<std::size_t>(nv) * nv), b(BD.size());
    if (fp32) {
        std::vector<float> a32(2 * static_cast<std::size_t>(nv) * nv);
        hip_check(hipMemcpyAsync(a32.data(), A.data(), A.bytes(), hipMemcpyDeviceToHost, s), "A download");
        hip_check(hipStreamSynchronize(s), "A sync");
        std::copy(a32.begin(), a32.end(), a.begin());
    } else {
        hip_check(hipMemcpyAsync(a.data(), A.data(), A.bytes(), hipMemcpyDeviceToHost, s), "A download");
    }
    hip_check(hipMemcpyAsync(b.data(), BD.data(), BD.bytes(), hipMemcpyDeviceToHost, s), "B download");
    hip_check(hipStreamSynchronize(s), "assembly download");
    // internal node order -> original vertex order
    for (std::uint32_t c = 0; c < nv; ++c)
        for (std::uint32_t r = 0; r < nv; ++r) {
            const std::size_t src = 2 * (static_cast<std::size_t>(c) * nv + r);
            const std::size_t dst = 2 * (static_cast<std::size_t>(vertex_of_[c]) * nv + vertex_of_[r]);
            A_out[dst] = a[src];
            A_out[dst + 1] = a[src + 1];
        }
    for (std::uint32_t c = 0; c < nt; ++c)
        for (std::uint32_t r = 0; r < nv; ++r) {
            const std::size_t src = 2 * (static_cast<std::size_t>(c) * nv + r);
            const std::size_t dst = 2 * (static_cast<std::size_t>(c) * nv + vertex_of_[r]);
            B_out[dst] = b[src];
            B_out[dst + 1] = b[src + 1];
        }
}

GmresWorkspace::~GmresWorkspace() { release(); }

void GmresWorkspace::release() {
    if (step_graph)
        (void)hipGraphExecDestroy(step_graph);
    step_graph = nullptr;
    step_key = {};
    cycle_current.release();
    cycle_w.release();
    cycle_z.release();
    cycle_small.release();
    cycle_state.release();
    if (cycle_pinned)
        (void)hipHostFree(cycle_pinned);
    cycle_pinned = nullptr;
    cycle_pinned_ints = 0;
    for (hipEvent_t e : cycle_events)
        (void)hipEventDestroy(e);
    cycle_events.clear();
    multi_basis.release();
    multi_in.release();
    multi_out.release();
    multi_small.release();
    if (multi_pinned)
        (void)hipHostFree(multi_pinned);
    multi_pinned = nullptr;
    multi_pinned_doubles = 0;
    basis.release();
    r.release();
    z.release();
    scratch.release();
    small.release();
    if (pinned)
        (void)hipHostFree(pinned);
    pinned = nullptr;
    pinned_doubles = 0;
    for (auto e : events)
        (void)hipEventDestroy(e);
    events.clear();
}

// Restarted GMRES with twice classical Gram-Schmidt on the device; host
// Givens rotations (device_gmres.cu:197-279). Right preconditioning by the
// block-Jacobi inverses when given. The Arnoldi process is normalised on the
// device, so step j + 1 is queued before the host reads step j's column (from
// pinned memory, by event): the device never waits for the host's rotations,
// and at most one speculative step per restart cycle is discarded.
namespace {

// Restarted GMRES on the dense operator with device-resident cycles: one Arnoldi
// step (v_j, block-Jacobi, the product with A, CGS2, the rotation) is captured once
// into a HIP graph and replayed once per iteration; the host reads the cycle state
// every kPoll steps, one poll behind the queue, so it never waits for a step to
// finish before queuing the next, and solves the small triangular system once per
// cycle. Same arithmetic as gmres(): the same reductions in the same order and the
// same rotation formulas (evaluated on the device).
constexpr std::uint32_t kPoll = 4;

void dense_gmres_graph(hipStream_t s, const void *A, bool fp32, std::uint32_t n, const double *b, double *x,
                       std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved,
                       std::uint32_t *iterations, const volatile std::int32_t *cancel, GmresWorkspace &ws,
                       const BlockJacobi &jacobi) {
    using C = std::complex<double>;
    const std::size_t vec = 2 * static_cast<std::size_t>(n), stride = restart + 1;
    if (ws.basis.size() < vec * stride)
        ws.basis.resize(vec * stride);
    if (ws.r.size() < vec)
        ws.r.resize(vec);
    for (auto *buf : {&ws.cycle_current, &ws.cycle_w, &ws.cycle_z})
        if (buf->size() < vec)
            buf->resize(vec);
    const std::size_t scratch =
        std::max({dev::gemv_scratch_doubles(n, 1), dev::gmres_cycle_scratch_doubles(n, restart),
                  dev::reduction_scratch_doubles(n, 1)});
    if (ws.scratch.size() < scratch)
        ws.scratch.resize(scratch);
    if (ws.cycle_small.size() < dev::gmres_cycle_small_doubles(restart))
        ws.cycle_small.resize(dev::gmres_cycle_small_doubles(restart));
    if (ws.cycle_state.size() < 8)
        ws.cycle_state.resize(8);
    if (ws.small.size() < 16)
        ws.small.resize(16);
    const std::size_t polls = restart / kPoll + 2;
    if (ws.cycle_pinned_ints < 8 * polls) {
        if (ws.cycle_pinned)
            (void)hipHostFree(ws.cycle_pinned);
        ws.cycle_pinned = nullptr;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&ws.cycle_pinned), sizeof(int) * 8 * polls),
                  "pinned GMRES cycle state");
        ws.cycle_pinned_ints = 8 * polls;
    }
    while (ws.cycle_events.size() < polls) {
        hipEvent_t e;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        ws.cycle_events.push_back(e);
    }
    const dev::GmresDeviceCycle cyc{ws.basis.data(), ws.cycle_current.data(), ws.cycle_w.data(), ws.scratch.data(),
                                    ws.cycle_small.data(), ws.cycle_state.data(), n, restart};
    double *nrm = ws.small.data();
    auto norm = [&](const double *v) {
        dev::launch_norm2(v, n, nullptr, nrm, ws.scratch.data(), s);
        double out = 0;
        hip_check(hipMemcpyAsync(&out, nrm, sizeof(double), hipMemcpyDeviceToHost, s), "norm download");
        hip_check(hipStreamSynchronize(s), "norm sync");
        return std::sqrt(out);
    };
    // The captured step: its arguments (buffers, A, sizes) are fixed; recapture when they change.
    const auto key = std::make_tuple(A, fp32, n, restart, jacobi.inverses, jacobi.width,
                                     static_cast<const double *>(ws.basis.data()),
                                     static_cast<const double *>(ws.scratch.data()));
    if (!ws.step_graph || ws.step_key != key) {
        if (ws.step_graph)
            (void)hipGraphExecDestroy(ws.step_graph);
        ws.step_graph = nullptr;
        // one eager product first: the column chunk is tuned outside the capture
        dev::launch_gemv(A, fp32, n, b, ws.cycle_w.data(), 1, ws.scratch.data(), s);
        hip_check(hipStreamSynchronize(s), "product warm-up");
        hipGraph_t graph = nullptr;
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture begin");
        dev::gmres_cycle_load(cyc, s);
        const double *in = ws.cycle_current.data();
        if (jacobi) {
            apply_block_jacobi(jacobi, n, in, ws.cycle_z.data(), s);
            in = ws.cycle_z.data();
        }
        dev::launch_gemv(A, fp32, n, in, ws.cycle_w.data(), 1, ws.scratch.data(), s, ws.cycle_state.data() + 1);
        dev::gmres_cycle_orthogonalize(cyc, s);
        const hipError_t captured = hipStreamEndCapture(s, &graph);
        hip_check(captured, "capture end");
        const hipError_t made = hipGraphInstantiate(&ws.step_graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        hip_check(made, "graph instantiate");
        ws.step_key = key;
    }
    hip_check(hipMemsetAsync(x, 0, vec * sizeof(double), s), "gmres clear");
    const double bnorm = norm(b);
    *iterations = 0;
    if (bnorm == 0) {
        *achieved = 0;
        return;
    }
    const double target = tol * bnorm;
    hip_check(hipMemcpyAsync(ws.r.data(), b, vec * sizeof(double), hipMemcpyDeviceToDevice, s), "gmres copy");
    double beta = bnorm;
    std::uint32_t it = 0;
    std::vector<double> host_small(dev::gmres_cycle_small_doubles(restart));
    const std::size_t hoff = dev::gmres_cycle_h_offset(restart), goff = dev::gmres_cycle_g_offset(restart);
    for (;;) {
        if (beta <= target) {
            *achieved = beta / bnorm;
            *iterations = it;
            return;
        }
        if (it >= max_it)
            break;
        dev::launch_scale(n, 1.0 / beta, ws.r.data(), ws.basis.data(), s);
        const std::uint32_t limit = std::min(restart, max_it - it);
        dev::gmres_cycle_begin(cyc, beta, target, limit, s);
        // queue the steps; every kPoll steps copy the state out and look at the
        // previous copy (one poll behind), stopping once a copy shows the cycle done
        std::uint32_t launched = 0, poll = 0;
        bool done = false;
        while (launched < limit && !done) {
            if (cancel && *cancel) {
                (void)hipStreamSynchronize(s);
                throw Cancelled();
            }
            hip_check(hipGraphLaunch(ws.step_graph, s), "step launch");
            ++launched;
            if (launched % kPoll == 0 || launched == limit) {
                int *slot = ws.cycle_pinned + 8 * poll;
                hip_check(hipMemcpyAsync(slot, ws.cycle_state.data(), 8 * sizeof(int), hipMemcpyDeviceToHost, s),
                          "cycle state");
                hip_check(hipEventRecord(ws.cycle_events[poll], s), "cycle event");
                if (poll > 0) {
                    hip_check(hipEventSynchronize(ws.cycle_events[poll - 1]), "cycle poll");
                    done = ws.cycle_pinned[8 * (poll - 1) + 1] != 0;
                }
                ++poll;
            }
        }
        hip_check(hipMemcpyAsync(host_small.data(), ws.cycle_small.data(), sizeof(double) * host_small.size(),
                                 hipMemcpyDeviceToHost, s),
                  "cycle download");
        int state[8] = {};
        hip_check(hipMemcpyAsync(state, ws.cycle_state.data(), sizeof(state), hipMemcpyDeviceToHost, s), "state");
        hip_check(hipStreamSynchronize(s), "cycle sync");
        if (state[4])
            throw std::runtime_error("GMRES: the Arnoldi process broke down");
        const std::uint32_t used = static_cast<std::uint32_t>(state[2]);
        it += used;
        const C *H = reinterpret_cast<const C *>(host_small.data() + hoff);
        const C *g = reinterpret_cast<const C *>(host_small.data() + goff);
        std::vector<C> y(used);
        for (std::uint32_t row = used; row-- > 0;) {
            C v = g[row];
            for (std::uint32_t c = row + 1; c < used; ++c)
                v -= H[static_cast<std::size_t>(c) * stride + row] * y[c];
            y[row] = v / H[static_cast<std::size_t>(row) * stride + row];
        }
        std::vector<double> neg(2 * used);
        for (std::uint32_t j = 0; j < used; ++j) {
            neg[2 * j] = -y[j].real();
            neg[2 * j + 1] = -y[j].imag();
        }
        double *ybuf = ws.cycle_small.data(); // h1 slot: free between cycles
        if (used) {
            hip_check(hipMemcpyAsync(ybuf, neg.data(), sizeof(double) * neg.size(), hipMemcpyHostToDevice, s),
                      "y upload");
            if (jacobi) { // x += M^{-1} V y
                hip_check(hipMemsetAsync(ws.r.data(), 0, vec * sizeof(double), s), "gmres clear");
                dev::launch_subtract_combination(ws.basis.data(), n, used, ybuf, ws.r.data(), s);
                apply_block_jacobi(jacobi, n, ws.r.data(), ws.cycle_z.data(), s);
                dev::launch_add(n, ws.cycle_z.data(), x, s);
            } else {
                dev::launch_subtract_combination(ws.basis.data(), n, used, ybuf, x, s);
            }
        }
        dev::launch_residual(A, fp32, n, b, x, 1, ws.r.data(), ws.scratch.data(), s);
        beta = norm(ws.r.data());
    }
    *achieved = beta / bnorm;
    *iterations = it;
    std::ostringstream msg;
    msg << "GMRES stopped without converging after " << it << "/" << max_it << " iterations: relative residual " << *achieved;
    throw std::runtime_error(msg.str());
}

} // namespace

void apply_block_jacobi(const BlockJacobi &jacobi, std::uint32_t n, const double *v, double *out, hipStream_t s) {
    if (jacobi.width == 64)
        dev::jacobi_apply(jacobi.inverses, n, v, out, s);
    else
        dev::jacobi_apply_wide(jacobi.inverses, n, jacobi.width, v, out, s);
}

std::uint32_t dense_jacobi_width() {
    if (const char *v = std::getenv("HSBEM_DENSE_JACOBI_WIDTH")) {
        const unsigned long w = std::strtoul(v, nullptr, 10);
        if (w >= 64 && w <= dev::kMaxJacobiWidth && (w & (w - 1)) == 0)
            return static_cast<std::uint32_t>(w);
    }
    return 64;
}

namespace {

bool gmres_graph_enabled() {
    const char *v = std::getenv("HSBEM_GMRES_GRAPH");
    return !(v && v[0] == '0');
}

} // namespace

void dense_gmres(hipStream_t s, const void *A, bool fp32, std::uint32_t n, const double *b, double *x,
                 std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved, std::uint32_t *iterations,
                 const volatile std::int32_t *cancel, GmresWorkspace &ws, const BlockJacobi &jacobi) {
    if (gmres_graph_enabled()) {
        dense_gmres_graph(s, A, fp32, n, b, x, restart, max_it, tol, achieved, iterations, cancel, ws, jacobi);
        return;
    }
    const std::size_t scratch = dev::gemv_scratch_doubles(n, 1);
    if (ws.scratch.size() < scratch)
        ws.scratch.resize(scratch);
    KrylovOperator op;
    op.apply = [&, A, fp32, n](const double *v, double *w, hipStream_t st) {
        dev::launch_gemv(A, fp32, n, v, w, 1, ws.scratch.data(), st);
    };
    op.residual = [&, A, fp32, n](const double *bb, const double *xx, double *r, hipStream_t st) {
        dev::launch_residual(A, fp32, n, bb, xx, 1, r, ws.scratch.data(), st);
    };
    if (jacobi)
        op.precondition = [jacobi, n](const double *v, double *z, hipStream_t st) {
            apply_block_jacobi(jacobi, n, v, z, st);
        };
    gmres(s, op, n, b, x, restart, max_it, tol, achieved, iterations, cancel, ws);
}

void dense_gmres_multi(hipStream_t s, const void *A, bool fp32, std::uint32_t n, std::uint32_t nrhs, const double *b,
                       double *x, std::uint32_t restart, std::uint32_t max_it, double tol, double *achieved,
                       std::uint32_t *iterations, const volatile std::int32_t *cancel, GmresWorkspace &ws,
                       const BlockJacobi &jacobi) {
    using C = std::complex<double>;
    if (nrhs == 0 || nrhs > kGmresBatch)
        throw std::invalid_argument("dense_gmres_multi takes 1 to 8 right-hand sides");
    const std::size_t vec = 2 * static_cast<std::size_t>(n), stride = restart + 1;
    const std::size_t column = 4 * stride + 1, slot = column + 8; // per system: h1, h2, |w|^2 (+ a norm cell)
    if (ws.multi_basis.size() < nrhs * stride * vec)
        ws.multi_basis.resize(nrhs * stride * vec);
    if (ws.multi_in.size() < nrhs * vec)
        ws.multi_in.resize(nrhs * vec);
    if (ws.multi_out.size() < nrhs * vec)
        ws.multi_out.resize(nrhs * vec);
    if (ws.multi_small.size() < nrhs * slot)
        ws.multi_small.resize(nrhs * slot);
    const std::size_t scratch = std::max(dev::gemv_scratch_doubles(n, nrhs), dev::reduction_scratch_doubles(n, stride));
    if (ws.scratch.size() < scratch)
        ws.scratch.resize(scratch);
    if (ws.multi_pinned_doubles < nrhs * slot) {
        if (ws.multi_pinned)
            (void)hipHostFree(ws.multi_pinned);
        ws.multi_pinned = nullptr;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&ws.multi_pinned), sizeof(double) * nrhs * slot),
                  "pinned GMRES columns");
        ws.multi_pinned_doubles = nrhs * slot;
    }
    double *sc = ws.scratch.data(), *basis0 = ws.multi_basis.data(), *in = ws.multi_in.data(),
           *out = ws.multi_out.data(), *small0 = ws.multi_small.data(), *pinned = ws.multi_pinned;
    auto basis = [&](std::uint32_t e, std::uint32_t j) { return basis0 + (e * stride + j) * vec; };
    auto h1 = [&](std::uint32_t e) { return small0 + e * slot; };
    auto h2 = [&](std::uint32_t e) { return small0 + e * slot + 2 * stride; };
    auto nrm = [&](std::uint32_t e) { return small0 + e * slot + 4 * stride; };
    auto cell = [&](std::uint32_t e) { return small0 + e * slot + column; };
    auto xe = [&](std::uint32_t e) { return x + e * vec; };
    auto re = [&](std::uint32_t e) { return out + e * vec; }; // residuals live in the output pack between cycles
    // |v_e|^2 of the given vectors to the host (one synchronisation for all)
    auto norms = [&](auto vector_of, const std::vector<std::uint32_t> &which) {
        for (auto e : which)
            dev::launch_norm2(vector_of(e), n, nullptr, cell(e), sc, s);
        hip_check(hipMemcpyAsync(pinned, small0, sizeof(double) * nrhs * slot, hipMemcpyDeviceToHost, s), "norms");
        hip_check(hipStreamSynchronize(s), "norm sync");
        std::vector<double> v(nrhs, 0.0);
        for (auto e : which)
            v[e] = std::sqrt(pinned[e * slot + column]);
        return v;
    };
    std::vector<std::uint32_t> all(nrhs);
    for (std::uint32_t e = 0; e < nrhs; ++e)
        all[e] = e;
    hip_check(hipMemsetAsync(x, 0, sizeof(double) * nrhs * vec, s), "gmres clear");
    const auto bnorm = norms([&](std::uint32_t e) { return b + e * vec; }, all);
    hip_check(hipMemcpyAsync(out, b, sizeof(double) * nrhs * vec, hipMemcpyDeviceToDevice, s), "gmres copy");
    std::vector<double> beta = bnorm, target(nrhs);
    std::vector<std::uint32_t> it(nrhs, 0), open;
    for (std::uint32_t e = 0; e < nrhs; ++e) {
        target[e] = tol * bnorm[e];
        if (bnorm[e] == 0) {
            achieved[e] = 0;
            iterations[e] = 0;
        } else {
            open.push_back(e);
        }
    }
    std::vector<std::vector<C>> H(nrhs, std::vector<C>(stride * restart)), g(nrhs, std::vector<C>(stride)),
        sn(nrhs, std::vector<C>(restart));
    std::vector<std::vector<double>> cs(nrhs, std::vector<double>(restart));
    std::vector<std::uint32_t> used(nrhs);
    while (!open.empty()) {
        // systems whose true residual meets the target are done; the rest start a cycle
        std::vector<std::uint32_t> cycle;
        for (auto e : open) {
            if (beta[e] <= target[e]) {
                achieved[e] = beta[e] / bnorm[e];
                iterations[e] = it[e];
            } else if (it[e] >= max_it) {
                std::ostringstream msg;
                msg << "GMRES stopped without converging after " << it[e] << "/" << max_it
                    << " iterations: relative residual " << beta[e] / bnorm[e];
                throw std::runtime_error(msg.str());
            } else {
                cycle.push_back(e);
            }
        }
        open = cycle;
        if (open.empty())
            break;
        for (auto e : cycle) {
            dev::launch_scale(n, 1.0 / beta[e], re(e), basis(e, 0), s);
            std::fill(H[e].begin(), H[e].end(), C{});
            std::fill(g[e].begin(), g[e].end(), C{});
            g[e][0] = beta[e];
            used[e] = 0;
        }
        std::vector<std::uint32_t> active = cycle;
        for (std::uint32_t j = 0; j < restart && !active.empty(); ++j) {
            if (cancel && *cancel) {
                (void)hipStreamSynchronize(s);
                throw Cancelled();
            }
            const auto c = static_cast<std::uint32_t>(active.size());
            for (std::uint32_t a = 0; a < c; ++a) { // pack M^{-1} v_j of the active systems
                if (jacobi)
                    apply_block_jacobi(jacobi, n, basis(active[a], j), in + a * vec, s);
                else
                    hip_check(hipMemcpyAsync(in + a * vec, basis(active[a], j), sizeof(double) * vec,
                                             hipMemcpyDeviceToDevice, s), "pack");
            }
            dev::launch_gemv(A, fp32, n, in, out, c, sc, s);
            for (std::uint32_t a = 0; a < c; ++a) {
                const std::uint32_t e = active[a];
                double *w = basis(e, j + 1);
                hip_check(hipMemcpyAsync(w, out + a * vec, sizeof(double) * vec, hipMemcpyDeviceToDevice, s), "unpack");
                dev::launch_project(basis(e, 0), n, j + 1, w, h1(e), sc, s);
                dev::launch_subtract_combination(basis(e, 0), n, j + 1, h1(e), w, s);
                dev::launch_project(basis(e, 0), n, j + 1, w, h2(e), sc, s);
                dev::launch_subtract_combination(basis(e, 0), n, j + 1, h2(e), w, s);
                dev::launch_norm2(w, n, nullptr, nrm(e), sc, s);
                dev::launch_normalize(n, nrm(e), w, w, s);
            }
            hip_check(hipMemcpyAsync(pinned, small0, sizeof(double) * nrhs * slot, hipMemcpyDeviceToHost, s),
                      "arnoldi download");
            hip_check(hipStreamSynchronize(s), "arnoldi sync");
            std::vector<std::uint32_t> still;
            for (auto e : active) { // Givens rotations of system e's column j (as gmres)
                const double *host = pinned + e * slot;
                C *col = H[e].data() + static_cast<std::size_t>(j) * stride;
                for (std::uint32_t i = 0; i <= j; ++i)
                    col[i] = C(host[2 * i], host[2 * i + 1]) + C(host[2 * stride + 2 * i], host[2 * stride + 2 * i + 1]);
                const double next = std::sqrt(host[4 * stride]);
                col[j + 1] = next;
                const bool breakdown = !(next > 0);
                for (std::uint32_t i = 0; i < j; ++i) {
                    const C t = cs[e][i] * col[i] + sn[e][i] * col[i + 1];
                    col[i + 1] = -std::conj(sn[e][i]) * col[i] + cs[e][i] * col[i + 1];
                    col[i] = t;
                }
                const C a = col[j], bb = col[j + 1];
                const double len = std::hypot(std::abs(a), std::abs(bb));
                if (!(len > 0) || !std::isfinite(len))
                    throw std::runtime_error("GMRES: the Arnoldi process broke down");
                const C phase = std::abs(a) > 0 ? a / std::abs(a) : C(1, 0);
                cs[e][j] = std::abs(a) / len;
                sn[e][j] = phase * std::conj(bb) / len;
                col[j] = phase * len;
                col[j + 1] = 0;
                const C t = cs[e][j] * g[e][j] + sn[e][j] * g[e][j + 1];
                g[e][j + 1] = -std::conj(sn[e][j]) * g[e][j] + cs[e][j] * g[e][j + 1];
                g[e][j] = t;
                ++it[e];
                used[e] = j + 1;
                if (!(std::abs(g[e][j + 1]) <= target[e] || breakdown || it[e] >= max_it))
                    still.push_back(e);
            }
            active = still;
        }
        // x_e += M^{-1} V_e y_e, then the true residuals of the cycle's systems
        for (auto e : cycle) {
            const std::uint32_t m = used[e];
            std::vector<C> y(m);
            for (std::uint32_t row = m; row-- > 0;) {
                C v = g[e][row];
                for (std::uint32_t col = row + 1; col < m; ++col)
                    v -= H[e][static_cast<std::size_t>(col) * stride + row] * y[col];
                y[row] = v / H[e][static_cast<std::size_t>(row) * stride + row];
            }
            double *neg = pinned + e * slot; // host staging: system e's own slot (the stream is idle here)
            for (std::uint32_t j = 0; j < m; ++j) {
                neg[2 * j] = -y[j].real();
                neg[2 * j + 1] = -y[j].imag();
            }
            hip_check(hipMemcpyAsync(h1(e), neg, sizeof(double) * 2 * m, hipMemcpyHostToDevice, s), "y upload");
            if (jacobi) {
                hip_check(hipMemsetAsync(in, 0, sizeof(double) * vec, s), "gmres clear");
                dev::launch_subtract_combination(basis(e, 0), n, m, h1(e), in, s);
                apply_block_jacobi(jacobi, n, in, in + vec * (nrhs > 1 ? 1 : 0), s);
                dev::launch_add(n, in + vec * (nrhs > 1 ? 1 : 0), xe(e), s);
            } else {
                dev::launch_subtract_combination(basis(e, 0), n, m, h1(e), xe(e), s);
            }
        }
        dev::launch_residual(A, fp32, n, b, x, nrhs, out, sc, s);
        beta = norms(re, cycle);
    }
}

void gmres(hipStream_t s, const KrylovOperator &op, std::uint32_t n, const double *b, double *x, std::uint32_t restart,
           std::uint32_t max_it, double tol, double *achieved, std::uint32_t *iterations,
           const volatile std::int32_t *cancel, GmresWorkspace &ws) {
    using C = std::complex<double>;
    const bool jacobi = static_cast<bool>(op.precondition);
    auto &basis_ = ws.basis;
    auto &r_ = ws.r;
    auto &scratch_ = ws.scratch;
    if (basis_.size() < 2 * static_cast<std::size_t>(n) * (restart + 1))
        basis_.resize(2 * static_cast<std::size_t>(n) * (restart + 1));
    if (r_.size() < 2 * static_cast<std::size_t>(n))
        r_.resize(2 * static_cast<std::size_t>(n));
    if (jacobi && ws.z.size() < 2 * static_cast<std::size_t>(n))
        ws.z.resize(2 * static_cast<std::size_t>(n));
    const std::size_t scratch = std::max(dev::gemv_scratch_doubles(n, 1), dev::reduction_scratch_doubles(n, restart + 1));
    if (scratch_.size() < scratch)
        scratch_.resize(scratch);
    const std::size_t column = 4 * (restart + 1) + 1; // h1, h2 (complex, restart + 1 each) and |w|^2
    if (ws.small.size() < column + 8)
        ws.small.resize(column + 8);
    if (ws.pinned_doubles < column * restart) {
        if (ws.pinned)
            (void)hipHostFree(ws.pinned);
        ws.pinned = nullptr;
        hip_check(hipHostMalloc(reinterpret_cast<void **>(&ws.pinned), sizeof(double) * column * restart),
                  "pinned GMRES columns");
        ws.pinned_doubles = column * restart;
    }
    while (ws.events.size() < restart) {
        hipEvent_t e;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        ws.events.push_back(e);
    }
    auto &small = ws.small;
    double *h1 = small.data(), *h2 = small.data() + 2 * (restart + 1), *nrm = small.data() + 4 * (restart + 1);
    auto norm = [&](const double *v) {
        dev::launch_norm2(v, n, nullptr, nrm, scratch_.data(), s);
        double out = 0;
        hip_check(hipMemcpyAsync(&out, nrm, sizeof(double), hipMemcpyDeviceToHost, s), "norm download");
        hip_check(hipStreamSynchronize(s), "norm sync");
        return std::sqrt(out);
    };
    hip_check(hipMemsetAsync(x, 0, 2 * sizeof(double) * n, s), "gmres clear");
    const double bnorm = norm(b);
    *iterations = 0;
    if (bnorm == 0) {
        *achieved = 0;
        return;
    }
    const double target = tol * bnorm;
    hip_check(hipMemcpyAsync(r_.data(), b, 2 * sizeof(double) * n, hipMemcpyDeviceToDevice, s), "gmres copy");
    double beta = bnorm;
    const std::size_t stride = restart + 1;
    std::vector<C> H(stride * restart), g(stride), sn(restart);
    std::vector<double> cs(restart);
    std::uint32_t it = 0;
    // One Arnoldi step: w = A M^{-1} v_j, CGS2 against v_0..v_j, |w|^2, v_{j+1} = w/|w|,
    // then the column (h1, h2, |w|^2) to pinned slot j.
    auto enqueue_step = [&](std::uint32_t j) {
        double *w = basis_.data() + 2 * static_cast<std::size_t>(j + 1) * n;
        const double *vj = basis_.data() + 2 * static_cast<std::size_t>(j) * n;
        if (jacobi) {
            op.precondition(vj, ws.z.data(), s);
            vj = ws.z.data();
        }
        op.apply(vj, w, s);
        dev::launch_project(basis_.data(), n, j + 1, w, h1, scratch_.data(), s);
        dev::launch_subtract_combination(basis_.data(), n, j + 1, h1, w, s);
        dev::launch_project(basis_.data(), n, j + 1, w, h2, scratch_.data(), s);
        dev::launch_subtract_combination(basis_.data(), n, j + 1, h2, w, s);
        dev::launch_norm2(w, n, nullptr, nrm, scratch_.data(), s);
        dev::launch_normalize(n, nrm, w, w, s);
        hip_check(hipMemcpyAsync(ws.pinned + column * j, small.data(), sizeof(double) * column, hipMemcpyDeviceToHost, s),
                  "arnoldi download");
        hip_check(hipEventRecord(ws.events[j], s), "arnoldi event");
    };
    for (;;) {
        if (beta <= target) {
            *achieved = beta / bnorm;
            *iterations = it;
            return;
        }
        if (it >= max_it)
            break;
        dev::launch_scale(n, 1.0 / beta, r_.data(), basis_.data(), s);
        std::fill(H.begin(), H.end(), C{});
        std::fill(g.begin(), g.end(), C{});
        g[0] = beta;
        std::uint32_t used = 0;
        enqueue_step(0);
        for (std::uint32_t j = 0; j < restart && it < max_it; ++j) {
            if (cancel && *cancel) {
                (void)hipStreamSynchronize(s);
                throw Cancelled();
            }
            if (j + 1 < restart && it + 1 < max_it)
                enqueue_step(j + 1); // speculative: keeps the device busy while the host rotates column j
            hip_check(hipEventSynchronize(ws.events[j]), "arnoldi sync");
            const double *host = ws.pinned + column * j;
            C *col = H.data() + static_cast<std::size_t>(j) * stride;
            for (std::uint32_t i = 0; i <= j; ++i)
                col[i] = C(host[2 * i], host[2 * i + 1]) + C(host[2 * (restart + 1) + 2 * i], host[2 * (restart + 1) + 2 * i + 1]);
            const double next = std::sqrt(host[4 * (restart + 1)]);
            col[j + 1] = next;
            const bool breakdown = !(next > 0);
            for (std::uint32_t i = 0; i < j; ++i) {
                const C t = cs[i] * col[i] + sn[i] * col[i + 1];
                col[i + 1] = -std::conj(sn[i]) * col[i] + cs[i] * col[i + 1];
                col[i] = t;
            }
            const C a = col[j], bb = col[j + 1];
            const double len = std::hypot(std::abs(a), std::abs(bb));
            if (!(len > 0) || !std::isfinite(len)) {
                (void)hipStreamSynchronize(s);
                throw std::runtime_error("GMRES: the Arnoldi process broke down");
            }
            const C phase = std::abs(a) > 0 ? a / std::abs(a) : C(1, 0);
            cs[j] = std::abs(a) / len;
            sn[j] = phase * std::conj(bb) / len;
            col[j] = phase * len;
            col[j + 1] = 0;
            const C t = cs[j] * g[j] + sn[j] * g[j + 1];
            g[j + 1] = -std::conj(sn[j]) * g[j] + cs[j] * g[j + 1];
            g[j] = t;
            ++it;
            used = j + 1;
            if (std::abs(g[j + 1]) <= target || breakdown)
                break;
        }
        std::vector<C> y(used);
        for (std::uint32_t row = used; row-- > 0;) {
            C v = g[row];
            for (std::uint32_t c = row + 1; c < used; ++c)
                v -= H[static_cast<std::size_t>(c) * stride + row] * y[c];
            y[row] = v / H[static_cast<std::size_t>(row) * stride + row];
        }
        std::vector<double> neg(2 * used);
        for (std::uint32_t j = 0; j < used; ++j) {
            neg[2 * j] = -y[j].real();
            neg[2 * j + 1] = -y[j].imag();
        }
        // h1 is free once every queued step has read it (stream order).
        hip_check(hipMemcpyAsync(h1, neg.data(), sizeof(double) * neg.size(), hipMemcpyHostToDevice, s), "y upload");
        if (jacobi) { // x += M^{-1} V y
            hip_check(hipMemsetAsync(r_.data(), 0, 2 * sizeof(double) * n, s), "gmres clear");
            dev::launch_subtract_combination(basis_.data(), n, used, h1, r_.data(), s);
            op.precondition(r_.data(), ws.z.data(), s);
            dev::launch_add(n, ws.z.data(), x, s);
        } else {
            dev::launch_subtract_combination(basis_.data(), n, used, h1, x, s);
        }
        if (op.residual) {
            op.residual(b, x, r_.data(), s);
        } else {
            op.apply(x, r_.data(), s);
            dev::launch_subtract_from(n, b, r_.data(), s);
        }
        beta = norm(r_.data());
    }
    *achieved = beta / bnorm;
    *iterations = it;
    std::ostringstream msg;
    msg << "GMRES stopped without converging after " << it << "/" << max_it << " iterations: relative residual " << *achieved;
    throw std::runtime_error(msg.str());
}

std::shared_ptr<PreparedRequest> MeshSolver::prepare(const Request &req) const {
    using C = std::complex<double>;
    const Numerics &num = req.numerics;
    if (!std::isfinite(req.density) || req.density <= 0 || !std::isfinite(req.sound_speed) || req.sound_speed <= 0 ||
        req.frequencies.empty() || req.excitations.empty() ||
        (num.quadrature_order != 2 && num.quadrature_order != 4) || num.singular_order < 2 || num.singular_order > 20 ||
        !std::isfinite(num.near_ratio) || num.near_ratio < 0 || num.near_ratio > 10 ||
        !std::isfinite(num.near_tolerance) || num.near_tolerance < 1e-10 || num.near_tolerance > .01 ||
        num.near_depth > 10 || !std::isfinite(num.residual_tolerance) || num.residual_tolerance < 1e-12 ||
        num.residual_tolerance > .01 || num.refinement_steps > 10 || num.solver > 1 || num.precision > 1 ||
        num.preconditioner > 1 ||
        num.gmres_restart < 2 || num.gmres_restart > 300 || num.gmres_max_iterations < 1 || num.acceleration > 1)
        throw std::invalid_argument("invalid medium, frequencies, excitations or quadrature settings");
    // fast multipole settings (cuda_solver.cu:1057-1067)
    if (num.acceleration == 1 &&
        (num.fmm_leaf_capacity < 4 || num.fmm_leaf_capacity > 256 || !std::isfinite(num.fmm_separation) ||
         num.fmm_separation < 1.5 || num.fmm_separation > 10 || num.fmm_maximum_iterations < 1 ||
         num.fmm_maximum_iterations > 10000 || num.fmm_restart < 2 || num.fmm_restart > 300 ||
         num.fmm_block_width < 32 || num.fmm_block_width > 2048 ||
         (num.fmm_block_width & (num.fmm_block_width - 1)) != 0))
        throw std::invalid_argument("invalid Fast Multipole or GMRES settings (leaf capacity, separation, restart, iterations or block width)");
    if (req.excitations.size() > 1024 || req.observers.size() > 100000)
        throw std::invalid_argument("the request is too large for 32-bit indices");
    static std::atomic<std::uint64_t> next_id{1};
    auto P = std::make_shared<PreparedRequest>();
    P->id = next_id++;
    P->request = req;
    P->k_max = planning_wavenumber(req.frequencies, req.planning_frequencies, req.sound_speed);
    // observers (recentred, or unit directions about the far-field origin)
    if (req.far_field) {
        P->far_origin = vsub(req.far_origin, origin_);
        const Vec3 &o = P->far_origin;
        if (!std::isfinite(o.x) || !std::isfinite(o.y) || !std::isfinite(o.z) ||
            !std::isfinite(req.far_reference_distance) || req.far_reference_distance <= 0 ||
            !std::isfinite(P->k_max * req.far_reference_distance))
            throw std::invalid_argument("far field needs a finite origin and a finite positive reference distance");
    }
    for (const auto &w : req.observers) {
        Vec3 p = vsub(w, req.far_field ? req.far_origin : origin_);
        if (!std::isfinite(w.x) || !std::isfinite(w.y) || !std::isfinite(w.z) || !std::isfinite(p.x) ||
            !std::isfinite(p.y) || !std::isfinite(p.z))
            throw std::invalid_argument("observer coordinates must be finite");
        if (req.far_field) {
            const double d = std::hypot(p.x, p.y, p.z);
            if (!(d > 0) || !std::isfinite(d))
                throw std::invalid_argument("a far-field observer must not sit at the origin");
            p = {p.x / d, p.y / d, p.z / d};
        } else if (point_inside(space_.panels, p)) {
            throw std::invalid_argument("observers must be finite and outside the body");
        }
        P->observers.push_back(p);
    }
    // boundary conditions per face and panel velocities per excitation
    std::map<std::uint32_t, Boundary> conditions;
    for (const auto &t : mesh_.triangles) {
        const auto [it, inserted] = conditions.emplace(t.face_id, t.boundary);
        if (!inserted && it->second != t.boundary)
            throw std::invalid_argument("one face id carries both wall and driven triangles");
    }
    const std::uint32_t nt = this->nt(), nv = this->nv(), ne = static_cast<std::uint32_t>(req.excitations.size());
    std::set<std::uint32_t> active_faces;
    P->face_velocity.resize(ne);
    for (std::uint32_t e = 0; e < ne; ++e) {
        for (const auto &[face, v] : req.excitations[e].faces) {
            if (!conditions.count(face) || conditions.at(face) != Boundary::driven || !std::isfinite(v.real()) ||
                !std::isfinite(v.imag()) || !P->face_velocity[e].emplace(face, v).second)
                throw std::invalid_argument("an excitation lists a face twice or a face that is not driven");
            active_faces.insert(face);
            P->has_nonzero |= v != C(0);
        }
        if (P->face_velocity[e].empty())
            throw std::invalid_argument("every excitation needs at least one driven face");
    }
    P->flux_column.assign(nt, -1);
    for (std::uint32_t t = 0; t < nt; ++t)
        if (active_faces.count(mesh_.triangles[t].face_id))
            P->flux_column[t] = static_cast<std::int32_t>(P->nd++);
    P->ne = ne;
    P->nv = nv;
    P->nt = nt;
    // device bytes of one frequency point (buffers of bind() and dense_gmres())
    const bool mixed = num.precision == 1;
    const bool lu = !mixed && num.solver == 0;
    const std::uint64_t nv2 = static_cast<std::uint64_t>(nv) * nv, nd1 = std::max<std::uint32_t>(P->nd, 1);
    const std::uint64_t nobs = P->observers.size();
    const std::uint64_t scratch = std::max({dev::gemv_scratch_doubles(nv, std::min<std::uint32_t>(ne, 8)),
                                            dev::reduction_scratch_doubles(nv, num.gmres_restart + 1), std::size_t{4096}});
    const bool fmm = num.acceleration == 1;
    std::uint64_t bytes = (fmm ? 0 : (mixed ? 8 : 16) * nv2) + 16 * nv * nd1 + 16 * nd1 * ne + 3 * 16ull * nv * ne +
                          16ull * nt * ne + 8 * scratch + 4ull * nt;
    if (fmm) {
        // operators of the point (near entries, leaf moments, expansions, link workspaces,
        // audit / verification / comparison operators): the largest measured on this mesh
        // with these settings, else an estimate of about 200 KB per node (measured: 166 KB
        // per node on a 10k-node horn, 203 KB on a 59k-node flanged horn at 16 kHz)
        const std::uint64_t measured = const_cast<MeshSolver *>(this)->fmm_point_bytes(num);
        bytes += measured ? measured + measured / 10 : 16ull * nv * 12288 + (64ull << 20);
        bytes += 16ull * nv * (num.fmm_restart + 2) + 16ull * nv + 8ull * (4 * (num.fmm_restart + 2) + 8) +
                 8 * (num.fmm_block_width == 64 ? dev::jacobi_factor_doubles(nv)
                                                : 2 * dev::jacobi_wide_doubles(nv, num.fmm_block_width));
    }
    else if (lu)
        bytes += 16 * nv2 + 4ull * nv;
    else
        bytes += 16ull * nv * (num.gmres_restart + 2) + 16ull * nv + 8 * scratch + 8ull * (4 * (num.gmres_restart + 2) + 8) +
                 (ne > 1 ? 16ull * nv * std::min<std::uint32_t>(ne, kGmresBatch) * (num.gmres_restart + 3) +
                               8ull * std::min<std::uint32_t>(ne, kGmresBatch) * (4 * (num.gmres_restart + 2) + 16) +
                               8 * dev::gemv_scratch_doubles(nv, std::min<std::uint32_t>(ne, kGmresBatch))
                         : 0) +
                 (!num.preconditioner ? 0
                  : dense_jacobi_width() == 64
                      ? 8 * dev::jacobi_factor_doubles(nv)
                      : 16 * dev::jacobi_wide_doubles(nv, dense_jacobi_width()) +
                            4 * dev::jacobi_wide_doubles(nv, dense_jacobi_width()) / (2 * dense_jacobi_width()));
    if (nobs)
        bytes += 24 * nobs + 16 * nobs * ne + 8 * dev::field_partial_doubles(nt, static_cast<std::uint32_t>(nobs), ne);
    P->footprint = bytes;
    // resident plan: touching jobs (~13 per panel, 16 B), near CSR and patches, cluster
    // tiles (8 B per cluster pair); the FMM near structure adds its job lists (about
    // 27 leaves x leaf capacity near panels per panel, 16 B each)
    const std::uint64_t clusters = nt / 48 + 1;
    P->plan_estimate = 16ull * 16 * nt + 8 * clusters * clusters + 8ull * nt * 16;
    if (fmm)
        P->plan_estimate += 16ull * nt * 27 * num.fmm_leaf_capacity;
    return P;
}

void MeshSolver::bind(const PreparedRequest &P, Workspace &ws) {
    const Numerics &num = P.request.numerics;
    const std::uint32_t nt = this->nt(), nv = this->nv(), ne = P.ne;
    const std::size_t nv2 = static_cast<std::size_t>(nv) * nv;
    const bool mixed = num.precision == 1;
    const bool fmm = num.acceleration == 1;
    ws.A.resize(fmm ? 0 : (mixed ? 1 : 2) * nv2);
    ws.BD.resize(2 * static_cast<std::size_t>(nv) * std::max<std::uint32_t>(P.nd, 1));
    ws.q.resize(2 * static_cast<std::size_t>(std::max<std::uint32_t>(P.nd, 1)) * ne);
    ws.rhs.resize(2 * static_cast<std::size_t>(nv) * ne);
    ws.x.resize(2 * static_cast<std::size_t>(nv) * ne);
    ws.r.resize(2 * static_cast<std::size_t>(nv) * ne);
    ws.flux.resize(2 * static_cast<std::size_t>(nt) * ne);
    ws.flux_column.upload(P.flux_column, ws.stream);
    const std::size_t scratch = std::max({dev::gemv_scratch_doubles(nv, std::min<std::uint32_t>(ne, 8)),
                                          dev::reduction_scratch_doubles(nv, num.gmres_restart + 1), std::size_t{4096}});
    ws.scratch.resize(scratch);
    if (!fmm && !mixed && num.solver == 0) {
        ws.LU.resize(2 * nv2);
        ws.pivots.resize(nv);
    }
    if (!P.observers.empty()) {
        std::vector<double> o;
        for (const auto &p : P.observers) {
            o.push_back(p.x);
            o.push_back(p.y);
            o.push_back(p.z);
        }
        ws.obs.upload(o, ws.stream);
        ws.field.resize(2 * P.observers.size() * ne);
        ws.partials.resize(dev::field_partial_doubles(nt, static_cast<std::uint32_t>(P.observers.size()), ne));
    }
    hip_check(hipStreamSynchronize(ws.stream), "request upload");
    ws.bound = P.id;
}

Result MeshSolver::solve(const Request &req, ProgressFn progress, void *user, const volatile std::int32_t *cancel) {
    std::lock_guard lock(mutex_);
    if (cancel && *cancel)
        throw Cancelled();
    const auto P = prepare(req);
    Workspace &ws = *workspace_;
    Result result;
    result.nt = nt();
    result.nv = nv();
    result.nobs = static_cast<std::uint32_t>(P->observers.size());
    result.nexc = P->ne;
    std::uint32_t completed = 0;
    const auto total = static_cast<std::uint32_t>(req.frequencies.size());
    for (std::size_t i = 0; i < req.frequencies.size(); ++i) {
        const double f = req.frequencies[i];
        auto report = [&](std::int32_t phase) {
            if (progress)
                progress(user, completed, total, f, phase);
        };
        bool built = false;
        FrequencyResult fr = solve_frequency(*P, i, ws, cancel, &built, report);
        if (i != 0)
            fr.ms[5] = 0;
        result.frequencies.push_back(std::move(fr));
        ++completed;
        report(6);
    }
    result.peak_bytes = ws.capacity_bytes();
    return result;
}

FrequencyResult MeshSolver::solve_frequency(const PreparedRequest &P, std::size_t index, Workspace &ws,
                                            const volatile std::int32_t *cancel, bool *plan_built,
                                            const std::function<void(std::int32_t)> &report) {
    using C = std::complex<double>;
    const Request &req = P.request;
    const Numerics &num = req.numerics;
    auto cancelled = [&] {
        if (cancel && *cancel)
            throw Cancelled();
    };
    auto phase = [&](std::int32_t p) {
        if (report)
            report(p);
    };
    cancelled();
    hipStream_t s = ws.stream;
    bool built = false;
    const std::shared_ptr<DevicePlan> plan_ref = device_plan(num, P.k_max, s, &built, num.acceleration == 1 && device_estimates_);
    const DevicePlan &plan = *plan_ref;
    if (plan_built)
        *plan_built = built;
    if (ws.bound != P.id)
        bind(P, ws);
    cancelled();
    const std::uint32_t nt = this->nt(), nv = this->nv(), ne = P.ne, nd = P.nd;
    const bool mixed = num.precision == 1;
    const unsigned solver_kind = mixed ? 1u : num.solver;
    const std::size_t scratch = ws.scratch.size();
    const double f = req.frequencies[index];
    phase(0);
    FrequencyResult fr;
    fr.frequency = f;
    fr.ms[5] = built ? plan.plan_ms : 0;
    const double omega = 2 * std::numbers::pi * f, k = omega / req.sound_speed;
    double max_kh = 0;
    for (const auto &p : space_.panels)
        max_kh = std::max(max_kh, k * p.max_edge);
    // fluxes: q = -i rho omega v, per flux column and per panel
    std::vector<double> qd(2 * static_cast<std::size_t>(std::max<std::uint32_t>(nd, 1)) * ne, 0.0),
        qpanel(2 * static_cast<std::size_t>(nt) * ne, 0.0);
    for (std::uint32_t e = 0; e < ne; ++e)
        for (std::uint32_t t = 0; t < nt; ++t) {
            const auto it = P.face_velocity[e].find(mesh_.triangles[t].face_id);
            if (it == P.face_velocity[e].end())
                continue;
            const C q = C(0, -req.density * omega) * it->second;
            qpanel[2 * (static_cast<std::size_t>(e) * nt + t)] = q.real();
            qpanel[2 * (static_cast<std::size_t>(e) * nt + t) + 1] = q.imag();
            if (P.flux_column[t] >= 0) {
                qd[2 * (static_cast<std::size_t>(e) * nd + P.flux_column[t])] = q.real();
                qd[2 * (static_cast<std::size_t>(e) * nd + P.flux_column[t]) + 1] = q.imag();
            }
        }
    hip_check(hipMemcpyAsync(ws.q.data(), qd.data(), sizeof(double) * qd.size(), hipMemcpyHostToDevice, s), "q upload");
    hip_check(hipMemcpyAsync(ws.flux.data(), qpanel.data(), sizeof(double) * qpanel.size(), hipMemcpyHostToDevice, s),
              "flux upload");
    std::vector<std::uint32_t> iterations(ne, 0);
    std::uint32_t refinements = 0;
    std::vector<double> relative(ne);
    if (num.acceleration == 1) {
        phase(3);
        fmm_point(P, ws, k, omega, relative, iterations, fr, cancel);
    } else {
    const auto t_asm = Clock::now();
    auto L = assembly_launch(plan, num, k, ws.A.data(), ws.BD.data(), nd, ws.flux_column.data(), ws.error.data());
    L.args.samples = ws.samples.data();
    tune_regular_tiles(L, mixed, s); // its launches add into A, B, the error flag and the samples: cleared next
    hip_check(hipMemsetAsync(ws.A.data(), 0, ws.A.bytes(), s), "clear A");
    hip_check(hipMemsetAsync(ws.BD.data(), 0, ws.BD.bytes(), s), "clear B");
    hip_check(hipMemsetAsync(ws.error.data(), 0, sizeof(int), s), "clear error");
    hip_check(hipMemsetAsync(ws.samples.data(), 0, sizeof(unsigned long long), s), "clear samples");
    hipEvent_t *ev = ws.events;
    hip_check(hipEventRecord(ev[0], s), "event");
    dev::launch_regular(L, s);
    hip_check(hipEventRecord(ev[1], s), "event");
    dev::launch_touching(L, s);
    hip_check(hipEventRecord(ev[2], s), "event");
    dev::launch_near(L, s);
    hip_check(hipEventRecord(ev[3], s), "event");
    dev::launch_mass(L, s);
    hip_check(hipGetLastError(), "assembly launch");
    if (nd)
        dev::launch_flux_rhs(ws.BD.data(), nv, nd, ws.q.data(), ne, ws.rhs.data(), s);
    else
        hip_check(hipMemsetAsync(ws.rhs.data(), 0, ws.rhs.bytes(), s), "clear rhs");
    check_error_flag(ws.error, s);
    float e01 = 0, e12 = 0, e23 = 0;
    hip_check(hipEventElapsedTime(&e01, ev[0], ev[1]), "event time");
    hip_check(hipEventElapsedTime(&e12, ev[1], ev[2]), "event time");
    hip_check(hipEventElapsedTime(&e23, ev[2], ev[3]), "event time");
    unsigned long long regular_samples = 0;
    hip_check(hipMemcpyAsync(&regular_samples, ws.samples.data(), sizeof(regular_samples), hipMemcpyDeviceToHost, s),
              "samples download");
    hip_check(hipStreamSynchronize(s), "samples sync");
    fr.regular_samples = static_cast<double>(regular_samples);
    fr.touching_samples = plan.touching_samples;
    fr.near_samples = plan.near_samples;
    fr.regular_kernel_ms = e01;
    fr.touching_kernel_ms = e12;
    fr.near_kernel_ms = e23;
    fr.ms[0] = ms_since(t_asm);
    cancelled();
    // mass-weighted rhs norms (for the unscaled relative residual)
    std::vector<double> rhs_norm(ne);
    for (std::uint32_t e = 0; e < ne; ++e) {
        dev::launch_norm2(ws.rhs.data() + 2 * static_cast<std::size_t>(e) * nv, nv, lumped_.data(),
                          ws.scratch.data() + scratch - 8, ws.scratch.data(), s);
        double v = 0;
        hip_check(hipMemcpyAsync(&v, ws.scratch.data() + scratch - 8, sizeof(double), hipMemcpyDeviceToHost, s), "norm");
        hip_check(hipStreamSynchronize(s), "norm sync");
        rhs_norm[e] = std::sqrt(v);
        if (!std::isfinite(rhs_norm[e]))
            throw std::runtime_error("the assembled right-hand side is not finite");
    }
    phase(3);
    const auto t_solve = Clock::now();
    if (solver_kind == 0) {
        hip_check(hipMemcpyAsync(ws.LU.data(), ws.A.data(), ws.A.bytes(), hipMemcpyDeviceToDevice, s), "LU copy");
        hip_check(hipMemsetAsync(ws.error.data(), 0, sizeof(int), s), "clear info");
        const auto t_f = Clock::now();
        dev::lu_factor(ws.LU.data(), nv, ws.pivots.data(), ws.error.data(), nullptr, s);
        int info = 0;
        hip_check(hipMemcpyAsync(&info, ws.error.data(), sizeof(int), hipMemcpyDeviceToHost, s), "info");
        hip_check(hipStreamSynchronize(s), "LU sync");
        fr.ms[2] = ms_since(t_f);
        if (info)
            throw std::runtime_error("LU factorization hit a zero pivot at column " + std::to_string(info));
        const auto t_t = Clock::now();
        hip_check(hipMemcpyAsync(ws.x.data(), ws.rhs.data(), ws.rhs.bytes(), hipMemcpyDeviceToDevice, s), "rhs copy");
        dev::lu_solve(ws.LU.data(), nv, ws.pivots.data(), ws.x.data(), ne, s);
        hip_check(hipStreamSynchronize(s), "solve sync");
        fr.ms[3] = ms_since(t_t);
    } else {
        BlockJacobi jacobi;
        if (num.preconditioner == 1) {
            const auto t_f = Clock::now();
            jacobi.width = dense_jacobi_width();
            if (jacobi.width == 64) {
                ws.jacobi.resize(dev::jacobi_factor_doubles(nv));
                dev::jacobi_factor(ws.A.data(), mixed, nv, ws.jacobi.data(), s);
            } else {
                const std::size_t doubles = dev::jacobi_wide_doubles(nv, jacobi.width);
                ws.jacobi.resize(doubles);
                ws.jacobi_blocks.resize(doubles);
                ws.jacobi_pivots.resize(doubles / (2 * jacobi.width));
                dev::jacobi_gather_dense(ws.A.data(), mixed, nv, jacobi.width, ws.jacobi_blocks.data(), s);
                dev::jacobi_factor_wide(ws.jacobi_blocks.data(), nv, jacobi.width, ws.jacobi.data(),
                                        ws.jacobi_pivots.data(), s);
            }
            jacobi.inverses = ws.jacobi.data();
            hip_check(hipGetLastError(), "block-Jacobi factor");
            hip_check(hipStreamSynchronize(s), "block-Jacobi sync");
            fr.ms[2] = ms_since(t_f);
        }
        if (ne == 1) {
            double achieved = 0;
            dense_gmres(s, ws.A.data(), mixed, nv, ws.rhs.data(), ws.x.data(), num.gmres_restart,
                        num.gmres_max_iterations, num.residual_tolerance * 0.1, &achieved, &iterations[0], cancel,
                        ws.gmres, jacobi);
        } else { // right-hand sides in lockstep batches sharing each product with A
            std::vector<double> achieved(kGmresBatch);
            for (std::uint32_t e0 = 0; e0 < ne; e0 += kGmresBatch) {
                const std::uint32_t m = std::min(kGmresBatch, ne - e0);
                dense_gmres_multi(s, ws.A.data(), mixed, nv, m, ws.rhs.data() + 2 * static_cast<std::size_t>(e0) * nv,
                                  ws.x.data() + 2 * static_cast<std::size_t>(e0) * nv, num.gmres_restart,
                                  num.gmres_max_iterations, num.residual_tolerance * 0.1, achieved.data(),
                                  &iterations[e0], cancel, ws.gmres, jacobi);
            }
        }
    }
    // FP64 residual of the unscaled equation; LU corrections while above tolerance
    for (;;) {
        dev::launch_residual(ws.A.data(), mixed, nv, ws.rhs.data(), ws.x.data(), ne, ws.r.data(), ws.scratch.data(), s);
        double largest = 0;
        for (std::uint32_t e = 0; e < ne; ++e) {
            dev::launch_norm2(ws.r.data() + 2 * static_cast<std::size_t>(e) * nv, nv, lumped_.data(),
                              ws.scratch.data() + scratch - 8, ws.scratch.data(), s);
            double v = 0;
            hip_check(hipMemcpyAsync(&v, ws.scratch.data() + scratch - 8, sizeof(double), hipMemcpyDeviceToHost, s),
                      "residual norm");
            hip_check(hipStreamSynchronize(s), "residual sync");
            relative[e] = rhs_norm[e] > 0 ? std::sqrt(v) / rhs_norm[e]
                                           : (v == 0 ? 0 : std::numeric_limits<double>::infinity());
            if (!std::isfinite(relative[e]))
                throw std::runtime_error("relative residual is not finite.");
            largest = std::max(largest, relative[e]);
        }
        if (largest <= num.residual_tolerance)
            break;
        if (solver_kind != 0 || refinements >= num.refinement_steps) {
            std::ostringstream msg;
            msg.precision(9);
            msg << "relative residual " << largest << " is above the requested tolerance " << num.residual_tolerance << " at " << f
                << " Hz after " << refinements << " refinement steps.";
            throw std::runtime_error(msg.str());
        }
        dev::lu_solve(ws.LU.data(), nv, ws.pivots.data(), ws.r.data(), ne, s);
        for (std::uint32_t e = 0; e < ne; ++e)
            dev::launch_axpy(nv, unit_.data(), ws.r.data() + 2 * static_cast<std::size_t>(e) * nv,
                             ws.x.data() + 2 * static_cast<std::size_t>(e) * nv, s);
        hip_check(hipStreamSynchronize(s), "refinement sync");
        ++refinements;
    }
    fr.ms[1] = ms_since(t_solve);
    } // dense form
    // fields
    phase(5);
    const auto t_field = Clock::now();
    const auto &observers = P.observers;
    std::vector<double> observed;
    if (!observers.empty() && P.has_nonzero) {
        auto evaluate = [&](unsigned degree) {
            cancelled();
            const auto slot = endpoint_rule_slots_.find(degree);
            if (slot == endpoint_rule_slots_.end())
                throw std::logic_error("no resident observer rule of degree " + std::to_string(degree));
            dev::FieldArgs fa{};
            fa.panels = {geo_.data(), nodes_.data(), bounds_.data(), shape_.data(), nt, nv};
            fa.pressure = ws.x.data();
            fa.flux = ws.flux.data();
            fa.nrhs = ne;
            fa.points = ws.obs.data();
            fa.count = static_cast<std::uint32_t>(observers.size());
            fa.rule = endpoint_rules_.data() + 4 * static_cast<std::size_t>(slot->second.first);
            fa.rule_count = slot->second.second;
            fa.k = k;
            fa.near_tolerance = std::min(num.near_tolerance, .025 * num.residual_tolerance);
            fa.max_depth = num.near_depth;
            fa.origin[0] = P.far_origin.x;
            fa.origin[1] = P.far_origin.y;
            fa.origin[2] = P.far_origin.z;
            fa.reference_distance = req.far_reference_distance;
            fa.out = ws.field.data();
            fa.error = ws.error.data();
            hip_check(hipMemsetAsync(ws.error.data(), 0, sizeof(int), s), "clear error");
            if (req.far_field)
                dev::launch_far_field(fa, ws.partials.data(), s);
            else
                dev::launch_observer_field(fa, ws.partials.data(), s);
            hip_check(hipGetLastError(), "field launch");
            check_error_flag(ws.error, s);
            std::vector<double> out(ws.field.size());
            hip_check(hipMemcpyAsync(out.data(), ws.field.data(), ws.field.bytes(), hipMemcpyDeviceToHost, s),
                      "field download");
            hip_check(hipStreamSynchronize(s), "field sync");
            return out;
        };
        unsigned degree = minimum_endpoint_degree(max_kh, num.quadrature_order);
        observed = evaluate(degree);
        for (;;) {
            const unsigned next = refined_endpoint_degree(degree);
            auto refined = evaluate(next);
            bool passed = true;
            for (std::uint32_t e = 0; e < ne; ++e) {
                double diff = 0, norm = 0;
                for (std::size_t j = 0; j < observers.size(); ++j) {
                    const std::size_t i = 2 * (e * observers.size() + j);
                    const C hi(refined[i], refined[i + 1]), lo(observed[i], observed[i + 1]);
                    diff += std::norm(hi - lo);
                    norm += std::norm(hi);
                }
                const double ind = norm > 0 ? std::sqrt(diff / norm) : diff == 0 ? 0 : INFINITY;
                passed = passed && ind <= .025 * num.residual_tolerance;
            }
            observed = std::move(refined);
            if (passed)
                break;
            if (degree == 12)
                throw std::runtime_error("observer field: the degree 12/14 panel rule did not agree within "
                                         "0.025 x the residual tolerance; use a finer mesh or observers farther from the surface");
            degree = next;
        }
    } else if (!observers.empty()) {
        observed.assign(2 * observers.size() * ne, 0.0);
    }
    // nodal pressures to host
    std::vector<double> nodal(2 * static_cast<std::size_t>(nv) * ne);
    hip_check(hipMemcpyAsync(nodal.data(), ws.x.data(), sizeof(double) * nodal.size(), hipMemcpyDeviceToHost, s),
              "pressure download");
    hip_check(hipStreamSynchronize(s), "pressure sync");
    fr.ms[4] = ms_since(t_field);
    for (std::uint32_t e = 0; e < ne; ++e) {
        ExcitationResult er;
        er.residual = relative[e];
        er.iterations = iterations[e];
        er.refinements = refinements;
        er.pressure_trace.resize(nv);
        for (std::uint32_t i = 0; i < nv; ++i)
            er.pressure_trace[vertex_of_[i]] = C(nodal[2 * (static_cast<std::size_t>(e) * nv + i)],
                                                 nodal[2 * (static_cast<std::size_t>(e) * nv + i) + 1]);
        er.surface_pressure.resize(nt);
        er.normal_derivative.resize(nt);
        for (std::uint32_t t = 0; t < nt; ++t) {
            C avg = 0;
            for (auto v : space_.panels[t].nodes)
                avg += er.pressure_trace[v] / 3.0;
            if (!std::isfinite(avg.real()) || !std::isfinite(avg.imag()))
                throw std::runtime_error("the solved pressure is not finite");
            er.surface_pressure[t] = avg;
            const auto it = P.face_velocity[e].find(mesh_.triangles[t].face_id);
            er.normal_derivative[t] = it == P.face_velocity[e].end() ? C{} : C(0, -req.density * omega) * it->second;
        }
        for (std::size_t j = 0; j < observers.size(); ++j)
            er.observer_pressure.emplace_back(observed[2 * (e * observers.size() + j)],
                                              observed[2 * (e * observers.size() + j) + 1]);
        for (const auto &[face, v] : req.excitations[e].faces) {
            if (std::abs(v) == 0)
                continue;
            double area = 0;
            C integrated = 0;
            for (std::uint32_t t = 0; t < nt; ++t)
                if (mesh_.triangles[t].face_id == face) {
                    area += space_.panels[t].area;
                    integrated += er.surface_pressure[t] * space_.panels[t].area;
                }
            const C mean = integrated / area, volume = v * area;
            er.impedances.push_back({face, area, mean, volume, mean / volume});
        }
        fr.excitations.push_back(std::move(er));
    }
    return fr;
}

} // namespace hsbem
