#include "fmm_operator.hpp"

#include "autotune.hpp"

#include "host/fmm_levels.hpp"
#include "host/fmm_rotation.hpp"
#include "host/fmm_scaled.hpp"
#include "host/qualification.hpp"
#include "host/rules.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <future>
#include <map>
#include <numeric>
#include <stdexcept>
#include <thread>
#include <sstream>
#include <numbers>
#include <complex>

namespace hsbem {
namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

using Triple = std::array<std::int64_t, 3>;

Triple reduced(Triple t) {
    std::int64_t g = 0;
    for (auto v : t)
        g = std::gcd(g, v < 0 ? -v : v);
    if (g > 1)
        for (auto &v : t)
            v /= g;
    return t;
}

std::int64_t squared(const Triple &t) { return t[0] * t[0] + t[1] * t[1] + t[2] * t[2]; }

// Runs f(i) for i < count on up to hardware_concurrency threads.
template <class F> void parallel_for(std::size_t count, F &&f) {
    const unsigned workers =
        static_cast<unsigned>(std::min<std::size_t>(count, std::max(1u, std::thread::hardware_concurrency())));
    if (workers <= 1) {
        for (std::size_t i = 0; i < count; ++i)
            f(i);
        return;
    }
    std::atomic<std::size_t> next{0};
    std::vector<std::future<void>> jobs;
    for (unsigned w = 0; w < workers; ++w)
        jobs.push_back(std::async(std::launch::async, [&] {
            for (std::size_t i = next++; i < count; i = next++)
                f(i);
        }));
    for (auto &j : jobs)
        j.get(); // rethrows the first worker exception
}

// Appends a plane-wave rule's ring table (cos theta, sin theta, sample weight) and
// its orthonormal Legendre rows P_lm(cos theta) = Y_lm(theta, 0), l <= bmax.
void append_hf_rule(const SphereRule &R, std::uint32_t bmax, std::vector<dev::HfRule> &rules, std::vector<double> &polar,
                    std::vector<double> &legendre) {
    const std::uint32_t rings = R.order + 1, na = 2 * rings;
    const std::size_t row = static_cast<std::size_t>(bmax + 1) * (bmax + 2) / 2;
    rules.push_back({legendre.size(), polar.size(), rings, na, bmax, 0});
    std::vector<double> ring(3 * rings), table(row * rings);
    for (std::uint32_t i = 0; i < rings; ++i) {
        const SphereSample &x = R.samples[static_cast<std::size_t>(i) * na];
        const double mu = x.direction.z;
        ring[3 * i] = mu;
        ring[3 * i + 1] = std::sqrt(std::max(0.0, 1 - mu * mu));
        ring[3 * i + 2] = x.weight;
    }
    parallel_for(rings, [&](std::size_t i) {
        const auto y = spherical_harmonics(bmax, {ring[3 * i + 1], 0.0, ring[3 * i]});
        for (std::uint32_t l = 0; l <= bmax; ++l)
            for (std::uint32_t m = 0; m <= l; ++m)
                table[i * row + static_cast<std::size_t>(l) * (l + 1) / 2 + m] = y[static_cast<std::size_t>(l) * l + l + m].real();
    });
    polar.insert(polar.end(), ring.begin(), ring.end());
    legendre.insert(legendre.end(), table.begin(), table.end());
}

} // namespace

std::vector<std::complex<double>> device_angular_transfer(dev::HfTransferKind kind, unsigned source_order,
                                                          unsigned target_order, unsigned bandlimit, double k,
                                                          double radius, Vec3 child_minus_parent,
                                                          const std::vector<std::complex<double>> &input,
                                                          hipStream_t s) {
    using C = std::complex<double>;
    const bool up = kind == dev::HfTransferKind::up_handoff || kind == dev::HfTransferKind::up_resample;
    const bool in_spherical = kind == dev::HfTransferKind::up_handoff,
               out_spherical = kind == dev::HfTransferKind::down_handoff;
    auto rule_samples = [](unsigned q) { return static_cast<std::size_t>(q + 1) * 2 * (q + 1); };
    if (!std::isfinite(k) || k <= 0 || !std::isfinite(radius) || radius <= 0 ||
        (in_spherical && source_order != bandlimit) || (out_spherical && target_order != bandlimit) ||
        (!in_spherical && (bandlimit > source_order || 2 * (source_order + 1) > unsigned(dev::kMaxHfAzimuths))) ||
        (!out_spherical && (bandlimit > target_order || 2 * (target_order + 1) > unsigned(dev::kMaxHfAzimuths))) ||
        bandlimit > 1023)
        throw std::invalid_argument("Invalid device angular transfer orders, wavenumber or radius.");
    const std::size_t n_in = in_spherical ? static_cast<std::size_t>(source_order + 1) * (source_order + 1)
                                          : rule_samples(source_order);
    const std::size_t n_out = out_spherical ? static_cast<std::size_t>(target_order + 1) * (target_order + 1)
                                            : rule_samples(target_order);
    if (input.size() != n_in)
        throw std::invalid_argument("Device angular transfer input has the wrong size.");
    std::vector<dev::HfRule> rules;
    std::vector<double> polar, legendre;
    dev::HfLink lk{};
    if (!in_spherical) {
        append_hf_rule(make_sphere_rule(source_order), bandlimit, rules, polar, legendre);
        lk.in_rule = 0;
    }
    if (!out_spherical) {
        append_hf_rule(make_sphere_rule(target_order), bandlimit, rules, polar, legendre);
        lk.out_rule = static_cast<std::uint32_t>(rules.size() - 1);
    }
    if (in_spherical)
        lk.in_rule = lk.out_rule;
    if (out_spherical)
        lk.out_rule = lk.in_rule;
    const double sign = up ? k : -k;
    lk.shift[0] = sign * child_minus_parent.x;
    lk.shift[1] = sign * child_minus_parent.y;
    lk.shift[2] = sign * child_minus_parent.z;
    lk.ka = k * radius;
    lk.order = bandlimit;
    lk.source = 0;
    lk.target = n_in;
    const std::uint32_t in_rings = in_spherical ? 0 : source_order + 1, out_rings = out_spherical ? 0 : target_order + 1;
    std::vector<double> pool(2 * (n_in + n_out), 0.0);
    for (std::size_t i = 0; i < n_in; ++i) {
        pool[2 * i] = input[i].real();
        pool[2 * i + 1] = input[i].imag();
    }
    DeviceBuffer<dev::HfRule> d_rules;
    DeviceBuffer<double> d_polar, d_legendre, d_pool, d_fourier, d_harmonic;
    DeviceBuffer<dev::HfLink> d_link;
    DeviceBuffer<dev::FmmDestination> d_dst;
    d_rules.upload(rules, s);
    d_polar.upload(polar, s);
    d_legendre.upload(legendre, s);
    d_pool.upload(pool, s);
    d_fourier.resize(2 * static_cast<std::size_t>(std::max(in_rings, out_rings)) * (2 * bandlimit + 1));
    d_harmonic.resize(2 * static_cast<std::size_t>(bandlimit + 1) * (bandlimit + 1));
    d_link.upload(std::vector<dev::HfLink>{lk}, s);
    d_dst.upload(std::vector<dev::FmmDestination>{{n_in, 0, 0, 1, 0}}, s);
    dev::HfTransferLaunch H{d_rules.data(), d_legendre.data(), d_polar.data(), d_link.data(), d_dst.data(),
                            d_fourier.data(), d_harmonic.data(), d_pool.data(), d_pool.data()};
    dev::launch_hf_transfer(H, kind, 1, 1, in_rings, out_rings, static_cast<std::uint32_t>(out_spherical ? 0 : n_out), s);
    hip_check(hipGetLastError(), "angular transfer launch");
    std::vector<double> back(pool.size());
    hip_check(hipMemcpyAsync(back.data(), d_pool.data(), d_pool.bytes(), hipMemcpyDeviceToHost, s), "angular download");
    hip_check(hipStreamSynchronize(s), "angular transfer sync");
    std::vector<C> out(n_out);
    for (std::size_t i = 0; i < n_out; ++i)
        out[i] = {back[2 * (n_in + i)], back[2 * (n_in + i) + 1]};
    return out;
}

namespace {

template <class T> void upload_nonempty(DeviceBuffer<T> &d, const std::vector<T> &v, hipStream_t s) {
    d.upload(v.empty() ? std::vector<T>(1) : v, s);
}

} // namespace

std::unique_ptr<FmmOperator> MeshSolver::fmm_operator(const Numerics &n, double k_max, double k, const FmmSettings &fs,
                                                      const std::vector<std::uint32_t> &faces) {
    std::lock_guard lock(mutex_);
    return build_fmm(n, k_max, k, fs, faces, ctx_->stream, nullptr);
}

std::unique_ptr<FmmOperator> MeshSolver::build_fmm(const Numerics &n, double k_max, double k, const FmmSettings &fs,
                                                   const std::vector<std::uint32_t> &faces, hipStream_t s,
                                                   const FmmOperator *share, unsigned endpoint_degree) {
    if (!std::isfinite(k) || k <= 0 || !std::isfinite(fs.accuracy_seed) || fs.accuracy_seed <= 0 ||
        fs.accuracy_seed >= 1)
        throw std::invalid_argument("FMM operator needs a positive wavenumber and an accuracy seed in (0, 1).");
    if (fs.block_width < 32 || fs.block_width > dev::kMaxJacobiWidth || (fs.block_width & (fs.block_width - 1)))
        throw std::invalid_argument("FMM preconditioner block width must be a power of two in [32, 2048].");
    auto op = std::unique_ptr<FmmOperator>(new FmmOperator());
    op->k_ = k;
    op->block_width_ = share ? share->block_width_ : fs.block_width;
    FmmStatistics &st = op->stats_;
    const std::uint32_t nv = this->nv(), nt = this->nt();
    op->n_ = nv;

    auto t0 = Clock::now();
    std::shared_ptr<const NearStructure> ns;
    if (share) {
        op->near_ = share->near_;
        op->jacobi_ = share->jacobi_;
        ns = op->near_->structure;
    } else {
        ns = near_structure(n, k_max, fs.tree, faces, s);
    }
    const SupportTree &tree = ns->tree;
    const SupportNearPlan &plan = ns->plan;
    const SupportExpansionPlan ep = make_support_expansion_plan(tree, plan.panel_face_column, k, fs.accuracy_seed);
    st.tree_ms = ms_since(t0);
    st.nodes = static_cast<std::uint32_t>(tree.nodes.size());
    st.levels = static_cast<std::uint32_t>(tree.levels.size());
    st.far_pairs = tree.pressure_pressure.far.size();

    t0 = Clock::now();
    if (!share) {
        auto dn = std::make_shared<DeviceNear>();
        integrate_near(n, k, ns, s, *dn);
        op->near_ = dn;
    }
    st.near_ms = ms_since(t0);
    st.near_nnz = op->near_->nnz;
    st.near_kernel_ms = op->near_->kernel_ms;

    // ---- levels: order, basis radius and representation of every active level
    const std::uint32_t nl = st.levels;
    std::vector<int> order(nl, -1);
    std::vector<double> radius(nl, 0);
    std::vector<char> plane(nl, 0);
    std::vector<unsigned> rule_order(nl, 0), translation_order(nl, 0);
    std::uint32_t top = nl;
    for (std::uint32_t l = 0; l < nl; ++l) {
        const WidebandLevelPlan &lp = ep.levels[l].expansion;
        if (lp.representation == LevelRepresentation::inactive)
            continue;
        if (lp.representation == LevelRepresentation::plane_wave) {
            if (2 * (lp.quadrature_order + 1) > static_cast<unsigned>(dev::kMaxHfAzimuths) ||
                lp.signature_order > lp.quadrature_order)
                throw std::runtime_error("FMM level " + std::to_string(l) + " needs a plane-wave rule of order " +
                                         std::to_string(lp.quadrature_order) + " (at most " +
                                         std::to_string(dev::kMaxHfAzimuths / 2 - 1) + ").");
            plane[l] = 1;
            rule_order[l] = lp.quadrature_order;
            translation_order[l] = lp.translation_order;
        } else if (lp.signature_order > static_cast<unsigned>(dev::kMaxBigOrder) ||
                   k * lp.active_radius > kMaximumSphericalElectricalRadius) {
            throw std::runtime_error("FMM level " + std::to_string(l) + " needs spherical order " +
                                     std::to_string(lp.signature_order) + " beyond the translations built here (at most " +
                                     std::to_string(dev::kMaxBigOrder) + ").");
        }
        order[l] = static_cast<int>(lp.signature_order);
        radius[l] = lp.active_radius;
        top = std::min(top, l);
    }
    st.top_level = top;
    for (std::uint32_t l = 0; l < nl; ++l) {
        st.level_orders.push_back(order[l] < 0 ? 0u : static_cast<std::uint32_t>(order[l]));
        st.level_representation.push_back(order[l] < 0 ? 0u : plane[l] ? 2u : 1u);
        st.level_rule_orders.push_back(rule_order[l]);
        st.plane_wave_levels += plane[l] ? 1 : 0;
    }
    for (std::uint32_t l = top; l < nl; ++l) {
        if (order[l] < 0)
            throw std::runtime_error("FMM active levels are not contiguous.");
        if (l > top && plane[l] && !plane[l - 1])
            throw std::runtime_error("FMM plan has a plane-wave level below a spherical one.");
    }
    std::vector<std::uint32_t> rule_index(nl, 0);
    std::vector<SphereRule> rules;
    for (std::uint32_t l = 0; l < nl; ++l)
        if (plane[l]) {
            rule_index[l] = static_cast<std::uint32_t>(rules.size());
            rules.push_back(make_sphere_rule(rule_order[l]));
        }
    auto samples = [&](std::uint32_t l) {
        return static_cast<std::uint64_t>(rule_order[l] + 1) * 2 * (rule_order[l] + 1);
    };
    auto coeffs = [&](std::uint32_t l) {
        return plane[l] ? samples(l) : static_cast<std::uint64_t>(order[l] + 1) * (order[l] + 1);
    };

    // ---- expansion pool: multipole and local of every node with P1 DOFs on an active level
    const std::uint64_t kNone = UINT64_MAX;
    std::vector<std::uint64_t> mult(tree.nodes.size(), kNone), loc(tree.nodes.size(), kNone);
    std::uint64_t pool = 0;
    for (std::uint32_t id = 0; id < tree.nodes.size(); ++id) {
        const SupportNode &nd = tree.nodes[id];
        if (!nd.pressure_count || order[nd.level] < 0)
            continue;
        mult[id] = pool;
        pool += coeffs(nd.level);
        loc[id] = pool;
        pool += coeffs(nd.level);
    }
    st.expansion_values = pool;

    // ---- leaves, their DOFs and (leaf, support panel) items; leaves on plane-wave
    // levels emit and test their signatures directly
    std::vector<dev::FmmLeaf> leaves;
    std::vector<std::uint32_t> leaf_dofs;
    std::vector<dev::FmmLeafItem> items;
    std::vector<dev::HfLeaf> hf_leaves;
    std::vector<dev::HfLeafItem> hf_items;
    std::vector<double> leaf_scale;
    std::uint32_t hf_max_samples = 0;
    std::vector<std::int32_t> local(nv, -1);
    std::uint64_t moments = 0;
    for (std::uint32_t id = 0; id < tree.nodes.size(); ++id) {
        const SupportNode &nd = tree.nodes[id];
        if (!nd.leaf() || !nd.pressure_count || order[nd.level] < 0)
            continue;
        const bool hf = plane[nd.level] != 0;
        if (nd.pressure_count > static_cast<std::uint32_t>(dev::kMaxLeafDofs) ||
            (!hf && order[nd.level] > dev::kMaxBigOrder))
            throw std::runtime_error("FMM leaf exceeds the supported DOF count or order.");
        const auto first_dof = static_cast<std::uint32_t>(leaf_dofs.size());
        for (std::uint32_t j = 0; j < nd.pressure_count; ++j) {
            const std::uint32_t v = tree.pressure_order[nd.first_pressure + j];
            local[v] = static_cast<std::int32_t>(j);
            leaf_dofs.push_back(v);
            leaf_scale.push_back(1.0 / space_.lumped_mass[v]);
        }
        auto corners = [&](std::uint32_t t, std::int32_t out[3]) {
            for (int c = 0; c < 3; ++c) {
                const std::uint32_t v = space_.panels[t].nodes[c];
                out[c] = tree.pressure_leaf[v] == id ? local[v] : -1;
            }
        };
        if (hf) {
            dev::HfLeaf hl{};
            hl.center[0] = nd.center.x;
            hl.center[1] = nd.center.y;
            hl.center[2] = nd.center.z;
            hl.multipole = mult[id];
            hl.local = loc[id];
            hl.rule = rule_index[nd.level];
            hl.first_item = static_cast<std::uint32_t>(hf_items.size());
            hl.first_dof = first_dof;
            hl.dof_count = nd.pressure_count;
            const auto leaf_index = static_cast<std::uint32_t>(hf_leaves.size());
            for (std::uint32_t q = plan.node_panel_offsets[id]; q < plan.node_panel_offsets[id + 1]; ++q) {
                const std::uint32_t t = plan.support_panels[q];
                const Panel &pn = space_.panels[t];
                dev::HfLeafItem it{};
                for (int x = 0; x < 3; ++x) {
                    const double v0 = x == 0 ? pn.vertices[0].x : x == 1 ? pn.vertices[0].y : pn.vertices[0].z;
                    const double v1 = x == 0 ? pn.vertices[1].x : x == 1 ? pn.vertices[1].y : pn.vertices[1].z;
                    const double v2 = x == 0 ? pn.vertices[2].x : x == 1 ? pn.vertices[2].y : pn.vertices[2].z;
                    it.v0[x] = v0;
                    it.e1[x] = v1 - v0;
                    it.e2[x] = v2 - v0;
                }
                it.normal[0] = pn.normal.x;
                it.normal[1] = pn.normal.y;
                it.normal[2] = pn.normal.z;
                it.area2 = 2 * pn.area;
                corners(t, it.corner);
                it.leaf = leaf_index;
                hf_items.push_back(it);
            }
            hl.item_count = static_cast<std::uint32_t>(hf_items.size()) - hl.first_item;
            hf_max_samples = std::max(hf_max_samples, static_cast<std::uint32_t>(samples(nd.level)));
            hf_leaves.push_back(hl);
            continue;
        }
        dev::FmmLeaf lf{};
        lf.center[0] = nd.center.x;
        lf.center[1] = nd.center.y;
        lf.center[2] = nd.center.z;
        lf.radius = radius[nd.level];
        lf.order = static_cast<std::uint32_t>(order[nd.level]);
        lf.first_dof = first_dof;
        lf.dof_count = nd.pressure_count;
        lf.emission = moments;
        lf.multipole = mult[id];
        lf.local = loc[id];
        moments += static_cast<std::uint64_t>(nd.pressure_count) * coeffs(nd.level);
        const auto leaf_index = static_cast<std::uint32_t>(leaves.size());
        for (std::uint32_t q = plan.node_panel_offsets[id]; q < plan.node_panel_offsets[id + 1]; ++q) {
            const std::uint32_t t = plan.support_panels[q];
            dev::FmmLeafItem it{leaf_index, t, {-1, -1, -1}, 0};
            corners(t, it.corner);
            items.push_back(it);
        }
        leaves.push_back(lf);
    }
    op->leaf_count_ = static_cast<std::uint32_t>(leaves.size());
    st.leaves = op->leaf_count_ + static_cast<std::uint32_t>(hf_leaves.size());
    st.plane_wave_leaves = hf_leaves.size();
    st.moment_values = moments;

    // ---- translation links
    t0 = Clock::now();
    // one rotation plan per (direction, order class): a launch's kernels are chosen by
    // its largest order (<= 32, <= 40, <= 48, <= 64, <= 128), so narrower levels never
    // share a plan with a wider one
    std::map<std::pair<Triple, int>, std::uint32_t> direction_index;
    std::vector<Triple> directions;
    std::vector<int> direction_order;
    auto direction = [&](Triple t, int ord) {
        t = reduced(t);
        const int order_class = ord > dev::kMaxWideOrder    ? 4
                                : ord > dev::kMidFmmOrder   ? 3
                                : ord > dev::kMaxFmmOrder   ? 2
                                : ord > dev::kSmallFmmOrder ? 1
                                                            : 0;
        auto [it, fresh] = direction_index.emplace(std::make_pair(t, order_class),
                                                   static_cast<std::uint32_t>(directions.size()));
        if (fresh) {
            directions.push_back(t);
            direction_order.push_back(ord);
        } else {
            direction_order[it->second] = std::max(direction_order[it->second], ord);
        }
        return it->second;
    };
    struct AxialKey {
        int kind;
        std::uint32_t source_level, target_level;
        std::int64_t length2;
        auto operator<=>(const AxialKey &) const = default;
    };
    std::map<AxialKey, std::uint32_t> axial_index;
    std::vector<AxialKey> axial_keys;
    auto axial = [&](AxialKey key) {
        auto [it, fresh] = axial_index.emplace(key, static_cast<std::uint32_t>(axial_keys.size()));
        if (fresh)
            axial_keys.push_back(key);
        return it->second;
    };
    auto half = [&](std::uint32_t l) { return tree.nodes[tree.levels[l].nodes.front()].half_width; };
    for (std::uint32_t l = 0; l < nl; ++l)
        for (std::uint32_t id : tree.levels[l].nodes)
            if (std::abs(tree.nodes[id].half_width - tree.nodes[tree.levels[l].nodes.front()].half_width) >
                1e-12 * tree.nodes[id].half_width)
                throw std::runtime_error("FMM level cells are not uniform.");
    // child - parent in units of half the child's cell width
    auto child_offset = [&](const SupportNode &c, const SupportNode &p) {
        Triple t;
        for (int a = 0; a < 3; ++a)
            t[a] = 2 * static_cast<std::int64_t>(c.coords[a]) + 1 - 2 * (2 * static_cast<std::int64_t>(p.coords[a]) + 1);
        return t;
    };
    // Links of one launch are stored in destination order with launch-relative
    // indices; each gets a workspace slot for the larger of its two orders.
    std::vector<dev::FmmDestination> destinations;
    std::vector<dev::FmmLink> links;
    std::vector<std::uint32_t> link_order, axial_order;
    std::vector<dev::FmmChunk> chunks, axial_chunks;
    std::uint32_t launch_link_base = 0;
    std::uint64_t slot = 0, workspace = 0;
    auto begin_destination = [&](std::uint64_t target, int ord) {
        destinations.push_back({target, static_cast<std::uint32_t>(ord),
                                static_cast<std::uint32_t>(links.size()) - launch_link_base, 0, 0});
    };
    auto add_link = [&](std::uint64_t source, int source_order, std::uint32_t rot, std::uint32_t ax) {
        const int target_order = static_cast<int>(destinations.back().order);
        const int big = std::max(source_order, target_order);
        links.push_back({source, slot, static_cast<std::uint32_t>(source_order), static_cast<std::uint32_t>(target_order),
                         rot, ax});
        slot += static_cast<std::uint64_t>(big + 1) * (big + 1);
        ++destinations.back().link_count;
    };
    // Chunks of at most kFmmChunk launch links sharing key(link), in key order;
    // chunk.first indexes the launch's own slice of the order array.
    auto make_chunks = [&](std::uint32_t count, auto key, std::vector<std::uint32_t> &order_out,
                           std::vector<dev::FmmChunk> &chunks_out) {
        std::vector<std::uint32_t> idx(count);
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](std::uint32_t a, std::uint32_t b) {
            return key(links[launch_link_base + a]) < key(links[launch_link_base + b]);
        });
        for (std::uint32_t i = 0; i < idx.size();) {
            const std::uint32_t kv = key(links[launch_link_base + idx[i]]);
            dev::FmmChunk ch{kv, i, 0, 0};
            while (i < idx.size() && ch.count < static_cast<std::uint32_t>(dev::kFmmChunk) &&
                   key(links[launch_link_base + idx[i]]) == kv) {
                const auto &lk = links[launch_link_base + idx[i]];
                ch.max_order = std::max({ch.max_order, lk.source_order, lk.target_order});
                ++ch.count;
                ++i;
            }
            chunks_out.push_back(ch);
        }
        order_out.insert(order_out.end(), idx.begin(), idx.end());
    };
    auto end_launch = [&](std::uint32_t first) {
        FmmOperator::Launch L{first, static_cast<std::uint32_t>(destinations.size()) - first, launch_link_base,
                              static_cast<std::uint32_t>(links.size()) - launch_link_base,
                              static_cast<std::uint32_t>(chunks.size()), 0,
                              static_cast<std::uint32_t>(axial_chunks.size()), 0};
        make_chunks(L.link_count, [](const dev::FmmLink &lk) { return lk.rotation; }, link_order, chunks);
        make_chunks(L.link_count, [](const dev::FmmLink &lk) { return lk.axial; }, axial_order, axial_chunks);
        L.chunk_count = static_cast<std::uint32_t>(chunks.size()) - L.first_chunk;
        L.axial_chunk_count = static_cast<std::uint32_t>(axial_chunks.size()) - L.first_axial_chunk;
        workspace = std::max(workspace, slot);
        slot = 0;
        launch_link_base = static_cast<std::uint32_t>(links.size());
        if (L.destination_count) {
            op->stages_.push_back({FmmOperator::StageKind::translate, static_cast<std::uint32_t>(op->translate_.size())});
            op->translate_.push_back(L);
        }
    };
    // plane-wave transfers: links in destination order, launch-relative indices and
    // workspace offsets
    std::vector<dev::HfLink> hf_links;
    std::vector<dev::FmmDestination> hf_destinations;
    std::uint64_t hf_fourier = 0, hf_harmonic = 0, hf_fourier_size = 0, hf_harmonic_size = 0;
    std::vector<std::uint32_t> rule_bandlimit(rules.size(), 0);
    FmmOperator::HfLaunch hl{};
    auto hf_begin = [&](dev::HfTransferKind kind) {
        hl = {};
        hl.kind = kind;
        hl.first_link = static_cast<std::uint32_t>(hf_links.size());
        hl.first_destination = static_cast<std::uint32_t>(hf_destinations.size());
        hf_fourier = hf_harmonic = 0;
    };
    auto hf_destination = [&](std::uint64_t target) {
        hf_destinations.push_back({target, 0, static_cast<std::uint32_t>(hf_links.size()) - hl.first_link, 0, 0});
    };
    // source: the child (upward) or the parent (downward) node; spherical levels pass no rule
    auto hf_link = [&](std::uint64_t source, std::uint64_t target, const SupportNode &child, const SupportNode &parent,
                       std::uint32_t child_level) {
        const bool up = hl.kind == dev::HfTransferKind::up_handoff || hl.kind == dev::HfTransferKind::up_resample;
        const std::uint32_t parent_level = child_level - 1;
        dev::HfLink lk{};
        lk.source = source;
        lk.target = target;
        const double sign = up ? k : -k;
        lk.shift[0] = sign * (child.center.x - parent.center.x);
        lk.shift[1] = sign * (child.center.y - parent.center.y);
        lk.shift[2] = sign * (child.center.z - parent.center.z);
        lk.ka = k * radius[child_level];
        lk.order = static_cast<std::uint32_t>(plane[child_level] ? std::min(order[child_level], order[parent_level])
                                                                 : order[child_level]);
        const std::uint32_t pr = rule_index[parent_level], cr = plane[child_level] ? rule_index[child_level] : pr;
        lk.in_rule = up ? cr : pr;
        lk.out_rule = up ? pr : cr;
        const std::uint64_t in_rings = hl.kind == dev::HfTransferKind::up_handoff ? 0 : rules[lk.in_rule].order + 1;
        const std::uint64_t out_rings = hl.kind == dev::HfTransferKind::down_handoff ? 0 : rules[lk.out_rule].order + 1;
        hl.in_rings = std::max(hl.in_rings, static_cast<std::uint32_t>(in_rings));
        hl.out_rings = std::max(hl.out_rings, static_cast<std::uint32_t>(out_rings));
        if (out_rings)
            hl.out_samples = std::max(hl.out_samples, static_cast<std::uint32_t>(samples(up ? parent_level : child_level)));
        lk.fourier = hf_fourier;
        lk.harmonic = hf_harmonic;
        hf_fourier += std::max(in_rings, out_rings) * (2 * lk.order + 1);
        hf_harmonic += static_cast<std::uint64_t>(lk.order + 1) * (lk.order + 1);
        rule_bandlimit[pr] = std::max(rule_bandlimit[pr], lk.order);
        rule_bandlimit[cr] = std::max(rule_bandlimit[cr], lk.order);
        hf_links.push_back(lk);
        ++hf_destinations.back().link_count;
        ++st.angular_links;
    };
    auto hf_end = [&] {
        hl.link_count = static_cast<std::uint32_t>(hf_links.size()) - hl.first_link;
        hl.destination_count = static_cast<std::uint32_t>(hf_destinations.size()) - hl.first_destination;
        if (hl.link_count > 65535 || hl.destination_count > 65535)
            throw std::runtime_error("FMM plane-wave level exceeds 65535 transfer links.");
        hf_fourier_size = std::max(hf_fourier_size, hf_fourier);
        hf_harmonic_size = std::max(hf_harmonic_size, hf_harmonic);
        if (hl.link_count) {
            op->stages_.push_back({FmmOperator::StageKind::angular, static_cast<std::uint32_t>(op->angular_.size())});
            op->angular_.push_back(hl);
        }
    };
    // plane-wave M2L: one diagonal translation per (level, cell offset)
    std::vector<dev::HfDiagonalLink> diag_links;
    std::vector<dev::FmmDestination> diag_destinations;
    std::map<std::pair<std::uint32_t, Triple>, std::uint64_t> diag_index;
    std::vector<std::pair<std::uint32_t, Triple>> diag_keys;
    std::uint64_t diag_values = 0;
    // upward: parents at level l from children at l + 1
    for (std::uint32_t l = nl - 1; l-- > top;) {
        if (plane[l]) {
            hf_begin(plane[l + 1] ? dev::HfTransferKind::up_resample : dev::HfTransferKind::up_handoff);
            for (std::uint32_t id : tree.levels[l].nodes) {
                const SupportNode &p = tree.nodes[id];
                if (p.leaf() || mult[id] == kNone ||
                    std::none_of(p.children.begin(), p.children.end(),
                                 [&](std::uint32_t c) { return c != kNoSupportNode && mult[c] != kNone; }))
                    continue; // no source below: the signature keeps its zero fill
                hf_destination(mult[id]);
                for (std::uint32_t c : p.children)
                    if (c != kNoSupportNode && mult[c] != kNone) {
                        hf_link(mult[c], 0, tree.nodes[c], p, l + 1);
                        ++st.m2m_links;
                    }
            }
            hf_end();
            continue;
        }
        const std::uint32_t first = static_cast<std::uint32_t>(destinations.size());
        for (std::uint32_t id : tree.levels[l].nodes) {
            const SupportNode &p = tree.nodes[id];
            if (p.leaf() || mult[id] == kNone)
                continue;
            begin_destination(mult[id], order[l]);
            for (std::uint32_t c : p.children) {
                if (c == kNoSupportNode || mult[c] == kNone)
                    continue;
                const Triple off = child_offset(tree.nodes[c], p);
                add_link(mult[c], order[l + 1], direction(off, std::max(order[l], order[l + 1])),
                         axial({static_cast<int>(AxialKind::multipole_to_parent), l + 1, l, squared(off)}));
                ++st.m2m_links;
            }
        }
        end_launch(first);
    }
    // downward: M2L from the far sources and L2L from the parent
    std::vector<std::vector<std::uint32_t>> far_sources(tree.nodes.size());
    for (const SupportPair &pr : tree.pressure_pressure.far)
        far_sources[pr.target].push_back(pr.source);
    auto cell_offset = [&](const SupportNode &t, const SupportNode &sn) {
        Triple off;
        for (int a = 0; a < 3; ++a)
            off[a] = static_cast<std::int64_t>(t.coords[a]) - static_cast<std::int64_t>(sn.coords[a]);
        return off;
    };
    for (std::uint32_t l = top; l < nl; ++l) {
        const bool parent_plane = l > top && plane[l - 1];
        if (plane[l]) {
            FmmOperator::DiagonalLaunch dl{static_cast<std::uint32_t>(diag_destinations.size()), 0,
                                           static_cast<std::uint32_t>(diag_links.size()), samples(l)};
            for (std::uint32_t id : tree.levels[l].nodes) {
                if (loc[id] == kNone)
                    continue;
                const SupportNode &t = tree.nodes[id];
                diag_destinations.push_back(
                    {loc[id], 0, static_cast<std::uint32_t>(diag_links.size()) - dl.first_link, 0, 0});
                for (std::uint32_t sid : far_sources[id]) {
                    if (mult[sid] == kNone)
                        continue;
                    const auto key = std::make_pair(l, cell_offset(t, tree.nodes[sid]));
                    auto [it, fresh] = diag_index.emplace(key, diag_values);
                    if (fresh) {
                        diag_keys.push_back(key);
                        diag_values += samples(l);
                    }
                    diag_links.push_back({mult[sid], it->second});
                    ++diag_destinations.back().link_count;
                    ++st.m2l_links;
                }
            }
            dl.destination_count = static_cast<std::uint32_t>(diag_destinations.size()) - dl.first_destination;
            if (dl.destination_count > 65535)
                throw std::runtime_error("FMM plane-wave level exceeds 65535 destinations.");
            if (dl.destination_count) {
                op->stages_.push_back({FmmOperator::StageKind::diagonal, static_cast<std::uint32_t>(op->diagonal_.size())});
                op->diagonal_.push_back(dl);
            }
        } else {
            const std::uint32_t first = static_cast<std::uint32_t>(destinations.size());
            for (std::uint32_t id : tree.levels[l].nodes) {
                if (loc[id] == kNone)
                    continue;
                const SupportNode &t = tree.nodes[id];
                begin_destination(loc[id], order[l]);
                for (std::uint32_t sid : far_sources[id]) {
                    if (mult[sid] == kNone)
                        continue;
                    const Triple off = cell_offset(t, tree.nodes[sid]);
                    add_link(mult[sid], order[l], direction(off, order[l]),
                             axial({static_cast<int>(AxialKind::multipole_to_local), l, l, squared(off)}));
                    ++st.m2l_links;
                }
                if (l > top && !parent_plane && t.parent != kNoSupportNode && loc[t.parent] != kNone) {
                    const Triple off = child_offset(t, tree.nodes[t.parent]);
                    add_link(loc[t.parent], order[l - 1], direction(off, std::max(order[l], order[l - 1])),
                             axial({static_cast<int>(AxialKind::local_to_child), l - 1, l, squared(off)}));
                    ++st.l2l_links;
                }
            }
            end_launch(first);
        }
        if (parent_plane) {
            // added after the level's own M2L sums
            hf_begin(plane[l] ? dev::HfTransferKind::down_resample : dev::HfTransferKind::down_handoff);
            for (std::uint32_t id : tree.levels[l].nodes) {
                const SupportNode &t = tree.nodes[id];
                if (loc[id] == kNone || t.parent == kNoSupportNode || loc[t.parent] == kNone)
                    continue;
                hf_destination(loc[id]);
                hf_link(loc[t.parent], loc[id], t, tree.nodes[t.parent], l);
                ++st.l2l_links;
            }
            hf_end();
        }
    }

    // the kernels of a launch are chosen by the largest order it touches: its links'
    // orders and the orders of the (level-shared) rotation plans it applies
    for (FmmOperator::Launch &L : op->translate_)
        for (std::uint32_t c = L.first_chunk; c < L.first_chunk + L.chunk_count; ++c)
            L.max_order = std::max({L.max_order, chunks[c].max_order,
                                    static_cast<std::uint32_t>(direction_order[chunks[c].rotation])});

    // ---- rotation and coaxial plans (host, in parallel)
    const TranslationPlanCacheStatistics cache_before = translation_plans_.statistics();
    // Wigner blocks per (order, polar class) from the mesh's cache, phases per direction
    std::vector<std::shared_ptr<const RotationPlan>> rotations(directions.size());
    std::vector<std::vector<std::complex<double>>> phases(directions.size());
    parallel_for(directions.size(), [&](std::size_t i) {
        const Triple &d = directions[i];
        const Vec3 axis{static_cast<double>(d[0]), static_cast<double>(d[1]), static_cast<double>(d[2])};
        const auto ord = static_cast<unsigned>(direction_order[i]);
        rotations[i] = translation_plans_.wigner(ord, rotation_polar_angle(axis));
        phases[i] = rotation_azimuth(ord, axis);
    });
    std::vector<std::shared_ptr<const AxialPlan>> plans(axial_keys.size());
    parallel_for(axial_keys.size(), [&](std::size_t i) {
        const AxialKey &key = axial_keys[i];
        const int ps = order[key.source_level], pt = order[key.target_level];
        // M2L offsets count whole cells (2 half widths) of the level, parent links
        // half widths of the child's level
        const double unit = key.kind == static_cast<int>(AxialKind::multipole_to_local)
                                ? 2 * half(key.source_level)
                                : half(std::max(key.source_level, key.target_level));
        const double distance = std::sqrt(static_cast<double>(key.length2)) * unit;
        plans[i] = translation_plans_.axial({static_cast<AxialKind>(key.kind), static_cast<unsigned>(ps),
                                             static_cast<unsigned>(pt), 2 * static_cast<unsigned>(std::max(ps, pt)) + 9,
                                             k, radius[key.source_level], radius[key.target_level], distance, 0.0});
    });
    {
        const TranslationPlanCacheStatistics after = translation_plans_.statistics();
        st.plan_cache_hits = after.hits - cache_before.hits;
        st.plans_built = after.plans_built - cache_before.plans_built;
    }
    st.rotations = rotations.size();
    st.axial_plans = plans.size();
    std::vector<dev::FmmRotation> rot_table;
    std::vector<double> small_d, azimuth;
    std::map<const RotationPlan *, std::uint64_t> block_offset; // one copy per polar class
    for (std::size_t i = 0; i < rotations.size(); ++i) {
        const RotationPlan &r = *rotations[i];
        auto [it, fresh] = block_offset.emplace(&r, small_d.size());
        if (fresh)
            small_d.insert(small_d.end(), r.small_d.begin(), r.small_d.end());
        rot_table.push_back({it->second, azimuth.size() / 2, r.order, 0});
        for (const auto &z : phases[i]) {
            azimuth.push_back(z.real());
            azimuth.push_back(z.imag());
        }
    }
    std::vector<std::uint64_t> axial_offsets;
    std::vector<double> axial_values;
    for (const auto &pp : plans) {
        axial_offsets.push_back(axial_values.size() / 2);
        for (const auto &z : pp->coefficients) {
            axial_values.push_back(z.real());
            axial_values.push_back(z.imag());
        }
    }
    // plane-wave rules (ring cosines, weights, Legendre rows up to each rule's largest
    // bandlimit) and the diagonal translations of every (level, cell offset)
    std::vector<dev::HfRule> hf_rules;
    std::vector<double> hf_polar, hf_legendre;
    for (std::size_t r = 0; r < rules.size(); ++r)
        append_hf_rule(rules[r], rule_bandlimit[r], hf_rules, hf_polar, hf_legendre);
    std::vector<double> diag_table(2 * diag_values);
    parallel_for(diag_keys.size(), [&](std::size_t i) {
        const auto &[l, off] = diag_keys[i];
        const double unit = 2 * half(l);
        const auto T = diagonal_translation(rules[rule_index[l]], translation_order[l], k,
                                            {off[0] * unit, off[1] * unit, off[2] * unit});
        const std::uint64_t base = diag_index.at(diag_keys[i]);
        for (std::size_t q = 0; q < T.size(); ++q) {
            diag_table[2 * (base + q)] = T[q].real();
            diag_table[2 * (base + q) + 1] = T[q].imag();
        }
    });
    st.diagonal_classes = diag_keys.size();
    st.plans_ms = ms_since(t0);

    // ---- device data
    std::vector<double> rule;
    const double max_kh = [&] {
        double h = 0;
        for (const Panel &p : space_.panels)
            h = std::max(h, p.max_edge);
        return k * h;
    }();
    const unsigned degree = std::max(minimum_endpoint_degree(max_kh, n.quadrature_order),
                                     endpoint_degree ? endpoint_degree : fs.endpoint_degree);
    if (degree > 14)
        throw std::runtime_error("fast multipole panel-rule comparison above degree 14 is not available.");
    st.endpoint_degree = degree;
    for (const auto &q : triangle_rule(degree)) {
        rule.insert(rule.end(), {q.bary[0], q.bary[1], q.bary[2], q.weight});
    }
    st.rule_points = static_cast<std::uint32_t>(rule.size() / 4);
    upload_nonempty(op->leaves_, leaves, s);
    upload_nonempty(op->leaf_dofs_, leaf_dofs, s);
    op->order_.upload(tree.pressure_order, s);
    DeviceBuffer<dev::FmmLeafItem> d_items;
    upload_nonempty(d_items, items, s);
    DeviceBuffer<double> &d_rule = op->leaf_rule_;
    d_rule.upload(rule, s);
    op->hf_leaf_count_ = static_cast<std::uint32_t>(hf_leaves.size());
    op->hf_leaf_item_count_ = static_cast<std::uint32_t>(hf_items.size());
    op->hf_max_samples_ = hf_max_samples;
    upload_nonempty(op->hf_leaves_, hf_leaves, s);
    upload_nonempty(op->hf_leaf_items_, hf_items, s);
    upload_nonempty(op->leaf_scale_, leaf_scale, s);
    op->hf_field_.resize(2 * std::max<std::size_t>(hf_items.size() * st.rule_points, 1));
    op->emission_.resize(2 * std::max<std::uint64_t>(moments, 1));
    op->testing_.resize(2 * std::max<std::uint64_t>(moments, 1));
    op->pool_.resize(2 * std::max<std::uint64_t>(pool, 1));
    hip_check(hipMemsetAsync(op->emission_.data(), 0, op->emission_.bytes(), s), "clear emission");
    hip_check(hipMemsetAsync(op->testing_.data(), 0, op->testing_.bytes(), s), "clear testing");
    hip_check(hipMemsetAsync(op->pool_.data(), 0, op->pool_.bytes(), s), "clear expansions");
    upload_nonempty(op->destinations_, destinations, s);
    upload_nonempty(op->links_, links, s);
    upload_nonempty(op->chunks_, chunks, s);
    upload_nonempty(op->link_order_, link_order, s);
    upload_nonempty(op->axial_chunks_, axial_chunks, s);
    upload_nonempty(op->axial_order_, axial_order, s);
    op->work_y_.resize(2 * std::max<std::uint64_t>(workspace, 1));
    op->work_z_.resize(2 * std::max<std::uint64_t>(workspace, 1));
    upload_nonempty(op->rotations_, rot_table, s);
    upload_nonempty(op->small_d_, small_d, s);
    upload_nonempty(op->azimuth_, azimuth, s);
    upload_nonempty(op->axial_, axial_values, s);
    upload_nonempty(op->axial_offsets_, axial_offsets, s);
    upload_nonempty(op->hf_rules_, hf_rules, s);
    upload_nonempty(op->hf_polar_, hf_polar, s);
    upload_nonempty(op->hf_legendre_, hf_legendre, s);
    upload_nonempty(op->hf_links_, hf_links, s);
    upload_nonempty(op->hf_destinations_, hf_destinations, s);
    upload_nonempty(op->diag_links_, diag_links, s);
    upload_nonempty(op->diag_destinations_, diag_destinations, s);
    upload_nonempty(op->diag_table_, diag_table, s);
    op->hf_fourier_.resize(2 * std::max<std::uint64_t>(hf_fourier_size, 1));
    op->hf_harmonic_.resize(2 * std::max<std::uint64_t>(hf_harmonic_size, 1));
    DeviceBuffer<int> d_err(1);
    hip_check(hipMemsetAsync(d_err.data(), 0, sizeof(int), s), "clear error");

    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event");
    dev::FmmLeafMomentsLaunch M{};
    M.geo = geo_.data();
    M.nt = nt;
    M.rule = d_rule.data();
    M.rule_points = st.rule_points;
    M.k = k;
    M.alpha_re = 0;
    M.alpha_im = -1 / k;
    M.row_scale = original_row_scale_.data();
    M.leaves = op->leaves_.data();
    M.leaf_dofs = op->leaf_dofs_.data();
    M.items = d_items.data();
    M.item_count = static_cast<std::uint32_t>(items.size());
    for (const dev::FmmLeaf &lf : leaves)
        M.max_order = std::max(M.max_order, lf.order);
    M.emission = op->emission_.data();
    M.testing = op->testing_.data();
    M.error = d_err.data();
    dev::launch_leaf_moments(M, s);
    hip_check(hipGetLastError(), "leaf moments launch");

    // near block-Jacobi in tree order
    DeviceBuffer<std::uint32_t> d_pos;
    DeviceBuffer<double> blocks;
    if (!share) {
        const std::uint32_t W = fs.block_width, nb = (nv + W - 1) / W;
        std::vector<std::uint32_t> position(nv);
        for (std::uint32_t i = 0; i < nv; ++i)
            position[tree.pressure_order[i]] = i;
        d_pos.upload(position, s);
        blocks.resize(2 * static_cast<std::size_t>(nb) * W * W);
        hip_check(hipMemsetAsync(blocks.data(), 0, blocks.bytes(), s), "clear blocks");
        dev::launch_gather_blocks(ns->row_offsets.data(), ns->columns.data(), op->near_->values.data(),
                                  op->order_.data(), d_pos.data(), nv, W, blocks.data(), s);
        std::shared_ptr<DeviceBuffer<double>> jac;
        if (W == static_cast<std::uint32_t>(dev::kFmmJacobi)) {
            jac = std::make_shared<DeviceBuffer<double>>(dev::jacobi_factor_doubles(nv));
            dev::jacobi_factor_gathered(blocks.data(), nv, jac->data(), s);
        } else {
            jac = std::make_shared<DeviceBuffer<double>>(dev::jacobi_wide_doubles(nv, W));
            DeviceBuffer<std::int32_t> pivots(static_cast<std::size_t>(nb) * W);
            dev::jacobi_factor_wide(blocks.data(), nv, W, jac->data(), pivots.data(), s);
            hip_check(hipStreamSynchronize(s), "block-Jacobi factor");
        }
        op->jacobi_ = jac;
    }
    hip_check(hipEventRecord(e1, s), "event");
    hip_check(hipGetLastError(), "jacobi launch");
    int code = 0;
    hip_check(hipMemcpyAsync(&code, d_err.data(), sizeof(int), hipMemcpyDeviceToHost, s), "error download");
    hip_check(hipStreamSynchronize(s), "fmm build sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.moments_ms = ms;
    if (code)
        throw std::runtime_error("FMM leaf moments left the bounded radial-table domain.");
    // direct audit data (small meshes) and the work model deciding between the audits
    op->geo_ = geo_.data();
    op->row_scale_ = original_row_scale_.data();
    op->nodes_ = original_nodes_.data();
    op->nt_ = nt;
    {
        const long double panels = nt, q = st.rule_points, ratio = 2; // FP32 : FP64 throughput of MI355X
        op->direct_work_ = static_cast<double>(panels * panels * (9.L + q * q * (13.L + 1.L / ratio)));
        long double multipole = 0;
        for (std::uint32_t l = 0; l < nl; ++l) {
            if (order[l] < 0)
                continue;
            const long double pairs = static_cast<long double>(tree.levels[l].pp_far) + tree.levels[l].pf_far;
            const long double p = order[l], n = p + 1;
            if (!plane[l]) {
                const long double rotation = n * (4 * p * p + 8 * p + 3) / 3, axial_work = n * n + p * n * (2 * p + 1) / 3;
                multipole += pairs * (2 * rotation + axial_work) +
                             static_cast<long double>(tree.levels[l].squared_offsets.size()) * n * n * n * n;
            } else {
                multipole += pairs * 2.L * (rule_order[l] + 1.L) * (rule_order[l] + 1.L);
            }
        }
        op->multipole_work_ = static_cast<double>(multipole);
    }
    if (nt <= kMaximumDirectAuditPanels) {
        std::vector<std::uint32_t> dense(tree.nodes.size(), UINT32_MAX), vertex_leaf(nv);
        std::uint32_t count = 0;
        for (std::uint32_t v = 0; v < nv; ++v) {
            const std::uint32_t leaf = tree.pressure_leaf[v];
            if (dense[leaf] == UINT32_MAX)
                dense[leaf] = count++;
            vertex_leaf[v] = dense[leaf];
        }
        std::vector<std::uint8_t> near(static_cast<std::size_t>(count) * count, 0);
        for (const SupportPair &pr : tree.pressure_pressure.near)
            if (dense[pr.target] != UINT32_MAX && dense[pr.source] != UINT32_MAX)
                near[static_cast<std::size_t>(dense[pr.target]) * count + dense[pr.source]] = 1;
        op->vertex_leaf_.upload(vertex_leaf, s);
        op->near_leaves_.upload(near, s);
        op->direct_leaf_count_ = count;
        hip_check(hipStreamSynchronize(s), "direct audit upload");
    }
    st.device_bytes = op->near_->values.bytes() + ns->columns.bytes() + op->emission_.bytes() +
                      op->testing_.bytes() + op->pool_.bytes() + op->small_d_.bytes() + op->azimuth_.bytes() +
                      op->axial_.bytes() + op->jacobi_->bytes() + op->links_.bytes() + op->destinations_.bytes() +
                      op->work_y_.bytes() + op->work_z_.bytes() + op->hf_legendre_.bytes() + op->diag_table_.bytes() +
                      op->hf_fourier_.bytes() + op->hf_harmonic_.bytes();
    st.shared_bytes = op->near_->values.bytes() + ns->columns.bytes() + op->jacobi_->bytes();
    // Stage-split target of the translations, chosen by the autotuner per (mesh, number
    // of translation launches: the tree's spherical levels) from whole applies on a zero
    // vector; operators of other orders on the same tree (audits, other frequencies)
    // take the stored choice. HSBEM_FMM_SPLIT pins it (0 = no split).
    if (!op->translate_.empty()) {
        std::ostringstream key;
        key << "mesh " << std::hex << hash_ << std::dec << " (" << nt << " panels), " << op->translate_.size()
            << " translation launches";
        const auto pinned = autotune::pinned_from_env("HSBEM_FMM_SPLIT", {0, 1024, 2048, 4096});
        op->graph_x_.resize(2 * static_cast<std::size_t>(nv));
        op->graph_y_.resize(2 * static_cast<std::size_t>(nv));
        hip_check(hipMemsetAsync(op->graph_x_.data(), 0, op->graph_x_.bytes(), s), "tuning input");
        FmmOperator &o = *op;
        o.split_target_ = autotune::choose(
            "fmm_split_target", key.str(), {0, 1024, 2048, 4096}, 2048, s,
            [&](std::uint32_t c) {
                o.split_target_ = c;
                o.apply(o.graph_x_.data(), o.graph_y_.data(), s);
            },
            1, pinned);
        hip_check(hipGetLastError(), "translation tuning");
    }
    return op;
}

void FmmOperator::apply(const double *x, double *y, hipStream_t s) const { apply_in(x, y, s, false); }

bool FmmOperator::fp32_available() const {
    if (!angular_.empty() || !diagonal_.empty() || hf_leaf_count_)
        return false;
    for (const Launch &L : translate_)
        if (L.link_count && L.max_order > static_cast<std::uint32_t>(dev::kMidFmmOrder))
            return false;
    return true;
}

void FmmOperator::apply_fp32(const double *x, double *y, hipStream_t s) const {
    if (!fp32_available())
        throw std::logic_error("the FP32-bulk apply needs spherical levels of order at most 48");
    if (!pool32_.size()) {
        pool32_.resize(pool_.size());
        work_y32_.resize(work_y_.size());
        work_z32_.resize(work_z_.size());
        hip_check(hipMemsetAsync(pool32_.data(), 0, pool32_.bytes(), s), "clear FP32 expansions");
    }
    apply_in(x, y, s, true);
}

void FmmOperator::apply_in(const double *x, double *y, hipStream_t s, bool fp32) const {
    double *pool = fp32 ? reinterpret_cast<double *>(pool32_.data()) : pool_.data();
    dev::launch_csr_apply(near_->structure->row_offsets.data(), near_->structure->columns.data(), near_->values.data(),
                          n_, x, y, s);
    dev::launch_leaf_emit(leaves_.data(), leaf_count_, leaf_dofs_.data(), emission_.data(), x, pool, s, fp32);
    dev::HfLeafLaunch HL{hf_leaves_.data(), hf_leaf_count_, hf_leaf_items_.data(), hf_leaf_item_count_,
                         leaf_dofs_.data(), leaf_scale_.data(), hf_rules_.data(), hf_polar_.data(), leaf_rule_.data(),
                         stats_.rule_points, hf_max_samples_, k_, hf_field_.data(), pool_.data()};
    dev::launch_hf_leaf_emit(HL, x, s);
    dev::FmmTranslateLaunch T{};
    T.rotations = rotations_.data();
    T.work_y = fp32 ? reinterpret_cast<double *>(work_y32_.data()) : work_y_.data();
    T.work_z = fp32 ? reinterpret_cast<double *>(work_z32_.data()) : work_z_.data();
    T.fp32 = fp32;
    T.small_d = small_d_.data();
    T.azimuth = azimuth_.data();
    T.axial = axial_.data();
    T.axial_offsets = axial_offsets_.data();
    T.source_expansions = pool;
    T.target_expansions = pool;
    dev::HfTransferLaunch H{};
    H.rules = hf_rules_.data();
    H.legendre = hf_legendre_.data();
    H.polar = hf_polar_.data();
    H.fourier_ws = hf_fourier_.data();
    H.harmonic_ws = hf_harmonic_.data();
    H.source_expansions = pool_.data();
    H.target_expansions = pool_.data();
    dev::HfDiagonalLaunch D{};
    D.coefficients = diag_table_.data();
    D.source_expansions = pool_.data();
    D.target_expansions = pool_.data();
    for (const Stage &stage : stages_) {
        if (stage.kind == StageKind::translate) {
            const Launch &L = translate_[stage.index];
            T.destinations = destinations_.data() + L.first_destination;
            T.links = links_.data() + L.first_link;
            T.order = link_order_.data() + L.first_link;
            T.chunks = chunks_.data() + L.first_chunk;
            T.axial_order = axial_order_.data() + L.first_link;
            T.axial_chunks = axial_chunks_.data() + L.first_axial_chunk;
            dev::launch_translate(T, L.destination_count, L.link_count, L.chunk_count, L.axial_chunk_count, L.max_order,
                                  split_target_, s);
        } else if (stage.kind == StageKind::angular) {
            const HfLaunch &A = angular_[stage.index];
            H.links = hf_links_.data() + A.first_link;
            H.destinations = hf_destinations_.data() + A.first_destination;
            dev::launch_hf_transfer(H, A.kind, A.link_count, A.destination_count, A.in_rings, A.out_rings, A.out_samples,
                                    s);
        } else {
            const DiagonalLaunch &G = diagonal_[stage.index];
            D.destinations = diag_destinations_.data() + G.first_destination;
            D.links = diag_links_.data() + G.first_link;
            D.samples = G.samples;
            dev::launch_hf_diagonal(D, G.destination_count, s);
        }
    }
    dev::launch_leaf_test(leaves_.data(), leaf_count_, leaf_dofs_.data(), testing_.data(), pool, y, s, fp32);
    dev::launch_hf_leaf_test(HL, y, s);
}

FmmOperator::~FmmOperator() {
    if (graph_exec_)
        (void)hipGraphExecDestroy(graph_exec_);
    if (graph_)
        (void)hipGraphDestroy(graph_);
}

void FmmOperator::apply_graph(const double *x, double *y, hipStream_t s) const {
    const std::size_t bytes = 2 * sizeof(double) * n_;
    if (!graph_exec_) {
        graph_x_.resize(2 * static_cast<std::size_t>(n_));
        graph_y_.resize(2 * static_cast<std::size_t>(n_));
        hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "FMM graph capture");
        apply(graph_x_.data(), graph_y_.data(), s);
        hip_check(hipStreamEndCapture(s, &graph_), "FMM graph capture end");
        hip_check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "FMM graph instantiate");
    }
    hip_check(hipMemcpyAsync(graph_x_.data(), x, bytes, hipMemcpyDeviceToDevice, s), "FMM graph input");
    hip_check(hipGraphLaunch(graph_exec_, s), "FMM graph launch");
    hip_check(hipMemcpyAsync(y, graph_y_.data(), bytes, hipMemcpyDeviceToDevice, s), "FMM graph output");
}

void FmmOperator::apply_direct(const double *x, double *y, hipStream_t s) const {
    if (!direct_available())
        throw std::runtime_error("The direct Galerkin audit is built for meshes of at most 20000 panels.");
    dev::launch_csr_apply(near_->structure->row_offsets.data(), near_->structure->columns.data(), near_->values.data(),
                          n_, x, y, s);
    dev::FmmDirectLaunch D{geo_, nodes_, nt_, n_, vertex_leaf_.data(), near_leaves_.data(), direct_leaf_count_,
                           leaf_rule_.data(), stats_.rule_points, k_, row_scale_, x, y};
    dev::launch_direct_far(D, s);
}

bool FmmOperator::prefer_direct() const {
    return direct_available() && stats_.levels && direct_work_ <= 2 * multipole_work_;
}

void FmmOperator::precondition(const double *v, double *z, hipStream_t s) const {
    if (block_width_ == static_cast<std::uint32_t>(dev::kFmmJacobi))
        dev::jacobi_apply(jacobi_->data(), n_, v, z, s, order_.data());
    else
        dev::jacobi_apply_wide(jacobi_->data(), n_, block_width_, v, z, s, order_.data());
}

namespace {

double weighted_norm(const double *v, std::uint32_t n, const double *weights, Workspace &ws, hipStream_t s) {
    double *out = ws.scratch.data() + ws.scratch.size() - 8;
    dev::launch_norm2(v, n, weights, out, ws.scratch.data(), s);
    double value = 0;
    hip_check(hipMemcpyAsync(&value, out, sizeof(double), hipMemcpyDeviceToHost, s), "norm download");
    hip_check(hipStreamSynchronize(s), "norm sync");
    return std::sqrt(value);
}

} // namespace

namespace {
// GMRES target of the FMM solve as a fraction of the residual tolerance: the
// reference's Krylov settings scale 0.05 times its device GMRES fraction 0.1
// (support_fmm_solve.hpp:11, device_gmres.cu:191-192).
constexpr double kKrylovFraction = 0.05 * 0.1;
constexpr unsigned kMaximumSeedRefinements = 4; // support_fmm_solve.hpp:8
} // namespace

void MeshSolver::fmm_point(const PreparedRequest &P, Workspace &ws, double k, double omega,
                           std::vector<double> &relative, std::vector<std::uint32_t> &iterations, FrequencyResult &fr,
                           const volatile std::int32_t *cancel) {
    using C = std::complex<double>;
    const Request &req = P.request;
    const Numerics &num = req.numerics;
    hipStream_t s = ws.stream;
    const std::uint32_t nv = this->nv(), nt = this->nt(), ne = P.ne;
    const double tol = num.residual_tolerance;
    // active faces: every face with a nonzero velocity in some excitation, ascending
    std::vector<std::uint32_t> faces;
    for (const auto &fv : P.face_velocity)
        for (const auto &[face, v] : fv)
            if (std::abs(v) > 0)
                faces.push_back(face);
    std::sort(faces.begin(), faces.end());
    faces.erase(std::unique(faces.begin(), faces.end()), faces.end());
    FmmSettings fs;
    fs.tree.leaf_capacity = num.fmm_leaf_capacity;
    fs.tree.separation = num.fmm_separation;
    fs.tree.near_ratio = num.near_ratio;
    fs.block_width = num.fmm_block_width;
    fs.accuracy_seed = 0.1 * tol;
    const auto t_build = Clock::now();
    const std::unique_ptr<FmmOperator> op = build_fmm(num, P.k_max, k, fs, faces, s, nullptr);
    fr.ms[0] = ms_since(t_build);
    fr.near_kernel_ms = op->statistics().near_kernel_ms;
    fr.fmm_plane_wave_levels = op->statistics().plane_wave_levels;
    fr.fmm_m2l_links = op->statistics().m2l_links;
    fr.fmm_near_nnz = op->statistics().near_nnz;
    // device bytes of the point: the first operator whole, later ones (audit,
    // verification, comparison, promotion; all share the first one's near entries
    // and preconditioner) without the shared part -- an upper bound of the peak
    std::uint64_t point_bytes = op->statistics().device_bytes;
    auto count_operator = [&point_bytes](const FmmOperator &o) {
        point_bytes += o.statistics().device_bytes - o.statistics().shared_bytes;
    };
    {
        std::lock_guard lock(plan_mutex_);
        if (!internal_of_device_.size())
            internal_of_device_.upload(internal_of_, s);
    }
    // right-hand sides b = B q, mesh vertex order: near flux CSR per face column + far
    // pairs integrated with the operator's endpoint rule
    const std::size_t nvec = 2 * static_cast<std::size_t>(nv) * ne;
    ws.fmm_b.resize(nvec);
    ws.fmm_x.resize(nvec);
    ws.fmm_r.resize(nvec);
    std::vector<double> qf(2 * std::max<std::size_t>(faces.size(), 1) * ne, 0.0);
    for (std::uint32_t e = 0; e < ne; ++e)
        for (std::size_t f = 0; f < faces.size(); ++f) {
            const auto it = P.face_velocity[e].find(faces[f]);
            if (it == P.face_velocity[e].end())
                continue;
            const C q = C(0, -req.density * omega) * it->second;
            qf[2 * (e * faces.size() + f)] = q.real();
            qf[2 * (e * faces.size() + f) + 1] = q.imag();
        }
    ws.fmm_q.upload(qf, s);
    const NearStructure &ns = *op->near().structure;
    DeviceBuffer<std::uint32_t> d_panels, d_pl, d_fl, d_no, d_ns;
    std::uint32_t flux_count = 0;
    {
        const SupportTree &tree = ns.tree;
        std::vector<std::uint32_t> flux_panels;
        for (std::uint32_t t = 0; t < nt; ++t)
            if (ns.plan.panel_face_column[t] != UINT32_MAX)
                flux_panels.push_back(t);
        flux_count = static_cast<std::uint32_t>(flux_panels.size());
        std::vector<std::uint32_t> near_offsets(tree.nodes.size() + 1, 0), near_sources;
        for (const SupportPair &pr : tree.pressure_flux.near)
            ++near_offsets[pr.target + 1];
        for (std::size_t i = 0; i < tree.nodes.size(); ++i)
            near_offsets[i + 1] += near_offsets[i];
        near_sources.resize(near_offsets.back());
        std::vector<std::uint32_t> fill(near_offsets.begin(), near_offsets.end() - 1);
        for (const SupportPair &pr : tree.pressure_flux.near)
            near_sources[fill[pr.target]++] = pr.source;
        d_panels.upload(flux_panels.empty() ? std::vector<std::uint32_t>{0} : flux_panels, s);
        d_pl.upload(tree.pressure_leaf, s);
        d_fl.upload(tree.flux_leaf, s);
        d_no.upload(near_offsets, s);
        d_ns.upload(near_sources.empty() ? std::vector<std::uint32_t>{0} : near_sources, s);
    }
    auto right_hand_sides = [&](const FmmOperator &A, DeviceBuffer<double> &out) {
        out.resize(nvec);
        hip_check(hipMemsetAsync(out.data(), 0, out.bytes(), s), "clear FMM rhs");
        const DeviceNear &dn = A.near();
        dev::launch_flux_csr_apply(ns.flux_offsets.data(), ns.flux_columns.data(), dn.flux_values.data(), nv,
                                   ws.fmm_q.data(), static_cast<std::uint32_t>(faces.size()), ne, out.data(), s);
        dev::FmmFarFluxLaunch F{};
        F.geo = geo_.data();
        F.nodes = original_nodes_.data();
        F.nt = nt;
        F.nv = nv;
        F.nt_all = nt;
        F.flux_panels = d_panels.data();
        F.flux_count = flux_count;
        F.pressure_leaf = d_pl.data();
        F.flux_leaf = d_fl.data();
        F.near_offsets = d_no.data();
        F.near_sources = d_ns.data();
        F.rule = A.leaf_rule_.data();
        F.rule_points = A.statistics().rule_points;
        F.k = k;
        F.row_scale = original_row_scale_.data();
        F.q = ws.flux.data();
        F.excitations = ne;
        F.out = out.data();
        dev::launch_far_flux(F, s);
        hip_check(hipGetLastError(), "far flux launch");
    };
    right_hand_sides(*op, ws.fmm_b);

    // GMRES per excitation on the FMM operator, then the audits of the solution: the
    // direct Galerkin audit or the seed/10 operator (with defect corrections), and
    // the endpoint-quadrature comparison against the next panel rule, promoting the
    // operator's rule until the comparison holds (support_fmm_solve.cu:526-636)
    const auto t_solve = Clock::now();
    // FP32-bulk form (precision mixed): the Krylov iterations apply operators with FP32
    // translations where they have only spherical levels of order <= 48; every residual,
    // audit and comparison applies FP64 operators. The explicit residual of an
    // FP32-translation apply stalls near 1e-7 of its right-hand side (measured 4e-8 on a
    // 3.9k-panel horn, 2.6e-7 on the 118k-panel flange), so the solve runs as refinement
    // cycles: each inner GMRES on the FP32 action reduces its right-hand side r by
    // kFp32CycleReduction, x += d, and r = b - A x is formed with the FP64 action, until
    // the FP64 residual reaches the plain form's kKrylovFraction tol of b; the defect
    // corrections below do the same.
    constexpr double kFp32CycleReduction = 1e-3;
    constexpr unsigned kMaxFp32Cycles = 8;
    const bool fp32_bulk = num.precision == 1;
    auto krylov = [&](const FmmOperator &A) {
        KrylovOperator kop;
        if (fp32_bulk && A.fp32_available()) {
            kop.apply = [&A](const double *v, double *w, hipStream_t st) { A.apply_fp32(v, w, st); };
            fr.fmm_fp32 = true;
        } else {
            kop.apply = [&A](const double *v, double *w, hipStream_t st) { A.apply(v, w, st); };
        }
        if (num.preconditioner == 1)
            kop.precondition = [&A](const double *v, double *z, hipStream_t st) { A.precondition(v, z, st); };
        return kop;
    };
    auto vec = [&](DeviceBuffer<double> &b, std::uint32_t e) { return b.data() + 2 * static_cast<std::size_t>(e) * nv; };
    auto wnorm = [&](const double *v) { return weighted_norm(v, nv, original_lumped_.data(), ws, s); };
    auto ratio = [](double a, double b) { return b > 0 ? a / b : (a == 0 ? 0 : std::numeric_limits<double>::infinity()); };
    std::vector<double> bnorm(ne);
    auto norms_of_b = [&] {
        for (std::uint32_t e = 0; e < ne; ++e) {
            bnorm[e] = wnorm(vec(ws.fmm_b, e));
            if (!std::isfinite(bnorm[e]))
                throw std::runtime_error("the assembled right-hand side is not finite");
        }
    };
    norms_of_b();
    DeviceBuffer<double> cycle_d, cycle_r;
    // x = A^{-1} b by FP32-bulk refinement cycles (kop applies A with FP32 translations,
    // the residuals r = b - A x apply A in FP64) until |r| <= kKrylovFraction tol |b|
    auto fp32_cycles = [&](const KrylovOperator &kop, const FmmOperator &A, const double *b, double *x,
                           std::uint32_t &its_total) {
        cycle_d.resize(2 * static_cast<std::size_t>(nv));
        cycle_r.resize(2 * static_cast<std::size_t>(nv));
        const double bn = wnorm(b);
        hip_check(hipMemsetAsync(x, 0, 2 * sizeof(double) * nv, s), "clear FP32-bulk solution");
        hip_check(hipMemcpyAsync(cycle_r.data(), b, 2 * sizeof(double) * nv, hipMemcpyDeviceToDevice, s),
                  "FP32-bulk residual");
        double rel = bn > 0 ? 1.0 : 0.0;
        for (unsigned cycle = 0; rel > kKrylovFraction * tol; ++cycle) {
            if (cycle == kMaxFp32Cycles) {
                std::ostringstream msg;
                msg << "FP32-bulk refinement stopped after " << cycle << " cycles at relative residual " << rel;
                throw std::runtime_error(msg.str());
            }
            double achieved = 0;
            std::uint32_t its = 0;
            gmres(s, kop, nv, cycle_r.data(), cycle_d.data(), num.fmm_restart, num.fmm_maximum_iterations,
                  std::max(kFp32CycleReduction, 0.5 * kKrylovFraction * tol / rel), &achieved, &its, cancel, ws.gmres);
            its_total += its;
            dev::launch_add(nv, cycle_d.data(), x, s);
            A.apply(x, cycle_r.data(), s);
            dev::launch_subtract_from(nv, b, cycle_r.data(), s);
            rel = wnorm(cycle_r.data()) / bn;
        }
    };
    {
        const KrylovOperator kop = krylov(*op);
        for (std::uint32_t e = 0; e < ne; ++e) {
            double achieved = 0;
            if (fr.fmm_fp32) {
                iterations[e] = 0;
                fp32_cycles(kop, *op, vec(ws.fmm_b, e), vec(ws.fmm_x, e), iterations[e]);
            } else {
                gmres(s, kop, nv, vec(ws.fmm_b, e), vec(ws.fmm_x, e), num.fmm_restart, num.fmm_maximum_iterations,
                      kKrylovFraction * tol, &achieved, &iterations[e], cancel, ws.gmres);
            }
        }
    }
    // r = b - A x per excitation with a given action; relative residuals
    auto residuals = [&](const auto &action) {
        for (std::uint32_t e = 0; e < ne; ++e) {
            action(vec(ws.fmm_x, e), vec(ws.fmm_r, e));
            dev::launch_subtract_from(nv, vec(ws.fmm_b, e), vec(ws.fmm_r, e), s);
            relative[e] = ratio(wnorm(vec(ws.fmm_r, e)), bnorm[e]);
        }
    };
    DeviceBuffer<double> ax(2 * static_cast<std::size_t>(nv)), t1(2 * static_cast<std::size_t>(nv)),
        t2(2 * static_cast<std::size_t>(nv)), b_next;
    // defect corrections x += A_op^{-1} r while the audited residual exceeds tol
    auto correct = [&](const FmmOperator &krylov_op, const auto &action) {
        const KrylovOperator kop = krylov(krylov_op);
        for (unsigned step = 0; *std::max_element(relative.begin(), relative.end()) > tol; ++step) {
            if (step == num.refinement_steps)
                break; // reported by the residual check below
            for (std::uint32_t e = 0; e < ne; ++e) {
                if (!(relative[e] > tol))
                    continue;
                double achieved = 0;
                std::uint32_t its = 0;
                // d = A_op^{-1} r (FP32-bulk: refinement cycles on the FP32 action with
                // FP64 residuals of A_op), x += d
                if (fp32_bulk && krylov_op.fp32_available())
                    fp32_cycles(kop, krylov_op, vec(ws.fmm_r, e), ax.data(), its);
                else
                    gmres(s, kop, nv, vec(ws.fmm_r, e), ax.data(), num.fmm_restart, num.fmm_maximum_iterations,
                          kKrylovFraction * tol, &achieved, &its, cancel, ws.gmres);
                dev::launch_add(nv, ax.data(), vec(ws.fmm_x, e), s);
                iterations[e] += its;
            }
            residuals(action);
            fr.fmm_corrected = true;
        }
    };
    std::unique_ptr<FmmOperator> promoted; // the operator after endpoint-rule promotions
    for (;;) {
        const FmmOperator &A = promoted ? *promoted : *op;
        const bool direct = A.prefer_direct();
        std::unique_ptr<FmmOperator> audit;
        std::function<void(const double *, double *)> audited;
        if (direct) {
            audited = [&](const double *x, double *y) { A.apply_direct(x, y, s); };
            residuals(audited);
            correct(A, audited);
            fr.fmm_audit = 0;
        } else {
            // seed audits (support_fmm_solve.cu:393-524): the solution is corrected with
            // the seed/10 operator until its residual holds, then checked against the
            // next operator with a distinct level plan (seed/100, ...): their actions at
            // the solution must agree to 0.25 tol; otherwise the finer one takes over
            auto action = [&](const FmmOperator &O) {
                return [&O, &s](const double *x, double *y) { O.apply(x, y, s); };
            };
            auto same_levels = [](const FmmOperator &a, const FmmOperator &b) {
                const FmmStatistics &x = a.statistics(), &y = b.statistics();
                return x.level_orders == y.level_orders && x.level_representation == y.level_representation &&
                       x.level_rule_orders == y.level_rule_orders;
            };
            FmmSettings aset = fs;
            aset.accuracy_seed = fs.accuracy_seed / 10;
            audit = build_fmm(num, P.k_max, k, aset, faces, s, &A, A.statistics().endpoint_degree);
            count_operator(*audit);
            residuals(action(*audit));
            DeviceBuffer<double> prod(nvec);
            for (unsigned refinements = 0;;) {
                correct(*audit, action(*audit));
                for (std::uint32_t e = 0; e < ne; ++e)
                    audit->apply(vec(ws.fmm_x, e), vec(prod, e), s);
                std::unique_ptr<FmmOperator> verify;
                for (;;) {
                    aset.accuracy_seed /= 10;
                    verify = build_fmm(num, P.k_max, k, aset, faces, s, &A, A.statistics().endpoint_degree);
                    count_operator(*verify);
                    if (!same_levels(*verify, *audit))
                        break;
                    if (refinements == kMaximumSeedRefinements)
                        throw std::runtime_error("FMM seed refinement exhausted without a distinct level plan.");
                    ++refinements;
                }
                residuals(action(*verify));
                double discrepancy = 0;
                for (std::uint32_t e = 0; e < ne; ++e) {
                    // A_audit x - A_verify x = prod - (b - r)
                    hip_check(hipMemcpyAsync(t1.data(), vec(ws.fmm_r, e), t1.bytes(), hipMemcpyDeviceToDevice, s), "copy");
                    dev::launch_subtract_from(nv, vec(ws.fmm_b, e), t1.data(), s); // b - r = A_verify x
                    dev::launch_subtract_from(nv, vec(prod, e), t1.data(), s);     // prod - A_verify x
                    discrepancy = std::max(discrepancy, ratio(wnorm(t1.data()), bnorm[e]));
                }
                fr.fmm_audit = discrepancy;
                fr.fmm_verification_seed = aset.accuracy_seed;
                audit = std::move(verify);
                if (*std::max_element(relative.begin(), relative.end()) <= tol && discrepancy <= 0.25 * tol)
                    break;
                if (refinements == kMaximumSeedRefinements) {
                    std::ostringstream msg;
                    msg.precision(6);
                    msg << "FMM accuracy-seed refinement exhausted: residual "
                        << *std::max_element(relative.begin(), relative.end()) << ", operator discrepancy "
                        << discrepancy << "; tolerance " << tol << ".";
                    throw std::runtime_error(msg.str());
                }
                ++refinements;
                fr.fmm_corrected = true;
            }
            audited = [&](const double *x, double *y) { audit->apply(x, y, s); };
        }
        fr.fmm_direct_audit = direct;
        // endpoint-quadrature comparison: the audited action and B at the next rule
        const unsigned degree = A.statistics().endpoint_degree, next = refined_endpoint_degree(degree);
        if (!next)
            throw std::runtime_error("fast multipole panel rule cannot be refined further for the endpoint comparison");
        FmmSettings cset = fs;
        if (!direct)
            cset.accuracy_seed = fr.fmm_verification_seed; // the verified seed, the next rule
        const std::unique_ptr<FmmOperator> cmp = build_fmm(num, P.k_max, k, cset, faces, s, &A, next);
        count_operator(*cmp);
        right_hand_sides(*cmp, b_next);
        double quad = 0, next_residual = 0;
        for (std::uint32_t e = 0; e < ne; ++e) {
            if (direct)
                cmp->apply_direct(vec(ws.fmm_x, e), t1.data(), s);
            else
                cmp->apply(vec(ws.fmm_x, e), t1.data(), s);
            audited(vec(ws.fmm_x, e), t2.data());
            dev::launch_subtract_from(nv, t1.data(), t2.data(), s); // A_next x - A x
            const double da = wnorm(t2.data());
            hip_check(hipMemcpyAsync(t2.data(), vec(b_next, e), t2.bytes(), hipMemcpyDeviceToDevice, s), "copy");
            dev::launch_subtract_from(nv, vec(ws.fmm_b, e), t2.data(), s); // b_next - b
            const double db = wnorm(t2.data());
            const double bn = wnorm(vec(b_next, e));
            dev::launch_subtract_from(nv, vec(b_next, e), t1.data(), s); // b_next - A_next x
            quad = std::max(quad, ratio(da + db, bn));
            next_residual = std::max(next_residual, ratio(wnorm(t1.data()), bn));
        }
        fr.fmm_quadrature_discrepancy = quad;
        fr.fmm_endpoint_degree = degree;
        if (quad <= 0.025 * tol && next_residual <= tol)
            break;
        if (degree >= 12) {
            std::ostringstream msg;
            msg.precision(6);
            msg << "fast multipole panel-rule comparison " << degree << "/" << next << " did not converge: A/B action difference "
                << quad << " exceeds " << 0.025 * tol << " or the higher-rule residual " << next_residual
                << " exceeds " << tol << "; use a finer mesh";
            throw std::runtime_error(msg.str());
        }
        // promote the operator's rule; the right-hand side follows the new rule
        promoted = build_fmm(num, P.k_max, k, fs, faces, s, &A, next);
        count_operator(*promoted);
        ++fr.fmm_promotions;
        right_hand_sides(*promoted, ws.fmm_b);
        norms_of_b();
        residuals([&](const double *x, double *y) { promoted->apply(x, y, s); });
        correct(*promoted, [&](const double *x, double *y) { promoted->apply(x, y, s); });
    }
    const FmmOperator &final_op = promoted ? *promoted : *op;
    fr.ms[1] = ms_since(t_solve);
    point_bytes += ws.fmm_b.capacity_bytes() + ws.fmm_x.capacity_bytes() + ws.fmm_r.capacity_bytes() +
                   ax.capacity_bytes() + t1.capacity_bytes() + t2.capacity_bytes() + b_next.capacity_bytes() +
                   2 * static_cast<std::uint64_t>(nvec) * sizeof(double); // the seed audit's product vectors
    fr.fmm_device_bytes = point_bytes;
    note_fmm_point_bytes(num, point_bytes);
    {
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "event");
        hip_check(hipEventCreate(&e1), "event");
        hip_check(hipEventRecord(e0, s), "event");
        final_op.apply(ws.fmm_x.data(), ax.data(), s);
        hip_check(hipEventRecord(e1, s), "event");
        hip_check(hipEventSynchronize(e1), "event sync");
        float ms = 0;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        fr.fmm_apply_ms = ms;
    }
    for (std::uint32_t e = 0; e < ne; ++e)
        if (!(relative[e] <= tol)) {
            std::ostringstream msg;
            msg.precision(9);
            msg << "relative residual " << relative[e] << " is above the requested tolerance " << tol << " at "
                << omega / (2 * std::numbers::pi) << " Hz (fast multipole form).";
            throw std::runtime_error(msg.str());
        }
    // pressures (and right-hand sides) in the internal order of the result path
    dev::launch_permute(ws.fmm_x.data(), ws.x.data(), internal_of_device_.data(), nv, ne, true, s);
    dev::launch_permute(ws.fmm_b.data(), ws.rhs.data(), internal_of_device_.data(), nv, ne, true, s);
    hip_check(hipStreamSynchronize(s), "FMM point sync");
}

} // namespace hsbem
