// Near part of the wideband FMM operator on the device.
//
// Replaces SupportNearOperator::assemble (modules/bem/cuda/support_near_operator.cu:24-851):
// the near panel-pair jobs of the support near plan (host planner) are split by
// the Galerkin interaction plan into touching pairs (Sauter-Schwab), adaptive
// near pairs (tensor patches per thread, projection fans per wavefront) and
// regular pairs (qualified degree), integrated in FP64 by the dense path's
// kernels and accumulated into the CSR entries their masks select; the mass
// terms close the near blocks. The SpMV applies A_near to a pressure vector.
//
// The split job lists and the CSR pattern depend on the mesh, the tree settings,
// the active faces and the planning wavenumber only: they are built once
// (NearStructure, device resident) and integrated per frequency.
#include "fmm_operator.hpp"
#include "host/support_near_plan.hpp"
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <thread>
#include <type_traits>

namespace hsbem {

std::shared_ptr<const NearStructure> MeshSolver::near_structure(const Numerics &n, double k_max,
                                                                const SupportTreeSettings &ts,
                                                                const std::vector<std::uint32_t> &faces,
                                                                hipStream_t s) {
    const auto key = std::make_tuple(ts.leaf_capacity, ts.maximum_depth, ts.separation, ts.near_ratio, faces,
                                     n.quadrature_order, n.singular_order, n.near_ratio, n.near_tolerance, n.near_depth,
                                     k_max);
    {
        std::lock_guard lock(plan_mutex_);
        if (near_cache_ && near_cache_key_ == key)
            return near_cache_;
    }
    const std::uint32_t nt = this->nt(), nv = this->nv();
    auto ns = std::make_shared<NearStructure>();
    ns->tree = make_support_tree(space_, ts);
    ns->plan = make_support_near_plan(space_, ns->tree, faces);
    const std::vector<SupportNearJob> jobs = support_near_jobs(space_, ns->tree, ns->plan);
    const std::shared_ptr<DevicePlan> dp_ref = device_plan(n, k_max, s, nullptr, device_estimates_);
    const InteractionPlan &ip = dp_ref->plan;
    ns->interaction = dp_ref;
    // split the jobs by integration route: contiguous job ranges per thread, each
    // list concatenated in job order
    struct Split {
        std::vector<std::uint32_t> regular, regular_masks, touching_masks, fan_masks, patch_masks;
        std::vector<TouchingJob> touching;
        std::vector<NearJob> fans;
        std::vector<dev::NearPatchItem> items;
    };
    const unsigned workers = static_cast<unsigned>(std::max<std::size_t>(
        1, std::min<std::size_t>(std::thread::hardware_concurrency(), jobs.size() / 65536 + 1)));
    std::vector<Split> parts(workers);
    auto classify = [&](unsigned w) {
        Split &out = parts[w];
        const std::size_t first = jobs.size() * w / workers, last = jobs.size() * (w + 1) / workers;
        for (std::size_t q = first; q < last; ++q) {
            const SupportNearJob &j = jobs[q];
            const std::uint32_t mask = j.pressure_mask | static_cast<std::uint32_t>(j.flux_mask) << 9;
            const Panel &a = space_.panels[j.test], &b = space_.panels[j.source];
            const Permutation perm = classify_pair_ids(a.nodes.data(), b.nodes.data());
            if (perm.topology != Topology::disjoint) {
                out.touching.push_back({j.test, j.source, perm});
                out.touching_masks.push_back(mask);
                continue;
            }
            const auto lo = ip.near_trials.begin() + ip.near_offsets[j.test],
                       hi = ip.near_trials.begin() + ip.near_offsets[j.test + 1];
            const auto it = std::lower_bound(lo, hi, j.source);
            if (it != hi && *it == j.source) {
                const NearJob &nj = ip.near[static_cast<std::size_t>(it - ip.near_trials.begin())];
                bool fan = false;
                for (std::uint32_t p = 0; p < nj.patch_count; ++p)
                    fan |= ip.patches[nj.first_patch + p].projection != 0;
                if (fan) {
                    out.fans.push_back(nj);
                    out.fan_masks.push_back(mask);
                } else {
                    for (std::uint32_t p = 0; p < nj.patch_count; ++p) {
                        out.items.push_back({nj.test, nj.trial, nj.first_patch + p, nj.swap});
                        out.patch_masks.push_back(mask);
                    }
                }
                continue;
            }
            out.regular.push_back(j.test);
            out.regular.push_back(j.source);
            out.regular_masks.push_back(mask);
        }
    };
    {
        std::vector<std::thread> threads;
        for (unsigned w = 1; w < workers; ++w)
            threads.emplace_back(classify, w);
        classify(0);
        for (auto &th : threads)
            th.join();
    }
    auto gather = [&](auto member) {
        std::remove_reference_t<decltype(parts[0].*member)> all;
        std::size_t total = 0;
        for (auto &p : parts)
            total += (p.*member).size();
        all.reserve(total);
        for (auto &p : parts) {
            all.insert(all.end(), (p.*member).begin(), (p.*member).end());
            std::remove_reference_t<decltype(p.*member)>().swap(p.*member);
        }
        return all;
    };
    const std::vector<std::uint32_t> regular = gather(&Split::regular), regular_masks = gather(&Split::regular_masks),
                                     touching_masks = gather(&Split::touching_masks),
                                     fan_masks = gather(&Split::fan_masks), patch_masks = gather(&Split::patch_masks);
    const std::vector<TouchingJob> touching = gather(&Split::touching);
    const std::vector<NearJob> fans = gather(&Split::fans);
    const std::vector<dev::NearPatchItem> items = gather(&Split::items);
    {
        // mesh-order node ids and row scales (the dense path uses a cluster order)
        std::lock_guard lock(plan_mutex_);
        if (!original_nodes_.size()) {
            std::vector<std::uint32_t> nodes(3 * static_cast<std::size_t>(nt));
            for (std::uint32_t t = 0; t < nt; ++t)
                for (int l = 0; l < 3; ++l)
                    nodes[static_cast<std::size_t>(l) * nt + t] = space_.panels[t].nodes[l];
            std::vector<double> scale(nv), mass(nv);
            for (std::uint32_t i = 0; i < nv; ++i) {
                scale[i] = 1.0 / space_.lumped_mass[i];
                mass[i] = space_.lumped_mass[i];
            }
            original_nodes_.upload(nodes, s);
            original_row_scale_.upload(scale, s);
            original_lumped_.upload(mass, s);
        }
    }
    auto up32 = [&](DeviceBuffer<std::uint32_t> &d, const std::vector<std::uint32_t> &v) {
        d.upload(v.empty() ? std::vector<std::uint32_t>{0} : v, s);
    };
    up32(ns->row_offsets, ns->plan.pressure.row_offsets);
    up32(ns->columns, ns->plan.pressure.columns);
    up32(ns->flux_offsets, ns->plan.flux.row_offsets);
    up32(ns->flux_columns, ns->plan.flux.columns);
    up32(ns->face_column, ns->plan.panel_face_column);
    up32(ns->regular, regular);
    up32(ns->regular_masks, regular_masks);
    up32(ns->touching_masks, touching_masks);
    up32(ns->fan_masks, fan_masks);
    up32(ns->patch_masks, patch_masks);
    ns->touching.upload(touching.empty() ? std::vector<TouchingJob>(1) : touching, s);
    ns->fans.upload(fans.empty() ? std::vector<NearJob>(1) : fans, s);
    ns->items.upload(items.empty() ? std::vector<dev::NearPatchItem>(1) : items, s);
    ns->nnz = ns->plan.pressure.columns.size();
    ns->fnnz = ns->plan.flux.columns.size();
    ns->jobs = jobs.size();
    ns->regular_count = regular_masks.size();
    ns->touching_count = touching.size();
    ns->patch_count = items.size();
    ns->fan_count = fans.size();
    hip_check(hipStreamSynchronize(s), "near structure upload");
    std::lock_guard lock(plan_mutex_);
    near_cache_ = ns;
    near_cache_key_ = key;
    near_bytes_ = ns->device_bytes();
    return ns;
}

void MeshSolver::integrate_near(const Numerics &n, double k, const std::shared_ptr<const NearStructure> &ns,
                                hipStream_t s, DeviceNear &out) {
    const std::uint32_t nt = this->nt(), nv = this->nv();
    out.structure = ns;
    out.nnz = ns->nnz;
    out.fnnz = ns->fnnz;
    out.values.resize(2 * std::max<std::size_t>(ns->nnz, 1));
    out.flux_values.resize(2 * std::max<std::size_t>(ns->fnnz, 1));
    DeviceBuffer<int> d_err(1);
    hip_check(hipMemsetAsync(out.values.data(), 0, out.values.bytes(), s), "clear near values");
    hip_check(hipMemsetAsync(out.flux_values.data(), 0, out.flux_values.bytes(), s), "clear near flux values");
    hip_check(hipMemsetAsync(d_err.data(), 0, sizeof(int), s), "clear error");

    const DevicePlan &dp = *ns->interaction;
    dev::NearOperatorLaunch L{};
    L.args.panels = {geo_.data(), original_nodes_.data(), bounds_.data(), shape_.data(), nt, nv};
    L.args.row_scale = original_row_scale_.data();
    L.args.k = k;
    L.args.tolerance = n.near_tolerance;
    L.args.regular_order = n.quadrature_order;
    L.args.error = d_err.data();
    L.csr = {nullptr,          ns->row_offsets.data(),  ns->columns.data(),      out.values.data(),
             ns->flux_offsets.data(), ns->flux_columns.data(), out.flux_values.data(), ns->face_column.data()};
    L.near = {dp.near_offsets.data(), dp.near_trials.data(), dp.whole.size() ? dp.whole.data() : nullptr};
    L.rules = rules_.data();
    L.regular = ns->regular.data();
    L.regular_count = static_cast<std::uint32_t>(ns->regular_count);
    L.touching = ns->touching.data();
    L.touching_count = static_cast<std::uint32_t>(ns->touching_count);
    L.gauss = dp.gauss.data();
    L.singular_order = n.singular_order;
    L.fan_jobs = ns->fans.data();
    L.fan_count = static_cast<std::uint32_t>(ns->fan_count);
    L.patch_items = ns->items.data();
    L.patch_count = static_cast<std::uint32_t>(ns->patch_count);
    L.patches = dp.patches.data();
    L.regular_masks = ns->regular_masks.data();
    L.touching_masks = ns->touching_masks.data();
    L.fan_masks = ns->fan_masks.data();
    L.patch_masks = ns->patch_masks.data();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, s), "event");
    dev::launch_near_operator(L, s);
    hip_check(hipEventRecord(e1, s), "event");
    hip_check(hipGetLastError(), "near operator launch");
    int code = 0;
    hip_check(hipMemcpyAsync(&code, d_err.data(), sizeof(int), hipMemcpyDeviceToHost, s), "error download");
    hip_check(hipStreamSynchronize(s), "near operator sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (code)
        throw std::runtime_error("BEM near operator quadrature failed (device error " + std::to_string(code) + ").");
    out.kernel_ms = ms;
}

NearOperator MeshSolver::near_operator(const Numerics &n, double k_max, double k, const SupportTreeSettings &ts,
                                       const std::vector<std::uint32_t> &faces, const double *x) {
    std::lock_guard lock(mutex_);
    hipStream_t s = ctx_->stream;
    const std::uint32_t nv = this->nv();
    DeviceNear dn;
    integrate_near(n, k, near_structure(n, k_max, ts, faces, s), s, dn);
    const NearStructure &ns = *dn.structure;
    NearOperator out;
    out.row_offsets = ns.plan.pressure.row_offsets;
    out.columns = ns.plan.pressure.columns;
    out.flux_offsets = ns.plan.flux.row_offsets;
    out.flux_columns = ns.plan.flux.columns;
    out.values.resize(2 * dn.nnz);
    out.flux_values.resize(2 * dn.fnnz);
    hip_check(hipMemcpyAsync(out.values.data(), dn.values.data(), sizeof(double) * out.values.size(),
                             hipMemcpyDeviceToHost, s),
              "near values download");
    hip_check(hipMemcpyAsync(out.flux_values.data(), dn.flux_values.data(), sizeof(double) * out.flux_values.size(),
                             hipMemcpyDeviceToHost, s),
              "near flux values download");
    if (x) {
        DeviceBuffer<double> dx, dy(2 * static_cast<std::size_t>(nv));
        dx.upload(std::vector<double>(x, x + 2 * static_cast<std::size_t>(nv)), s);
        dev::launch_csr_apply(ns.row_offsets.data(), ns.columns.data(), dn.values.data(), nv, dx.data(), dy.data(), s);
        out.applied.resize(2 * static_cast<std::size_t>(nv));
        hip_check(hipMemcpyAsync(out.applied.data(), dy.data(), dy.bytes(), hipMemcpyDeviceToHost, s), "apply download");
    }
    hip_check(hipStreamSynchronize(s), "near operator download");
    out.jobs = ns.jobs;
    out.regular = ns.regular_count;
    out.touching = ns.touching_count;
    out.patches = ns.patch_count;
    out.fan_jobs = ns.fan_count;
    out.kernel_ms = dn.kernel_ms;
    return out;
}

} // namespace hsbem
