// Near part of the wideband FMM operator on the device.
//
// Replaces SupportNearOperator::assemble (modules/bem/cuda/support_near_operator.cu:24-851):
// the near panel-pair jobs of the support near plan (host planner) are split by
// the Galerkin interaction plan into touching pairs (Sauter-Schwab), adaptive
// near pairs (tensor patches per thread, projection fans per wavefront) and
// regular pairs (qualified degree), integrated in FP64 by the dense path's
// kernels and accumulated into the CSR entries their masks select; the mass
// terms close the near blocks. The SpMV applies A_near to a pressure vector.
#include "fmm_operator.hpp"
#include "host/support_near_plan.hpp"
#include "solver.hpp"

#include <algorithm>
#include <chrono>

namespace hsbem {

void MeshSolver::build_near(const Numerics &n, double k_max, double k, const SupportTree &tree,
                            const SupportNearPlan &plan, hipStream_t s, DeviceNear &out) {
    const std::uint32_t nt = this->nt(), nv = this->nv();
    const std::vector<SupportNearJob> jobs = support_near_jobs(space_, tree, plan);
    const std::shared_ptr<DevicePlan> dp_ref = device_plan(n, k_max, s, nullptr);
    const InteractionPlan &ip = dp_ref->plan;
    // split the jobs by integration route
    std::vector<std::uint32_t> regular, regular_masks, touching_masks, fan_masks, patch_masks;
    std::vector<TouchingJob> touching;
    std::vector<NearJob> fans;
    std::vector<dev::NearPatchItem> items;
    for (const SupportNearJob &j : jobs) {
        const std::uint32_t mask = j.pressure_mask | static_cast<std::uint32_t>(j.flux_mask) << 9;
        const Panel &a = space_.panels[j.test], &b = space_.panels[j.source];
        const Permutation perm = classify_pair_ids(a.nodes.data(), b.nodes.data());
        if (perm.topology != Topology::disjoint) {
            touching.push_back({j.test, j.source, perm});
            touching_masks.push_back(mask);
            continue;
        }
        const auto first = ip.near_trials.begin() + ip.near_offsets[j.test],
                   last = ip.near_trials.begin() + ip.near_offsets[j.test + 1];
        const auto it = std::lower_bound(first, last, j.source);
        if (it != last && *it == j.source) {
            const NearJob &nj = ip.near[static_cast<std::size_t>(it - ip.near_trials.begin())];
            bool fan = false;
            for (std::uint32_t q = 0; q < nj.patch_count; ++q)
                fan |= ip.patches[nj.first_patch + q].projection != 0;
            if (fan) {
                fans.push_back(nj);
                fan_masks.push_back(mask);
            } else {
                for (std::uint32_t q = 0; q < nj.patch_count; ++q) {
                    items.push_back({nj.test, nj.trial, nj.first_patch + q, nj.swap});
                    patch_masks.push_back(mask);
                }
            }
            continue;
        }
        regular.push_back(j.test);
        regular.push_back(j.source);
        regular_masks.push_back(mask);
    }
    // mesh-order node ids and row scales (the dense path uses a cluster order)
    if (!original_nodes_.size()) {
        std::vector<std::uint32_t> nodes(3 * static_cast<std::size_t>(nt));
        for (std::uint32_t t = 0; t < nt; ++t)
            for (int l = 0; l < 3; ++l)
                nodes[static_cast<std::size_t>(l) * nt + t] = space_.panels[t].nodes[l];
        std::vector<double> scale(nv);
        for (std::uint32_t i = 0; i < nv; ++i)
            scale[i] = 1.0 / space_.lumped_mass[i];
        original_nodes_.upload(nodes, s);
        original_row_scale_.upload(scale, s);
    }
    auto up32 = [&](DeviceBuffer<std::uint32_t> &d, const std::vector<std::uint32_t> &v) {
        d.upload(v.empty() ? std::vector<std::uint32_t>{0} : v, s);
    };
    DeviceBuffer<std::uint32_t> &d_ro = out.row_offsets, &d_co = out.columns, &d_fo = out.flux_offsets,
                                &d_fc = out.flux_columns;
    DeviceBuffer<std::uint32_t> d_face, d_reg, d_rm, d_tm, d_fm, d_pm;
    up32(d_ro, plan.pressure.row_offsets);
    up32(d_co, plan.pressure.columns);
    up32(d_fo, plan.flux.row_offsets);
    up32(d_fc, plan.flux.columns);
    up32(d_face, plan.panel_face_column);
    up32(d_reg, regular);
    up32(d_rm, regular_masks);
    up32(d_tm, touching_masks);
    up32(d_fm, fan_masks);
    up32(d_pm, patch_masks);
    DeviceBuffer<TouchingJob> d_touch;
    d_touch.upload(touching.empty() ? std::vector<TouchingJob>(1) : touching, s);
    DeviceBuffer<NearJob> d_fans;
    d_fans.upload(fans.empty() ? std::vector<NearJob>(1) : fans, s);
    DeviceBuffer<dev::NearPatchItem> d_items;
    d_items.upload(items.empty() ? std::vector<dev::NearPatchItem>(1) : items, s);
    const std::size_t nnz = plan.pressure.columns.size(), fnnz = plan.flux.columns.size();
    DeviceBuffer<double> &d_val = out.values, &d_fval = out.flux_values;
    d_val.resize(2 * std::max<std::size_t>(nnz, 1));
    d_fval.resize(2 * std::max<std::size_t>(fnnz, 1));
    DeviceBuffer<int> d_err(1);
    hip_check(hipMemsetAsync(d_val.data(), 0, d_val.bytes(), s), "clear near values");
    hip_check(hipMemsetAsync(d_fval.data(), 0, d_fval.bytes(), s), "clear near flux values");
    hip_check(hipMemsetAsync(d_err.data(), 0, sizeof(int), s), "clear error");

    dev::NearOperatorLaunch L{};
    L.args.panels = {geo_.data(), original_nodes_.data(), bounds_.data(), shape_.data(), nt, nv};
    L.args.row_scale = original_row_scale_.data();
    L.args.k = k;
    L.args.tolerance = n.near_tolerance;
    L.args.regular_order = n.quadrature_order;
    L.args.error = d_err.data();
    L.csr = {nullptr, d_ro.data(), d_co.data(), d_val.data(), d_fo.data(), d_fc.data(), d_fval.data(), d_face.data()};
    L.near = {dp_ref->near_offsets.data(), dp_ref->near_trials.data(),
              dp_ref->whole.size() ? dp_ref->whole.data() : nullptr};
    L.rules = rules_.data();
    L.regular = d_reg.data();
    L.regular_count = static_cast<std::uint32_t>(regular_masks.size());
    L.touching = d_touch.data();
    L.touching_count = static_cast<std::uint32_t>(touching.size());
    L.gauss = dp_ref->gauss.data();
    L.singular_order = n.singular_order;
    L.fan_jobs = d_fans.data();
    L.fan_count = static_cast<std::uint32_t>(fans.size());
    L.patch_items = d_items.data();
    L.patch_count = static_cast<std::uint32_t>(items.size());
    L.patches = dp_ref->patches.data();
    L.regular_masks = d_rm.data();
    L.touching_masks = d_tm.data();
    L.fan_masks = d_fm.data();
    L.patch_masks = d_pm.data();
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
    out.nnz = nnz;
    out.fnnz = fnnz;
    out.jobs = jobs.size();
    out.regular = regular_masks.size();
    out.touching = touching.size();
    out.patches = items.size();
    out.fan_jobs = fans.size();
    out.kernel_ms = ms;
}

NearOperator MeshSolver::near_operator(const Numerics &n, double k_max, double k, const SupportTreeSettings &ts,
                                       const std::vector<std::uint32_t> &faces, const double *x) {
    std::lock_guard lock(mutex_);
    hipStream_t s = ctx_->stream;
    const std::uint32_t nv = this->nv();
    const SupportTree tree = make_support_tree(space_, ts);
    const SupportNearPlan plan = make_support_near_plan(space_, tree, faces);
    DeviceNear dn;
    build_near(n, k_max, k, tree, plan, s, dn);
    const std::size_t nnz = dn.nnz, fnnz = dn.fnnz;
    DeviceBuffer<double> &d_val = dn.values, &d_fval = dn.flux_values;
    DeviceBuffer<std::uint32_t> &d_ro = dn.row_offsets, &d_co = dn.columns;
    NearOperator out;
    out.row_offsets = plan.pressure.row_offsets;
    out.columns = plan.pressure.columns;
    out.flux_offsets = plan.flux.row_offsets;
    out.flux_columns = plan.flux.columns;
    out.values.resize(2 * nnz);
    out.flux_values.resize(2 * fnnz);
    hip_check(hipMemcpyAsync(out.values.data(), d_val.data(), sizeof(double) * out.values.size(), hipMemcpyDeviceToHost, s),
              "near values download");
    hip_check(hipMemcpyAsync(out.flux_values.data(), d_fval.data(), sizeof(double) * out.flux_values.size(),
                             hipMemcpyDeviceToHost, s),
              "near flux values download");
    if (x) {
        DeviceBuffer<double> dx, dy(2 * static_cast<std::size_t>(nv));
        dx.upload(std::vector<double>(x, x + 2 * static_cast<std::size_t>(nv)), s);
        dev::launch_csr_apply(d_ro.data(), d_co.data(), d_val.data(), nv, dx.data(), dy.data(), s);
        out.applied.resize(2 * static_cast<std::size_t>(nv));
        hip_check(hipMemcpyAsync(out.applied.data(), dy.data(), dy.bytes(), hipMemcpyDeviceToHost, s), "apply download");
    }
    hip_check(hipStreamSynchronize(s), "near operator download");
    out.jobs = dn.jobs;
    out.regular = dn.regular;
    out.touching = dn.touching;
    out.patches = dn.patches;
    out.fan_jobs = dn.fan_jobs;
    out.kernel_ms = dn.kernel_ms;
    return out;
}

} // namespace hsbem
