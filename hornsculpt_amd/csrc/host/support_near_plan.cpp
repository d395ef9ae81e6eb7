#include "support_near_plan.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>
#include <thread>

namespace hsbem {
namespace {

constexpr std::uint32_t kNone = UINT32_MAX;

// Offsets from per-row counts stored at [row + 1]; 32-bit indices.
void exclusive_scan(std::vector<std::uint32_t> &offsets) {
    std::uint64_t sum = 0;
    for (auto &o : offsets) {
        sum += o;
        if (sum > UINT32_MAX)
            throw std::overflow_error("near structure is too large for 32-bit indices");
        o = static_cast<std::uint32_t>(sum);
    }
}

// Near interactions of a partition grouped by target node (stable).
void group_by_target(const SupportTree &tree, const std::vector<SupportPair> &pairs, std::vector<std::uint32_t> &offsets,
                     std::vector<std::uint32_t> &indices) {
    if (pairs.size() > UINT32_MAX)
        throw std::overflow_error("near structure: block index overflow");
    offsets.assign(tree.nodes.size() + 1, 0);
    for (const SupportPair &p : pairs) {
        if (p.target >= tree.nodes.size() || p.source >= tree.nodes.size() || !tree.nodes[p.target].leaf() ||
            !tree.nodes[p.source].leaf())
            throw std::invalid_argument("near structure: a near block is not a pair of leaves");
        ++offsets[p.target + 1];
    }
    exclusive_scan(offsets);
    std::vector<std::uint32_t> fill(offsets.begin(), offsets.end() - 1);
    indices.assign(pairs.size(), 0);
    for (std::uint32_t i = 0; i < pairs.size(); ++i)
        indices[fill[pairs[i].target]++] = i;
}

} // namespace

SupportNearPlan make_support_near_plan(const Space &space, const SupportTree &tree,
                                       const std::vector<std::uint32_t> &faces) {
    const std::size_t nv = space.vertices.size(), nt = space.panels.size();
    if (nv != tree.pressure_order.size() || nt != tree.flux_order.size() || tree.nodes.empty() ||
        tree.pressure_leaf.size() != nv || tree.flux_leaf.size() != nt)
        throw std::invalid_argument("near structure: tree and mesh do not match");
    if (faces.size() > UINT32_MAX)
        throw std::overflow_error("near structure: too many driven face columns");
    SupportNearPlan plan;
    plan.active_faces = faces;
    for (std::size_t i = 0; i < faces.size(); ++i)
        if (std::find(faces.begin(), faces.begin() + i, faces[i]) != faces.begin() + i)
            throw std::invalid_argument("near structure: a driven face id is listed twice");
    plan.panel_face_column.assign(nt, kNone);
    std::vector<char> seen(faces.size(), 0);
    for (std::size_t p = 0; p < nt; ++p) {
        const auto it = std::find(faces.begin(), faces.end(), space.panels[p].face_id);
        if (it == faces.end())
            continue;
        if (space.panels[p].boundary != Boundary::driven)
            throw std::invalid_argument("near structure: a listed face has a wall panel");
        const auto column = static_cast<std::uint32_t>(it - faces.begin());
        plan.panel_face_column[p] = column;
        seen[column] = 1;
    }
    if (std::find(seen.begin(), seen.end(), 0) != seen.end())
        throw std::invalid_argument("near structure: a listed face has no driven panel");

    // flux CSR: one column set per target leaf, shared by its P1 rows
    group_by_target(tree, tree.pressure_flux.near, plan.flux_block_offsets, plan.flux_block_indices);
    std::vector<std::vector<std::uint32_t>> leaf_columns(tree.nodes.size());
    plan.flux.row_offsets.assign(nv + 1, 0);
    {
        std::vector<std::uint32_t> stamp(faces.size(), 0);
        for (std::size_t n = 0; n < tree.nodes.size(); ++n) {
            const SupportNode &node = tree.nodes[n];
            if (!node.leaf() || !node.pressure_count)
                continue;
            auto &cols = leaf_columns[n];
            const auto epoch = static_cast<std::uint32_t>(n + 1);
            for (std::uint32_t b = plan.flux_block_offsets[n]; b < plan.flux_block_offsets[n + 1]; ++b) {
                const SupportNode &src = tree.nodes[tree.pressure_flux.near[plan.flux_block_indices[b]].source];
                for (std::uint32_t j = src.first_flux; j < src.first_flux + src.flux_count; ++j) {
                    const std::uint32_t c = plan.panel_face_column[tree.flux_order[j]];
                    if (c != kNone && stamp[c] != epoch) {
                        stamp[c] = epoch;
                        cols.push_back(c);
                    }
                }
            }
            std::sort(cols.begin(), cols.end());
            for (std::uint32_t j = node.first_pressure; j < node.first_pressure + node.pressure_count; ++j)
                plan.flux.row_offsets[tree.pressure_order[j] + 1] = static_cast<std::uint32_t>(cols.size());
        }
    }
    exclusive_scan(plan.flux.row_offsets);
    plan.flux.columns.assign(plan.flux.row_offsets.back(), 0);
    for (std::size_t n = 0; n < tree.nodes.size(); ++n) {
        const SupportNode &node = tree.nodes[n];
        if (!node.leaf() || !node.pressure_count)
            continue;
        for (std::uint32_t j = node.first_pressure; j < node.first_pressure + node.pressure_count; ++j)
            std::copy(leaf_columns[n].begin(), leaf_columns[n].end(),
                      plan.flux.columns.begin() + plan.flux.row_offsets[tree.pressure_order[j]]);
    }

    // pressure CSR: sorted union of the near source leaves' P1 nodes
    std::vector<std::uint32_t> pp_offsets, pp_indices;
    group_by_target(tree, tree.pressure_pressure.near, pp_offsets, pp_indices);
    plan.pressure.row_offsets.assign(nv + 1, 0);
    for (std::size_t n = 0; n < tree.nodes.size(); ++n) {
        const SupportNode &node = tree.nodes[n];
        if (!node.leaf() || !node.pressure_count)
            continue;
        std::uint64_t count = 0;
        for (std::uint32_t b = pp_offsets[n]; b < pp_offsets[n + 1]; ++b)
            count += tree.nodes[tree.pressure_pressure.near[pp_indices[b]].source].pressure_count;
        if (count > nv)
            throw std::invalid_argument("near structure: a source node appears twice in a leaf");
        for (std::uint32_t j = node.first_pressure; j < node.first_pressure + node.pressure_count; ++j)
            plan.pressure.row_offsets[tree.pressure_order[j] + 1] = static_cast<std::uint32_t>(count);
    }
    exclusive_scan(plan.pressure.row_offsets);
    if (plan.pressure.row_offsets.back() != tree.pressure_pressure.represented_near)
        throw std::invalid_argument("near structure: row lengths do not add up to the pair count");
    plan.pressure.columns.assign(plan.pressure.row_offsets.back(), 0);
    std::vector<std::uint32_t> cols;
    for (std::size_t n = 0; n < tree.nodes.size(); ++n) {
        const SupportNode &node = tree.nodes[n];
        if (!node.leaf() || !node.pressure_count)
            continue;
        cols.clear();
        for (std::uint32_t b = pp_offsets[n]; b < pp_offsets[n + 1]; ++b) {
            const SupportNode &src = tree.nodes[tree.pressure_pressure.near[pp_indices[b]].source];
            cols.insert(cols.end(), tree.pressure_order.begin() + src.first_pressure,
                        tree.pressure_order.begin() + src.first_pressure + src.pressure_count);
        }
        std::sort(cols.begin(), cols.end());
        if (std::adjacent_find(cols.begin(), cols.end()) != cols.end())
            throw std::invalid_argument("near structure: a row repeats a column");
        for (std::uint32_t j = node.first_pressure; j < node.first_pressure + node.pressure_count; ++j)
            std::copy(cols.begin(), cols.end(), plan.pressure.columns.begin() + plan.pressure.row_offsets[tree.pressure_order[j]]);
    }

    // sorted unique panels of each leaf's P1 supports
    plan.node_panel_offsets.assign(tree.nodes.size() + 1, 0);
    plan.support_panels.reserve(space.incidences.size());
    std::vector<std::uint32_t> mark(nt, 0);
    for (std::size_t n = 0; n < tree.nodes.size(); ++n) {
        const SupportNode &node = tree.nodes[n];
        const std::size_t first = plan.support_panels.size();
        if (node.leaf())
            for (std::uint32_t j = node.first_pressure; j < node.first_pressure + node.pressure_count; ++j) {
                const std::uint32_t v = tree.pressure_order[j];
                for (std::size_t k = space.incidence_offsets[v]; k < space.incidence_offsets[v + 1]; ++k) {
                    const std::uint32_t p = space.incidences[k].panel;
                    if (mark[p] != n + 1) {
                        mark[p] = static_cast<std::uint32_t>(n + 1);
                        plan.support_panels.push_back(p);
                    }
                }
            }
        if (plan.support_panels.size() > UINT32_MAX)
            throw std::overflow_error("near structure: panel reference overflow");
        std::sort(plan.support_panels.begin() + static_cast<std::ptrdiff_t>(first), plan.support_panels.end());
        plan.node_panel_offsets[n + 1] = static_cast<std::uint32_t>(plan.support_panels.size());
    }
    return plan;
}

std::vector<SupportNearJob> support_near_jobs(const Space &space, const SupportTree &tree, const SupportNearPlan &plan) {
    const std::size_t nt = space.panels.size();
    if (plan.panel_face_column.size() != nt || plan.node_panel_offsets.size() != tree.nodes.size() + 1 ||
        plan.flux_block_offsets.size() != tree.nodes.size() + 1 ||
        plan.flux_block_indices.size() != tree.pressure_flux.near.size())
        throw std::invalid_argument("near structure: job list and mesh do not match");
    // test panels in contiguous ranges per thread, concatenated in order: the
    // result equals the single-threaded iterator order
    const unsigned workers = static_cast<unsigned>(
        std::max<std::size_t>(1, std::min<std::size_t>(std::thread::hardware_concurrency(), nt / 256 + 1)));
    std::vector<std::vector<SupportNearJob>> parts(workers);
    auto run = [&](unsigned w) {
        const std::uint32_t first = static_cast<std::uint32_t>(nt * w / workers),
                            last = static_cast<std::uint32_t>(nt * (w + 1) / workers);
        std::vector<SupportNearJob> &jobs = parts[w];
        std::vector<std::uint32_t> stamp(nt, 0), order;
        std::vector<std::uint16_t> mask(nt, 0);
        for (std::uint32_t t = first; t < last; ++t) {
            order.clear();
            auto touch = [&](std::uint32_t p) -> std::uint16_t & {
                if (stamp[p] != t + 1) {
                    stamp[p] = t + 1;
                    mask[p] = 0;
                    order.push_back(p);
                }
                return mask[p];
            };
            for (unsigned i = 0; i < 3; ++i) {
                const std::uint32_t v = space.panels[t].nodes[i];
                for (std::uint32_t j = plan.pressure.row_offsets[v]; j < plan.pressure.row_offsets[v + 1]; ++j) {
                    const std::uint32_t u = plan.pressure.columns[j];
                    for (std::size_t k = space.incidence_offsets[u]; k < space.incidence_offsets[u + 1]; ++k)
                        touch(space.incidences[k].panel) |=
                            static_cast<std::uint16_t>(1u << (3 * i + space.incidences[k].local));
                }
                const std::uint32_t leaf = tree.pressure_leaf[v];
                for (std::uint32_t b = plan.flux_block_offsets[leaf]; b < plan.flux_block_offsets[leaf + 1]; ++b) {
                    const SupportNode &src = tree.nodes[tree.pressure_flux.near[plan.flux_block_indices[b]].source];
                    for (std::uint32_t k = src.first_flux; k < src.first_flux + src.flux_count; ++k) {
                        const std::uint32_t p = tree.flux_order[k];
                        if (plan.panel_face_column[p] != kNone)
                            touch(p) |= static_cast<std::uint16_t>(1u << (9 + i));
                    }
                }
            }
            for (std::uint32_t p : order)
                jobs.push_back({t, p, static_cast<std::uint16_t>(mask[p] & 0x1ffu), static_cast<std::uint8_t>(mask[p] >> 9)});
        }
    };
    std::vector<std::thread> threads;
    for (unsigned w = 1; w < workers; ++w)
        threads.emplace_back(run, w);
    run(0);
    for (auto &th : threads)
        th.join();
    std::size_t total = 0;
    for (const auto &p : parts)
        total += p.size();
    std::vector<SupportNearJob> jobs;
    jobs.reserve(total);
    for (auto &p : parts) {
        jobs.insert(jobs.end(), p.begin(), p.end());
        std::vector<SupportNearJob>().swap(p);
    }
    return jobs;
}

} // namespace hsbem
