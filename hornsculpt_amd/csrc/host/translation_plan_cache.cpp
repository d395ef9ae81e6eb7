#include "translation_plan_cache.hpp"

#include <bit>
#include <chrono>
#include <stdexcept>

namespace hsbem {
namespace {

std::uint64_t bits(double v) { return std::bit_cast<std::uint64_t>(v); }

std::size_t plan_bytes(const AxialPlan &p) {
    return sizeof(AxialPlan) + p.coefficients.capacity() * sizeof(std::complex<double>) +
           p.block_offsets.capacity() * sizeof(std::size_t);
}

std::size_t plan_bytes(const RotationPlan &p) {
    return sizeof(RotationPlan) + p.small_d.capacity() * sizeof(double) +
           p.azimuth.capacity() * sizeof(std::complex<double>) + p.degree_offsets.capacity() * sizeof(std::size_t);
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

} // namespace

TranslationPlanCache::TranslationPlanCache(std::size_t maximum_bytes) : limit_(maximum_bytes) {
    if (!limit_)
        throw std::invalid_argument("Translation plan cache needs a positive byte budget.");
}

void TranslationPlanCache::insert(const Key &key, Entry entry) {
    // caller holds the lock; the new entry is never the one evicted
    entry.touch = ++clock_;
    stats_.live_bytes += entry.bytes;
    entries_.emplace(key, std::move(entry));
    while (stats_.live_bytes > limit_ && entries_.size() > 1) {
        auto oldest = entries_.end();
        for (auto it = entries_.begin(); it != entries_.end(); ++it)
            if (it->first != key && (oldest == entries_.end() || it->second.touch < oldest->second.touch))
                oldest = it;
        if (oldest == entries_.end())
            break;
        stats_.live_bytes -= oldest->second.bytes;
        entries_.erase(oldest);
        ++stats_.evictions;
    }
    stats_.peak_bytes = std::max(stats_.peak_bytes, stats_.live_bytes);
}

std::shared_ptr<const AxialPlan> TranslationPlanCache::axial(const AxialPlanRequest &r) {
    const Key key{0,
                  static_cast<std::uint64_t>(r.kind),
                  r.source_order,
                  r.target_order,
                  r.polar_count,
                  bits(r.k),
                  bits(r.source_radius),
                  bits(r.target_radius),
                  bits(r.distance),
                  bits(r.projection_radius)};
    {
        std::lock_guard lock(mutex_);
        const auto it = entries_.find(key);
        if (it != entries_.end()) {
            ++stats_.hits;
            it->second.touch = ++clock_;
            return it->second.axial;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto plan = std::make_shared<const AxialPlan>(make_axial_plan(r.kind, r.source_order, r.target_order, r.k,
                                                                  r.source_radius, r.target_radius, r.distance,
                                                                  r.polar_count, r.projection_radius));
    const double built = ms_since(t0);
    std::lock_guard lock(mutex_);
    ++stats_.plans_built;
    stats_.builder_ms += built;
    const auto it = entries_.find(key);
    if (it != entries_.end()) { // built concurrently: keep the stored one
        it->second.touch = ++clock_;
        return it->second.axial;
    }
    insert(key, Entry{plan, nullptr, plan_bytes(*plan), 0});
    return plan;
}

std::shared_ptr<const RotationPlan> TranslationPlanCache::wigner(unsigned order, double theta) {
    const Key key{1, order, bits(theta), 0, 0, 0, 0, 0, 0, 0};
    {
        std::lock_guard lock(mutex_);
        const auto it = entries_.find(key);
        if (it != entries_.end()) {
            ++stats_.hits;
            it->second.touch = ++clock_;
            return it->second.rotation;
        }
    }
    const auto t0 = std::chrono::steady_clock::now();
    auto plan = std::make_shared<const RotationPlan>(wigner_blocks(order, theta));
    const double built = ms_since(t0);
    std::lock_guard lock(mutex_);
    ++stats_.plans_built;
    stats_.builder_ms += built;
    const auto it = entries_.find(key);
    if (it != entries_.end()) {
        it->second.touch = ++clock_;
        return it->second.rotation;
    }
    insert(key, Entry{nullptr, plan, plan_bytes(*plan), 0});
    return plan;
}

TranslationPlanCacheStatistics TranslationPlanCache::statistics() const {
    std::lock_guard lock(mutex_);
    return stats_;
}

} // namespace hsbem
