// Bounded repository of immutable translation plans shared by the FMM operators of
// one mesh (a solve, its audit and comparison operators, later frequencies).
//
// Replaces TranslationPlanCache (modules/bem/fmm/translation_plan_cache.hpp:23-42,
// translation_plan_cache.cpp): coaxial plans keyed by the exact request (kind,
// orders, polar count and the bit patterns of k, radii, distance and probe radius),
// least-recently-used eviction above a byte budget (512 MB by default), hit and
// build statistics. The Wigner blocks of rotations depend on the order and the polar
// angle only (every direction of one polar class shares them), so they share the
// repository under their own keys and are reused across wavenumbers. Plans are built outside the lock; two threads
// asking for the same missing plan may both build it, the first stored one wins.
#pragma once

#include "fmm_rotation.hpp"
#include "fmm_scaled.hpp"

#include <array>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>

namespace hsbem {

struct AxialPlanRequest {
    AxialKind kind{};
    unsigned source_order{}, target_order{}, polar_count{};
    double k{}, source_radius{}, target_radius{}, distance{}, projection_radius{};
};

struct TranslationPlanCacheStatistics {
    std::size_t hits{}, plans_built{}, evictions{}, live_bytes{}, peak_bytes{};
    double builder_ms{}; // summed over builder calls (parallel builds can exceed wall time)
};

class TranslationPlanCache {
  public:
    explicit TranslationPlanCache(std::size_t maximum_bytes = std::size_t{1} << 30);
    std::shared_ptr<const AxialPlan> axial(const AxialPlanRequest &request);
    // Wigner blocks of one polar class (the plan's azimuth is left empty).
    std::shared_ptr<const RotationPlan> wigner(unsigned order, double theta);
    TranslationPlanCacheStatistics statistics() const;

  private:
    using Key = std::array<std::uint64_t, 10>;
    struct Entry {
        std::shared_ptr<const AxialPlan> axial;
        std::shared_ptr<const RotationPlan> rotation;
        std::size_t bytes{}, touch{};
    };
    void insert(const Key &key, Entry entry);
    mutable std::mutex mutex_;
    std::map<Key, Entry> entries_;
    std::size_t limit_{}, clock_{};
    TranslationPlanCacheStatistics stats_;
};

} // namespace hsbem
