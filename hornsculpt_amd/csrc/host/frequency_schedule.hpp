// Progressive order of a frequency raster for realtime solves
// (apps/desktop/frequency_schedule.hpp:11-32).
#pragma once

#include <cstddef>
#include <vector>

namespace hsbem {

// Indices 0..count-1 in solve order: the highest, then the lowest, then the
// midpoint of the widest remaining index gap, repeatedly; among gaps of equal
// width the one higher in the raster goes first.
std::vector<std::size_t> frequency_schedule(std::size_t count);

} // namespace hsbem
