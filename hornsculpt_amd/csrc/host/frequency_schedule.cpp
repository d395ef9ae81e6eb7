#include "frequency_schedule.hpp"

#include <algorithm>

namespace hsbem {

namespace {

struct Gap {
    std::size_t low, high;
    std::size_t width() const { return high - low; }
};

// Heap order: wider first; equal widths: higher in the raster first.
bool before(const Gap &a, const Gap &b) {
    if (a.width() != b.width())
        return a.width() < b.width();
    return a.low < b.low;
}

} // namespace

std::vector<std::size_t> frequency_schedule(std::size_t count) {
    std::vector<std::size_t> order;
    if (count == 0)
        return order;
    order.reserve(count);
    order.push_back(count - 1);
    if (count == 1)
        return order;
    order.push_back(0);
    std::vector<Gap> heap{{0, count - 1}};
    while (!heap.empty()) {
        std::pop_heap(heap.begin(), heap.end(), before);
        const Gap g = heap.back();
        heap.pop_back();
        if (g.width() < 2)
            continue;
        const std::size_t mid = g.low + g.width() / 2;
        order.push_back(mid);
        for (const Gap half : {Gap{g.low, mid}, Gap{mid, g.high}}) {
            heap.push_back(half);
            std::push_heap(heap.begin(), heap.end(), before);
        }
    }
    return order;
}

} // namespace hsbem
