// Launch-shape autotuner: timing of candidates with HIP events, one decision per
// (tunable, key), kept for the life of the process.
#include "autotune.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

namespace hsbem::autotune {
namespace {

struct Entry {
    std::mutex mutex;
    bool done{};
    Record record;
    std::uint64_t order{};
};

std::mutex table_mutex;
std::map<std::pair<std::string, std::string>, std::shared_ptr<Entry>> table;
std::uint64_t next_order = 0;

std::shared_ptr<Entry> entry_of(const std::string &tunable, const std::string &key) {
    std::lock_guard lock(table_mutex);
    auto &e = table[{tunable, key}];
    if (!e)
        e = std::make_shared<Entry>();
    return e;
}

// Mean milliseconds of `repeats` runs of candidate c after one warm run; negative on
// a failed event.
float time_candidate(std::uint32_t c, hipStream_t stream, const std::function<void(std::uint32_t)> &run,
                     int repeats, hipEvent_t start, hipEvent_t stop) {
    run(c);
    if (hipEventRecord(start, stream) != hipSuccess)
        return -1;
    for (int r = 0; r < repeats; ++r)
        run(c);
    if (hipEventRecord(stop, stream) != hipSuccess)
        return -1;
    float ms = 0;
    if (hipEventSynchronize(stop) != hipSuccess || hipEventElapsedTime(&ms, start, stop) != hipSuccess)
        return -1;
    return ms / static_cast<float>(repeats);
}

} // namespace

std::uint32_t choose(const std::string &tunable, const std::string &key, const std::vector<std::uint32_t> &candidates,
                     std::uint32_t fallback, hipStream_t stream, const std::function<void(std::uint32_t)> &run,
                     int repeats, std::optional<std::uint32_t> pinned) {
    if (pinned)
        return *pinned; // bypasses the table: a later unpinned call still tunes
    const auto e = entry_of(tunable, key);
    std::lock_guard lock(e->mutex);
    if (e->done)
        return e->record.chosen;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    if (candidates.empty() || hipStreamIsCapturing(stream, &capture) != hipSuccess ||
        capture != hipStreamCaptureStatusNone)
        return fallback;
    Record rec;
    rec.tunable = tunable;
    rec.key = key;
    rec.chosen = candidates[0];
    if (candidates.size() > 1) {
        hipEvent_t start = nullptr, stop = nullptr;
        if (hipEventCreate(&start) != hipSuccess)
            return fallback;
        if (hipEventCreate(&stop) != hipSuccess) {
            (void)hipEventDestroy(start);
            return fallback;
        }
        float best = -1;
        rec.chosen = fallback;
        for (const std::uint32_t c : candidates) {
            const float ms = time_candidate(c, stream, run, std::max(repeats, 1), start, stop);
            rec.timings.push_back({c, ms});
            if (ms >= 0 && (best < 0 || ms < best)) {
                best = ms;
                rec.chosen = c;
            }
        }
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
    }
    {
        std::lock_guard table_lock(table_mutex);
        e->order = next_order++;
    }
    e->record = std::move(rec);
    e->done = true;
    return e->record.chosen;
}

std::optional<std::uint32_t> lookup(const std::string &tunable, const std::string &key) {
    std::shared_ptr<Entry> e;
    {
        std::lock_guard lock(table_mutex);
        const auto it = table.find({tunable, key});
        if (it == table.end())
            return std::nullopt;
        e = it->second;
    }
    std::lock_guard lock(e->mutex);
    if (!e->done)
        return std::nullopt;
    return e->record.chosen;
}

std::vector<Record> records() {
    std::vector<std::pair<std::uint64_t, std::shared_ptr<Entry>>> entries;
    {
        std::lock_guard lock(table_mutex);
        for (const auto &[k, e] : table)
            entries.emplace_back(0, e);
    }
    std::vector<std::pair<std::uint64_t, Record>> done;
    for (auto &[o, e] : entries) {
        std::lock_guard lock(e->mutex);
        if (e->done)
            done.emplace_back(e->order, e->record);
    }
    std::sort(done.begin(), done.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    std::vector<Record> out;
    for (auto &d : done)
        out.push_back(std::move(d.second));
    return out;
}

std::string records_json() {
    auto quoted = [](const std::string &s) {
        std::string q = "\"";
        for (const char ch : s) {
            if (ch == '"' || ch == '\\')
                q += '\\';
            q += ch;
        }
        return q + "\"";
    };
    std::ostringstream o;
    o << "[";
    bool first = true;
    for (const auto &r : records()) {
        o << (first ? "" : ", ") << "{\"tunable\": " << quoted(r.tunable) << ", \"key\": " << quoted(r.key)
          << ", \"chosen\": " << r.chosen << ", \"timings\": [";
        for (std::size_t i = 0; i < r.timings.size(); ++i)
            o << (i ? ", " : "") << "[" << r.timings[i].candidate << ", " << r.timings[i].ms << "]";
        o << "]}";
        first = false;
    }
    o << "]";
    return o.str();
}

std::optional<std::uint32_t> pinned_from_env(const char *name, const std::vector<std::uint32_t> &allowed) {
    const char *v = std::getenv(name);
    if (!v || !*v)
        return std::nullopt;
    char *end = nullptr;
    const unsigned long x = std::strtoul(v, &end, 10);
    if (!end || *end)
        return std::nullopt;
    for (const std::uint32_t a : allowed)
        if (a == x)
            return a;
    return std::nullopt;
}

} // namespace hsbem::autotune
