// Launch-shape autotuner shared by the kernels that have more than one shape.
//
// A tunable (a kernel's launch shape: the regular tiles' trial block, the FMM
// translations' stage-split target, the GMRES product's column chunk and rows per
// lane) is chosen per key (a problem-size class: mesh, precision, operator levels,
// matrix order). The first request of a (tunable, key) times every candidate on
// the caller's stream and data with HIP events — one warm run, then `repeats`
// timed runs — and keeps the fastest; later requests return the stored choice.
// A pinned value (from the tunable's environment variable, parsed by the caller) is
// returned as it is, without timing and without being stored. While the stream is
// being captured into a graph, nothing is timed: the fallback is returned and nothing
// is stored.
//
// Thread safety: the table is guarded by one mutex held only for lookups and
// inserts; each entry has its own mutex, so a tuning run blocks only callers of
// the same (tunable, key), which then take the stored result.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <optional>
#include <string>
#include <vector>

namespace hsbem::autotune {

struct Timing {
    std::uint32_t candidate;
    float ms; // mean of the timed runs; negative when the timing failed
};

struct Record {
    std::string tunable, key;
    std::uint32_t chosen{};
    std::vector<Timing> timings;
};

// The choice of `tunable` for `key` among `candidates` (non-empty); `run(c)` enqueues
// one unit of the work with candidate c on `stream`.
std::uint32_t choose(const std::string &tunable, const std::string &key, const std::vector<std::uint32_t> &candidates,
                     std::uint32_t fallback, hipStream_t stream, const std::function<void(std::uint32_t)> &run,
                     int repeats = 2, std::optional<std::uint32_t> pinned = std::nullopt);

// The stored choice, if any, without tuning.
std::optional<std::uint32_t> lookup(const std::string &tunable, const std::string &key);

// Every decision taken so far in this process, in the order they were taken.
std::vector<Record> records();

// The decisions as a JSON array of {"tunable", "key", "chosen", "timings": [[c, ms], ...]}.
std::string records_json();

// Integer value of environment variable `name` when it is set to one of `allowed`.
std::optional<std::uint32_t> pinned_from_env(const char *name, const std::vector<std::uint32_t> &allowed);

} // namespace hsbem::autotune
