// Serving loop: continuous batching of (request, frequency) points over several
// concurrent solve slots on one GPU.
//
// The reference serves design sweeps and live edits by calling
// CudaBemSolver::solve() from worker threads, one solve per mesh at a time
// (apps/desktop/simulation_dock.cpp:2425-2600, apps/desktop/parameter_sweep.cpp:215-250,
// cuda_solver.cu:1024). Here a request is split into its frequency points; slots
// (each a thread with its own HIP stream and Workspace) take points in arrival
// order as soon as a slot is free and the point's device footprint fits the
// memory budget, so points of requests that arrive while others run are
// admitted immediately and their kernels overlap on the device. Mesh state
// (geometry, clusters, interaction plans, FMM near structures) stays resident in
// the MeshSolver, is shared by every slot and counts against the same budget.
#pragma once

#include "solver.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <thread>

namespace hsbem {

struct ServerOptions {
    unsigned slots{3};             // concurrent frequency points (streams)
    std::uint64_t memory_budget{}; // device bytes for slot workspaces; 0: 90 % of free memory
};

struct ServerStats {
    std::uint64_t submitted{}, completed{}, failed{}, points{};
    std::uint32_t running{}, queued{}, max_running{};
    std::uint64_t budget{}, held_bytes{}, peak_held_bytes{};
    std::uint64_t resident_bytes{}; // device bytes meshes keep between points (geometry, plans, near structures)
    std::uint64_t peak_total_bytes{}; // largest held_bytes + resident_bytes seen at an admission
};

class Server {
  public:
    Server(Context *ctx, ServerOptions options);
    ~Server();
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;

    // Validates the request (throws std::invalid_argument like a direct solve, or
    // when one frequency point cannot fit the memory budget) and queues its points.
    std::uint64_t submit(std::shared_ptr<MeshSolver> mesh, const Request &request);
    // Waits up to wait_ms (0: no wait, negative: until done). Returns true with the
    // result when every point is done; false while pending. A failed or cancelled
    // request rethrows its first exception. A finished ticket is forgotten.
    bool poll(std::uint64_t ticket, int wait_ms, Result *out);
    void cancel(std::uint64_t ticket);
    ServerStats stats() const;

  private:
    using Clock = std::chrono::steady_clock;
    struct Ticket {
        std::uint64_t id{};
        std::shared_ptr<MeshSolver> mesh;
        std::shared_ptr<PreparedRequest> prepared;
        std::vector<FrequencyResult> results;
        std::size_t remaining{}, running{};
        std::exception_ptr error;
        volatile std::int32_t cancel{};
        Clock::time_point submitted;
    };
    struct Item {
        std::shared_ptr<Ticket> ticket;
        std::uint32_t index{};
    };
    void worker_loop(unsigned slot, Workspace &ws);
    void fail(Ticket &t, std::exception_ptr e);
    // Device bytes resident in the meshes this server has seen (mutex_ held).
    std::uint64_t resident_locked();

    Context *ctx_;
    ServerOptions options_;
    mutable std::mutex mutex_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<Item> queue_;
    std::vector<std::weak_ptr<MeshSolver>> meshes_; // meshes whose resident state counts against the budget
    std::map<std::uint64_t, std::shared_ptr<Ticket>> tickets_;
    std::vector<std::uint64_t> slot_bytes_; // footprint reserved by each slot's workspace
    std::uint64_t held_{};
    std::vector<std::thread> threads_;
    bool stop_{};
    std::uint64_t next_ticket_{1};
    ServerStats stats_;
    std::string start_error_;
};

} // namespace hsbem
