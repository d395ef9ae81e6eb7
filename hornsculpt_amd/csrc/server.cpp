#include "server.hpp"

#include <algorithm>
#include <sstream>

namespace hsbem {
namespace {

// True when a workspace sized for `held` has room for every buffer of `next`,
// so the slot can switch requests without freeing (hipFree synchronises the device).
// Buffers only grow, so `next` must need no buffer kind `held` lacks and no more
// bytes than `held` reserved: the same solver, form and preconditioner, Krylov
// bases no longer than held's, and every dimension within held's.
bool covers(const PreparedRequest &held, const PreparedRequest &next) {
    const Numerics &a = held.request.numerics, &b = next.request.numerics;
    const bool fmm = b.acceleration == 1;
    const bool same_kind = a.acceleration == b.acceleration && a.precision == b.precision &&
                           (fmm || a.solver == b.solver) && a.preconditioner == b.preconditioner &&
                           (!fmm || a.fmm_block_width == b.fmm_block_width);
    const bool krylov = fmm ? b.fmm_restart <= a.fmm_restart : b.gmres_restart <= a.gmres_restart;
    return same_kind && krylov && next.footprint <= held.footprint && next.nv <= held.nv && next.nt <= held.nt &&
           std::max(next.nd, 1u) <= std::max(held.nd, 1u) && next.ne <= held.ne &&
           next.observers.size() <= held.observers.size();
}

} // namespace

Server::Server(Context *ctx, ServerOptions options) : ctx_(ctx), options_(options) {
    if (options_.slots < 1 || options_.slots > 16)
        throw std::invalid_argument("Serving slots must be between 1 and 16.");
    hip_check(hipSetDevice(ctx_->device), "set device");
    if (!options_.memory_budget) {
        std::size_t free_bytes = 0, total_bytes = 0;
        hip_check(hipMemGetInfo(&free_bytes, &total_bytes), "device memory info");
        options_.memory_budget = static_cast<std::uint64_t>(0.9 * static_cast<double>(free_bytes));
    }
    stats_.budget = options_.memory_budget;
    slot_bytes_.assign(options_.slots, 0);
    std::vector<hipStream_t> streams(options_.slots);
    for (auto &s : streams)
        hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream create");
    for (unsigned i = 0; i < options_.slots; ++i)
        threads_.emplace_back([this, i, s = streams[i]] {
            // The workspace owns (and destroys) the stream.
            std::unique_ptr<Workspace> ws;
            try {
                hip_check(hipSetDevice(ctx_->device), "set device");
                ws = std::make_unique<Workspace>(s, true);
            } catch (const std::exception &e) {
                std::lock_guard lock(mutex_);
                start_error_ = e.what();
                (void)hipStreamDestroy(s);
                return;
            }
            worker_loop(i, *ws);
        });
}

Server::~Server() {
    {
        std::lock_guard lock(mutex_);
        stop_ = true;
        for (auto &[id, t] : tickets_)
            t->cancel = 1;
    }
    work_cv_.notify_all();
    for (auto &t : threads_)
        t.join();
}

std::uint64_t Server::submit(std::shared_ptr<MeshSolver> mesh, const Request &request) {
    if (!mesh)
        throw std::invalid_argument("BEM serving request needs a mesh.");
    auto prepared = mesh->prepare(request); // validation, as a direct solve
    if (prepared->footprint > options_.memory_budget) {
        std::ostringstream msg;
        msg << "BEM request needs " << prepared->footprint << " device bytes per frequency point; the serving budget is "
            << options_.memory_budget << " bytes.";
        throw std::invalid_argument(msg.str());
    }
    auto t = std::make_shared<Ticket>();
    t->mesh = std::move(mesh);
    t->prepared = std::move(prepared);
    const std::size_t n = request.frequencies.size();
    t->results.resize(n);
    t->remaining = n;
    t->submitted = Clock::now();
    std::uint64_t id;
    {
        std::lock_guard lock(mutex_);
        if (!start_error_.empty())
            throw DeviceError("BEM serving slots failed to start: " + start_error_);
        bool known = false;
        for (const auto &w : meshes_)
            known = known || w.lock() == t->mesh;
        if (!known)
            meshes_.push_back(t->mesh);
        id = t->id = next_ticket_++;
        tickets_.emplace(id, t);
        for (std::uint32_t i = 0; i < n; ++i)
            queue_.push_back({t, i});
        ++stats_.submitted;
        stats_.queued = static_cast<std::uint32_t>(queue_.size());
    }
    work_cv_.notify_all();
    return id;
}

std::uint64_t Server::resident_locked() {
    std::uint64_t total = 0;
    for (auto it = meshes_.begin(); it != meshes_.end();) {
        if (const auto m = it->lock()) {
            total += m->resident_bytes();
            ++it;
        } else {
            it = meshes_.erase(it);
        }
    }
    stats_.resident_bytes = total;
    return total;
}

void Server::fail(Ticket &t, std::exception_ptr e) {
    if (!t.error)
        t.error = e;
    t.cancel = 1; // points of this request still running stop at their next check
}

void Server::worker_loop(unsigned slot, Workspace &ws) {
    std::shared_ptr<PreparedRequest> held; // request whose exact footprint the workspace was sized for
    for (;;) {
        Item item;
        bool release = false;
        {
            std::unique_lock lock(mutex_);
            for (;;) {
                if (stop_)
                    return;
                // Drop points of failed or cancelled requests; take the first point
                // that fits the budget (arrival order, smaller points may pass a
                // blocked larger one).
                bool found = false;
                const std::uint64_t resident = resident_locked();
                for (auto it = queue_.begin(); it != queue_.end();) {
                    Ticket &t = *it->ticket;
                    if (t.error || t.cancel) {
                        if (!t.error)
                            t.error = std::make_exception_ptr(Cancelled());
                        --t.remaining;
                        it = queue_.erase(it);
                        done_cv_.notify_all();
                        continue;
                    }
                    // the slot's workspace, the other slots' and the meshes' resident state
                    // (plus this point's plan if it is not built yet) must fit the budget
                    const std::uint64_t need = t.prepared->footprint;
                    const std::uint64_t state = resident + t.mesh->pending_plan_bytes(*t.prepared);
                    const bool reuse = held && covers(*held, *t.prepared);
                    const std::uint64_t total = (reuse ? held_ : held_ - slot_bytes_[slot] + need) + state;
                    if (total <= options_.memory_budget) {
                        stats_.peak_total_bytes = std::max(stats_.peak_total_bytes, total);
                        item = *it;
                        queue_.erase(it);
                        found = true;
                        break;
                    }
                    ++it;
                }
                if (found)
                    break;
                if (!queue_.empty() && !stats_.running && !held_) {
                    // Nothing runs and no slot holds memory, yet no point fits: drop the
                    // meshes' unused cached plans; if that frees nothing, the front
                    // request cannot be served under this budget.
                    std::uint64_t released = 0;
                    for (const auto &w : meshes_)
                        if (const auto m = w.lock())
                            released += m->trim_plans(0);
                    if (released)
                        continue;
                    std::ostringstream msg;
                    msg << "BEM serving budget of " << options_.memory_budget
                        << " device bytes cannot hold the resident mesh state (" << resident
                        << " bytes) and one frequency point of this request.";
                    Ticket &t = *queue_.front().ticket;
                    fail(t, std::make_exception_ptr(std::runtime_error(msg.str())));
                    continue;
                }
                if (!queue_.empty() && slot_bytes_[slot]) {
                    release = true; // an idle slot gives its memory back to blocked points
                    break;
                }
                stats_.queued = static_cast<std::uint32_t>(queue_.size());
                work_cv_.wait(lock);
            }
            if (!release) {
                Ticket &t = *item.ticket;
                if (!(held && covers(*held, *t.prepared))) {
                    held_ = held_ - slot_bytes_[slot] + t.prepared->footprint;
                    slot_bytes_[slot] = t.prepared->footprint;
                    stats_.peak_held_bytes = std::max(stats_.peak_held_bytes, held_);
                }
                ++t.running;
                ++stats_.running;
                stats_.max_running = std::max(stats_.max_running, stats_.running);
                stats_.queued = static_cast<std::uint32_t>(queue_.size());
                stats_.held_bytes = held_;
            }
        }
        if (release) {
            try {
                hip_check(hipStreamSynchronize(ws.stream), "stream synchronize");
                ws.release();
            } catch (...) {
            }
            held.reset();
            {
                std::lock_guard lock(mutex_);
                held_ -= slot_bytes_[slot];
                slot_bytes_[slot] = 0;
                stats_.held_bytes = held_;
            }
            work_cv_.notify_all();
            continue;
        }
        Ticket &t = *item.ticket;
        std::exception_ptr error;
        FrequencyResult fr;
        try {
            if (!(held && covers(*held, *t.prepared))) {
                ws.release(); // the slot then holds exactly this request's footprint
                held = t.prepared;
            }
            fr = t.mesh->solve_frequency(*t.prepared, item.index, ws, &t.cancel);
            fr.served_ms = std::chrono::duration<double, std::milli>(Clock::now() - t.submitted).count();
        } catch (...) {
            error = std::current_exception();
        }
        {
            std::lock_guard lock(mutex_);
            --t.running;
            --t.remaining;
            --stats_.running;
            if (error)
                fail(t, error);
            else {
                t.results[item.index] = std::move(fr);
                ++stats_.points;
            }
        }
        done_cv_.notify_all();
        work_cv_.notify_all();
    }
}

bool Server::poll(std::uint64_t ticket, int wait_ms, Result *out) {
    std::unique_lock lock(mutex_);
    const auto it = tickets_.find(ticket);
    if (it == tickets_.end())
        throw std::invalid_argument("Unknown BEM serving ticket.");
    const std::shared_ptr<Ticket> t = it->second;
    auto finished = [&] { return t->running == 0 && (t->remaining == 0 || (t->error && !t->running)); };
    // A failed request is finished once none of its points is still running;
    // its queued points are dropped by the slots or here.
    auto settle = [&] {
        if (t->error && !t->running) {
            for (auto q = queue_.begin(); q != queue_.end();)
                if (q->ticket == t) {
                    --t->remaining;
                    q = queue_.erase(q);
                } else
                    ++q;
        }
        return finished();
    };
    if (!settle()) {
        if (wait_ms == 0)
            return false;
        if (wait_ms < 0)
            done_cv_.wait(lock, settle);
        else if (!done_cv_.wait_for(lock, std::chrono::milliseconds(wait_ms), settle))
            return false;
    }
    tickets_.erase(ticket);
    stats_.queued = static_cast<std::uint32_t>(queue_.size());
    if (t->error) {
        ++stats_.failed;
        std::rethrow_exception(t->error);
    }
    ++stats_.completed;
    Result r;
    r.nt = t->mesh->nt();
    r.nv = t->mesh->nv();
    r.nobs = static_cast<std::uint32_t>(t->prepared->observers.size());
    r.nexc = t->prepared->ne;
    r.peak_bytes = t->prepared->footprint;
    r.frequencies = std::move(t->results);
    *out = std::move(r);
    return true;
}

void Server::cancel(std::uint64_t ticket) {
    {
        std::lock_guard lock(mutex_);
        const auto it = tickets_.find(ticket);
        if (it == tickets_.end())
            throw std::invalid_argument("Unknown BEM serving ticket.");
        fail(*it->second, std::make_exception_ptr(Cancelled()));
    }
    work_cv_.notify_all();
    done_cv_.notify_all();
}

ServerStats Server::stats() const {
    std::lock_guard lock(mutex_);
    return stats_;
}

} // namespace hsbem
