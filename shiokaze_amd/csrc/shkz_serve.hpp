// shkz_serve.hpp -- request queue, HBM-budget admission and continuous batching of the projection
// server (host C++ only, no HIP: the executor is a callback, so the same scheduler runs against a
// stand-in executor in the CPU tests).
//
// Continuous batching: every executor (one HIP stream and one device workspace in the server) takes
// a new batch from the queue as soon as it is free, so requests that arrive while others are on
// the GPU join the next batch that any executor forms, and batches of different executors overlap
// on the device.  Batch formation walks the queue in arrival order and
//   * keeps each scene's requests in order: a request whose scene is in flight (or that follows a
//     request of the same scene that was passed over) waits;
//   * batches only requests of one group (requests with equal solver parameters);
//   * stops at the first request that would exceed the device budget: the estimated working set of
//     all batches in flight plus the resident scene state must stay within the budget.  Memory is
//     admitted in arrival order, so a large request is never starved by smaller ones behind it;
//   * caps a batch at max_batch_grids grids and max_batch_cells box cells.
// A request whose estimate cannot fit even on an idle device fails with SHKZ_ERR_BUDGET.  A request
// still queued can be cancelled (it completes with SHKZ_ERR_CANCELLED); one that started runs to
// its end.
#pragma once

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace shkz {

struct ServeRequest {
    uint64_t ticket = 0;
    uint64_t scene_id = 0;   // 0 = no scene (no ordering constraint)
    uint64_t group = 0;      // requests batch together only within one group
    uint64_t est_bytes = 0;  // estimated device working set
    int64_t cells = 0;       // box cells
    double t_submit_ms = 0;
    void *payload = nullptr;  // owned by the caller of the scheduler
};

struct ServeOutcome {
    int status = 0;
    int batch_size = 0;
    int executor = -1;
    double t_submit_ms = 0, t_start_ms = 0, t_end_ms = 0;
    void *payload = nullptr;
};

struct ServeLimits {
    int executors = 2;
    int max_batch_grids = 64;
    int64_t max_batch_cells = (int64_t)64 << 20;
    uint64_t budget = 0;
};

struct ServeStats {
    uint64_t submitted = 0, completed = 0, rejected = 0, batches = 0, batched_requests = 0, cancelled = 0;
    int max_batch = 0;
    int max_in_flight_batches = 0;
    uint64_t max_in_flight_bytes = 0;
};

static inline double serve_now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

class ServeCore {
public:
    // run(executor, batch, status): runs one batch synchronously on `executor` and sets status[i]
    // of each request.  resident(): bytes held outside the batches (resident scene state).
    using RunBatch = std::function<void(int, std::vector<ServeRequest> &, std::vector<int> &)>;
    using Resident = std::function<uint64_t()>;

    ServeCore(const ServeLimits &lim, RunBatch run, Resident resident)
        : lim_(lim), run_(std::move(run)), resident_(std::move(resident)) {
        if (lim_.executors < 1) lim_.executors = 1;
        if (lim_.max_batch_grids < 1) lim_.max_batch_grids = 1;
        if (lim_.max_batch_cells < 1) lim_.max_batch_cells = 1;
        for (int k = 0; k < lim_.executors; ++k) threads_.emplace_back([this, k] { executor_loop(k); });
    }

    ~ServeCore() { shutdown(); }

    // Finishes every queued and running request, then stops the executors.
    void shutdown() {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stopping_) return;
            stopping_ = true;
        }
        cv_work_.notify_all();
        for (std::thread &t : threads_) t.join();
        threads_.clear();
    }

    // Queues a request; returns 0, or -1 when the scheduler is shutting down.
    int submit(ServeRequest r) {
        {
            std::lock_guard<std::mutex> g(mu_);
            if (stopping_) return -1;
            r.t_submit_ms = serve_now_ms();
            state_[r.ticket] = Entry{r, ServeOutcome{}, false};
            queue_.push_back(r.ticket);
            ++stats_.submitted;
        }
        cv_work_.notify_all();
        return 0;
    }

    // 1 = still queued or running, 0 = done (outcome filled, ticket forgotten), -5 = unknown ticket.
    int poll(uint64_t ticket, ServeOutcome *out) {
        std::lock_guard<std::mutex> g(mu_);
        return take_locked(ticket, out);
    }

    // Waits up to timeout_ms (< 0: forever) for the ticket; returns like poll().
    int wait(uint64_t ticket, double timeout_ms, ServeOutcome *out) {
        std::unique_lock<std::mutex> lk(mu_);
        auto ready = [&] {
            auto it = state_.find(ticket);
            return it == state_.end() || it->second.done;
        };
        if (timeout_ms < 0) cv_done_.wait(lk, ready);
        else cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
        return take_locked(ticket, out);
    }

    // Cancels a request that is still queued: it completes at once with status -7 (cancelled) and
    // is collected by poll / wait like any other.  Returns 0 when cancelled, 1 when the request
    // already started or completed (it runs to the end; its outcome is unchanged), -5 for an
    // unknown ticket.  Later requests of the same scene no longer wait for it.
    int cancel(uint64_t ticket) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto it = state_.find(ticket);
            if (it == state_.end()) return -5;
            auto q = std::find(queue_.begin(), queue_.end(), ticket);
            if (q == queue_.end()) return 1;
            queue_.erase(q);
            const double t = serve_now_ms();
            finish_locked(it->second.req, -7 /* SHKZ_ERR_CANCELLED */, 0, -1, t, t);
            ++stats_.cancelled;
        }
        cv_done_.notify_all();
        cv_work_.notify_all();  // requests of its scene may start now
        return 0;
    }

    // A scene with a queued or running request.
    bool scene_busy(uint64_t scene_id) {
        std::lock_guard<std::mutex> g(mu_);
        if (in_flight_scenes_.count(scene_id)) return true;
        for (uint64_t t : queue_)
            if (state_[t].req.scene_id == scene_id) return true;
        return false;
    }

    ServeStats stats() {
        std::lock_guard<std::mutex> g(mu_);
        return stats_;
    }

    // Takes the outcomes nobody collected (after shutdown()).
    std::vector<ServeOutcome> take_all() {
        std::lock_guard<std::mutex> g(mu_);
        std::vector<ServeOutcome> out;
        for (auto it = state_.begin(); it != state_.end();) {
            if (it->second.done) {
                out.push_back(it->second.out);
                it = state_.erase(it);
            } else {
                ++it;
            }
        }
        return out;
    }

    int pending() {
        std::lock_guard<std::mutex> g(mu_);
        return (int)queue_.size();
    }

private:
    struct Entry {
        ServeRequest req;
        ServeOutcome out;
        bool done;
    };

    int take_locked(uint64_t ticket, ServeOutcome *out) {
        auto it = state_.find(ticket);
        if (it == state_.end()) return -5;
        if (!it->second.done) return 1;
        if (out) *out = it->second.out;
        state_.erase(it);
        return 0;
    }

    void finish_locked(const ServeRequest &r, int status, int batch, int executor, double t0, double t1) {
        Entry &e = state_[r.ticket];
        e.done = true;
        e.out.status = status;
        e.out.batch_size = batch;
        e.out.executor = executor;
        e.out.t_submit_ms = r.t_submit_ms;
        e.out.t_start_ms = t0;
        e.out.t_end_ms = t1;
        e.out.payload = r.payload;
        ++stats_.completed;
    }

    // Forms the next batch (under the lock); empty when nothing can start now.  Requests that can
    // never fit are completed with the budget error.
    std::vector<ServeRequest> form_batch_locked(uint64_t &bytes) {
        std::vector<ServeRequest> batch;
        bytes = 0;
        if (queue_.empty()) return batch;
        const uint64_t resident = resident_ ? resident_() : 0;
        std::unordered_set<uint64_t> passed, in_batch;
        int64_t cells = 0;
        bool have_group = false;
        uint64_t group = 0;
        for (auto it = queue_.begin(); it != queue_.end();) {
            const ServeRequest &r = state_[*it].req;
            const bool busy = r.scene_id && (in_flight_scenes_.count(r.scene_id) || passed.count(r.scene_id) ||
                                             in_batch.count(r.scene_id));
            if (busy || (have_group && r.group != group)) {
                if (r.scene_id) passed.insert(r.scene_id);
                ++it;
                continue;
            }
            // a request that cannot fit even alone on an idle device fails now
            if (r.est_bytes + resident > lim_.budget) {
                const double t = serve_now_ms();
                finish_locked(r, -4 /* SHKZ_ERR_BUDGET */, 0, -1, t, t);
                ++stats_.rejected;
                it = queue_.erase(it);
                cv_done_.notify_all();
                continue;
            }
            if (in_flight_bytes_ + bytes + r.est_bytes + resident > lim_.budget) break;  // memory in arrival order
            if ((int)batch.size() >= lim_.max_batch_grids) break;
            if (!batch.empty() && cells + r.cells > lim_.max_batch_cells) break;
            batch.push_back(r);
            bytes += r.est_bytes;
            cells += r.cells;
            if (r.scene_id) in_batch.insert(r.scene_id);
            have_group = true;
            group = r.group;
            it = queue_.erase(it);
        }
        return batch;
    }

    void executor_loop(int k) {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            uint64_t bytes = 0;
            std::vector<ServeRequest> batch = form_batch_locked(bytes);
            if (batch.empty()) {
                if (stopping_ && queue_.empty()) break;
                cv_work_.wait(lk);
                continue;
            }
            for (const ServeRequest &r : batch)
                if (r.scene_id) in_flight_scenes_.insert(r.scene_id);
            in_flight_bytes_ += bytes;
            ++in_flight_batches_;
            ++stats_.batches;
            stats_.batched_requests += batch.size();
            if ((int)batch.size() > stats_.max_batch) stats_.max_batch = (int)batch.size();
            if (in_flight_batches_ > stats_.max_in_flight_batches) stats_.max_in_flight_batches = in_flight_batches_;
            if (in_flight_bytes_ > stats_.max_in_flight_bytes) stats_.max_in_flight_bytes = in_flight_bytes_;
            lk.unlock();
            const double t0 = serve_now_ms();
            std::vector<int> status(batch.size(), 0);
            run_(k, batch, status);
            const double t1 = serve_now_ms();
            lk.lock();
            for (size_t i = 0; i < batch.size(); ++i) {
                finish_locked(batch[i], status[i], (int)batch.size(), k, t0, t1);
                if (batch[i].scene_id) in_flight_scenes_.erase(batch[i].scene_id);
            }
            in_flight_bytes_ -= bytes;
            --in_flight_batches_;
            cv_done_.notify_all();
            cv_work_.notify_all();  // scenes and memory freed: other executors may form batches
        }
    }

    ServeLimits lim_;
    RunBatch run_;
    Resident resident_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<uint64_t> queue_;
    std::unordered_map<uint64_t, Entry> state_;
    std::unordered_set<uint64_t> in_flight_scenes_;
    uint64_t in_flight_bytes_ = 0;
    int in_flight_batches_ = 0;
    bool stopping_ = false;
    ServeStats stats_;
    std::vector<std::thread> threads_;
};

}  // namespace shkz
