// serve_core_check.cpp -- CPU check of the projection server's scheduler (shkz_serve.hpp) against a
// stand-in executor that only sleeps and records what it ran.  Built with and without
// -fsanitize=thread by tests/test_serve_core.py.  Exit status 0 = every invariant held.
//
// Invariants checked:
//   * every submitted request completes exactly once (shutdown drains the queue);
//   * a scene never has two requests running at once, and its requests run in submission order;
//   * a batch holds one parameter group, at most max_batch_grids grids and at most
//     max_batch_cells cells (unless it is a single request);
//   * the estimated bytes of the batches in flight plus the resident bytes never exceed the budget;
//   * a request that cannot fit the budget alone fails with -4, others never do;
//   * a large request queued behind a stream of small ones is not starved;
//   * a request cancelled while queued never runs and reports -7; a started one is not cancelled.
#include "../shiokaze_amd/csrc/shkz_serve.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>

using namespace shkz;

static std::atomic<int> g_fail{0};
#define CHECK(c, ...)                                                                                                  \
    do {                                                                                                               \
        if (!(c)) {                                                                                                    \
            std::fprintf(stderr, "CHECK failed at line %d: %s: ", __LINE__, #c);                                       \
            std::fprintf(stderr, __VA_ARGS__);                                                                         \
            std::fprintf(stderr, "\n");                                                                                \
            g_fail = 1;                                                                                                \
        }                                                                                                              \
    } while (0)

struct Tag {
    uint64_t scene;
    int seq;  // position of the request in its scene's submission order
};

struct Recorder {
    std::mutex mu;
    std::set<uint64_t> running_scenes;
    std::map<uint64_t, int> next_seq;  // per scene: the sequence number expected next
    uint64_t in_flight_bytes = 0;
    uint64_t budget = 0;
    int max_grids = 0;
    int64_t max_cells = 0;
    uint64_t resident = 0;
    int batches = 0, max_batch = 0, max_concurrent = 0, concurrent = 0;
    bool allow_gaps = false;
};

static void run_fake(Recorder &R, int, std::vector<ServeRequest> &batch, std::vector<int> &status, int us_base) {
    uint64_t bytes = 0;
    int64_t cells = 0;
    {
        std::lock_guard<std::mutex> g(R.mu);
        for (const ServeRequest &r : batch) {
            bytes += r.est_bytes;
            cells += r.cells;
            CHECK(r.group == batch[0].group, "mixed groups");
            if (r.scene_id) {
                CHECK(!R.running_scenes.count(r.scene_id), "scene %llu runs twice", (unsigned long long)r.scene_id);
                R.running_scenes.insert(r.scene_id);
                const Tag *t = (const Tag *)r.payload;
                // in order; with cancellations a scene may skip the steps cancelled before they ran
                CHECK(R.allow_gaps ? t->seq >= R.next_seq[r.scene_id] : t->seq == R.next_seq[r.scene_id],
                      "scene %llu ran step %d, expected %d", (unsigned long long)r.scene_id, t->seq,
                      R.next_seq[r.scene_id]);
                R.next_seq[r.scene_id] = t->seq + 1;
            }
        }
        CHECK((int)batch.size() <= R.max_grids, "batch of %zu", batch.size());
        CHECK(batch.size() == 1 || cells <= R.max_cells, "batch cells %lld", (long long)cells);
        R.in_flight_bytes += bytes;
        CHECK(R.in_flight_bytes + R.resident <= R.budget, "in flight %llu over budget",
              (unsigned long long)R.in_flight_bytes);
        ++R.batches;
        ++R.concurrent;
        if (R.concurrent > R.max_concurrent) R.max_concurrent = R.concurrent;
        if ((int)batch.size() > R.max_batch) R.max_batch = (int)batch.size();
    }
    std::this_thread::sleep_for(std::chrono::microseconds(us_base + 20 * (int)batch.size()));
    {
        std::lock_guard<std::mutex> g(R.mu);
        for (const ServeRequest &r : batch)
            if (r.scene_id) R.running_scenes.erase(r.scene_id);
        R.in_flight_bytes -= bytes;
        --R.concurrent;
    }
    for (int &s : status) s = 0;
}

// Many clients, each stepping its own scenes (submit, wait, next step), plus one client that
// queues several steps of one scene at once.
static void scenario_clients() {
    Recorder R;
    ServeLimits lim;
    lim.executors = 3;
    lim.max_batch_grids = 8;
    lim.max_batch_cells = 1000;
    lim.budget = 10000;
    R.budget = lim.budget;
    R.max_grids = lim.max_batch_grids;
    R.max_cells = lim.max_batch_cells;
    R.resident = 500;
    std::atomic<uint64_t> next_ticket{1};
    {
        ServeCore core(
            lim, [&](int k, std::vector<ServeRequest> &b, std::vector<int> &st) { run_fake(R, k, b, st, 200); },
            [&] { return R.resident; });
        const int clients = 8, scenes_per_client = 3, steps = 15;
        std::vector<std::thread> th;
        std::atomic<int> done{0};
        for (int c = 0; c < clients; ++c)
            th.emplace_back([&, c] {
                std::mt19937 rng(c);
                std::vector<Tag> tags(scenes_per_client * steps);
                for (int s = 0; s < steps; ++s)
                    for (int q = 0; q < scenes_per_client; ++q) {
                        Tag &t = tags[s * scenes_per_client + q];
                        t.scene = 1000 * (c + 1) + q;
                        t.seq = s;
                        ServeRequest r;
                        r.ticket = next_ticket++;
                        r.scene_id = t.scene;
                        r.group = (c % 2);
                        r.est_bytes = 100 + rng() % 1500;
                        r.cells = 50 + rng() % 400;
                        r.payload = &t;
                        CHECK(core.submit(r) == 0, "submit");
                        ServeOutcome o;
                        CHECK(core.wait(r.ticket, -1, &o) == 0, "wait");
                        CHECK(o.status == 0, "status %d", o.status);
                        CHECK(o.payload == &t, "payload");
                        CHECK(o.t_end_ms >= o.t_start_ms && o.t_start_ms >= o.t_submit_ms, "times");
                        ++done;
                    }
            });
        // a client that queues all steps of one scene at once, then polls them
        th.emplace_back([&] {
            const int n = 40;
            std::vector<Tag> tags(n);
            std::vector<uint64_t> tickets;
            for (int s = 0; s < n; ++s) {
                tags[s] = Tag{777, s};
                ServeRequest r;
                r.ticket = next_ticket++;
                r.scene_id = 777;
                r.est_bytes = 300;
                r.cells = 100;
                r.payload = &tags[s];
                CHECK(core.submit(r) == 0, "submit");
                tickets.push_back(r.ticket);
            }
            for (uint64_t t : tickets) {
                ServeOutcome o;
                int rc;
                while ((rc = core.poll(t, &o)) == 1) std::this_thread::sleep_for(std::chrono::microseconds(50));
                CHECK(rc == 0 && o.status == 0, "burst rc %d", rc);
                ++done;
            }
        });
        for (auto &t : th) t.join();
        const ServeStats st = core.stats();
        CHECK(done == clients * scenes_per_client * steps + 40, "done %d", done.load());
        CHECK(st.completed == st.submitted, "completed %llu of %llu", (unsigned long long)st.completed,
              (unsigned long long)st.submitted);
        CHECK(st.rejected == 0, "rejected %llu", (unsigned long long)st.rejected);
        CHECK(core.poll(123456789, nullptr) == -5, "unknown ticket");
        std::printf("clients: %llu requests in %llu batches (max %d, %d executors busy at most)\n",
                    (unsigned long long)st.submitted, (unsigned long long)st.batches, st.max_batch,
                    st.max_in_flight_batches);
        CHECK(st.max_in_flight_batches > 1, "no overlap of executors");
        CHECK(st.max_batch > 1, "no batching");
    }
}

// Budget: a request larger than the budget fails with -4; one that fits only alone runs alone
// and is not starved by a stream of small requests submitted after it; shutdown drains.
static void scenario_budget() {
    Recorder R;
    ServeLimits lim;
    lim.executors = 2;
    lim.max_batch_grids = 16;
    lim.max_batch_cells = 1 << 20;
    lim.budget = 1000;
    R.budget = lim.budget;
    R.max_grids = lim.max_batch_grids;
    R.max_cells = lim.max_batch_cells;
    std::vector<Tag> tags(400);
    std::vector<uint64_t> small;
    uint64_t big = 0, huge = 0;
    {
        ServeCore core(
            lim, [&](int k, std::vector<ServeRequest> &b, std::vector<int> &st) { run_fake(R, k, b, st, 300); },
            [] { return (uint64_t)0; });
        uint64_t t = 1;
        auto make = [&](uint64_t bytes) {
            ServeRequest r;
            r.ticket = t++;
            r.est_bytes = bytes;
            r.cells = 10;
            r.payload = &tags[r.ticket % tags.size()];
            return r;
        };
        for (int i = 0; i < 50; ++i) {
            ServeRequest r = make(150);
            small.push_back(r.ticket);
            core.submit(r);
        }
        ServeRequest b = make(900);  // fits only when nothing else is in flight
        big = b.ticket;
        core.submit(b);
        ServeRequest h = make(5000);  // never fits
        huge = h.ticket;
        core.submit(h);
        for (int i = 0; i < 150; ++i) {
            ServeRequest r = make(150);
            small.push_back(r.ticket);
            core.submit(r);
        }
        ServeOutcome o;
        CHECK(core.wait(huge, -1, &o) == 0 && o.status == -4, "huge status %d", o.status);
        CHECK(core.wait(big, -1, &o) == 0 && o.status == 0, "big status %d", o.status);
        CHECK(o.batch_size == 1, "big ran in a batch of %d", o.batch_size);
        const double big_end = o.t_end_ms;
        int after = 0;
        for (uint64_t s : small) {
            CHECK(core.wait(s, -1, &o) == 0 && o.status == 0, "small status %d", o.status);
            if (o.t_start_ms >= big_end) ++after;
        }
        // the small requests behind the big one waited for it (memory granted in arrival order)
        CHECK(after >= 100, "only %d small requests ran after the big one", after);
        const ServeStats st = core.stats();
        CHECK(st.rejected == 1, "rejected %llu", (unsigned long long)st.rejected);
        std::printf("budget: %llu batches, max in flight %llu bytes of %llu\n", (unsigned long long)st.batches,
                    (unsigned long long)st.max_in_flight_bytes, (unsigned long long)lim.budget);
        CHECK(st.max_in_flight_bytes <= lim.budget, "in-flight bytes");
        // requests left uncollected at shutdown are drained
        for (int i = 0; i < 20; ++i) core.submit(make(100));
        core.shutdown();
        CHECK(core.take_all().size() == 20, "drain");
        CHECK(core.submit(make(1)) == -1, "submit after shutdown");
    }
}

// Cancellation: requests cancelled while queued complete with -7 and never run; cancelling a
// running or finished request changes nothing (1); an unknown ticket gives -5; the remaining
// requests of a scene still run in order after a cancelled one; clients cancel from several
// threads while executors run.
static void scenario_cancel() {
    Recorder R;
    ServeLimits lim;
    lim.executors = 2;
    lim.max_batch_grids = 4;
    lim.max_batch_cells = 1 << 20;
    lim.budget = 1 << 30;
    R.budget = lim.budget;
    R.max_grids = lim.max_batch_grids;
    R.max_cells = lim.max_batch_cells;
    R.allow_gaps = true;
    const int scenes = 6, steps = 30;
    std::vector<Tag> tags(scenes * steps);
    std::atomic<int> ran{0};
    {
        ServeCore core(
            lim,
            [&](int k, std::vector<ServeRequest> &b, std::vector<int> &st) {
                ran += (int)b.size();
                run_fake(R, k, b, st, 150);
            },
            [] { return (uint64_t)0; });
        std::vector<uint64_t> tickets(scenes * steps);
        uint64_t t = 1;
        for (int s = 0; s < steps; ++s)
            for (int q = 0; q < scenes; ++q) {
                Tag &tg = tags[q * steps + s];
                tg = Tag{(uint64_t)(500 + q), s};
                ServeRequest r;
                r.ticket = t++;
                r.scene_id = tg.scene;
                r.est_bytes = 10;
                r.cells = 10;
                r.payload = &tg;
                tickets[q * steps + s] = r.ticket;
                CHECK(core.submit(r) == 0, "submit");
            }
        // three threads cancel every third step of their scenes from the back (run_fake checks that
        // each scene's remaining steps run in increasing order)
        std::atomic<int> cancelled{0}, not_cancelled{0};
        std::vector<std::thread> th;
        for (int c = 0; c < 3; ++c)
            th.emplace_back([&, c] {
                for (int q = c; q < scenes; q += 3)
                    for (int s = steps - 1; s >= 0; s -= 3) {
                        const int i = q * steps + s;
                        const int rc = core.cancel(tickets[i]);
                        CHECK(rc == 0 || rc == 1, "cancel rc %d", rc);
                        if (rc == 0) ++cancelled;
                        else ++not_cancelled;
                    }
            });
        for (auto &x : th) x.join();
        // the first step of the first scene has long run: cancelling it changes nothing
        std::this_thread::sleep_for(std::chrono::milliseconds(100));
        CHECK(core.cancel(tickets[0]) == 1, "cancelled a finished request");
        // collect every request: cancelled ones report -7 and were never run
        int got_cancelled = 0, got_ok = 0;
        for (int i = 0; i < scenes * steps; ++i) {
            ServeOutcome o;
            CHECK(core.wait(tickets[i], -1, &o) == 0, "wait");
            if (o.status == -7) {
                ++got_cancelled;
                CHECK(o.executor == -1 && o.batch_size == 0, "cancelled request has an executor");
            } else {
                CHECK(o.status == 0, "status %d", o.status);
                ++got_ok;
            }
            CHECK(core.cancel(tickets[i]) == -5, "collected ticket still known");
        }
        CHECK(got_cancelled == cancelled.load(), "cancelled %d reported %d", cancelled.load(), got_cancelled);
        CHECK(got_ok + got_cancelled == scenes * steps, "outcomes");
        CHECK(ran.load() == got_ok, "ran %d ok %d", ran.load(), got_ok);
        CHECK(cancelled.load() > 0, "nothing was cancelled");
        const ServeStats st = core.stats();
        CHECK(st.cancelled == (uint64_t)cancelled.load(), "stats.cancelled %llu", (unsigned long long)st.cancelled);
        CHECK(core.cancel(987654321) == -5, "unknown ticket");
        std::printf("cancel: %d cancelled while queued, %d already started or done\n", cancelled.load(),
                    not_cancelled.load());
    }
}

int main() {
    scenario_clients();
    scenario_budget();
    scenario_cancel();
    if (g_fail) {
        std::printf("FAILED\n");
        return 1;
    }
    std::printf("OK\n");
    return 0;
}
