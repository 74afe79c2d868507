// serve.cpp -- continuous-batching request server over a device generator
// set (mtgp_server_* in mtgp_hip.h).  No counterpart in the reference,
// whose samples launch one kernel per call
// (cuda-sample/mtgp32-cuda.cu:171-240); each request's numbers are the ones
// the reference's genrand loop would produce at that stream position.
//
// Threads: callers submit / poll / release; one scheduler thread orders all
// batches and launches them; planner threads build the batches' work
// lists.  Each scheduling round the scheduler
//   1. retires finished batches (event query) and publishes completions,
//   2. admits pending requests in arrival order while their output fits
//      the arena (first fit, 256-byte aligned, so every output starts on a
//      fresh cache line),
//   3. forms batches of everything admitted, one per state-size class, in
//      arrival order (the classes taking turns for the free batch slots,
//      the one with the oldest waiting request first) -- fixing each
//      batch's state slots, stream and dependencies -- and hands them to
//      the planner threads, which build
//      them side by side (Plan::build, with its upload on the batch's
//      stream),
//   4. launches planned batches in formation order: a batch waits
//      (hipStreamWaitEvent) only for in-flight batches that advance one of
//      its generators, otherwise it overlaps whatever is already running.
// Generator states: a batch reads slot cur[g] and writes cur[g] ^ 1, and
// cur flips when the batch is formed; batches launch in formation order and
// the event wait orders two batches that share a generator, so two slots
// suffice.
#include <hip/hip_runtime_api.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mtgp_hip.h"

#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using mtgp::Error;
using mtgp::hip_check;

namespace {

using Clock = std::chrono::steady_clock;

// Numbers per batch before it is closed (one more request may overshoot).
constexpr uint64_t kBatchNumbers = (uint64_t)1 << 28;

// First-fit allocator over [0, cap) in 256-byte units.
class ArenaAlloc {
  public:
    explicit ArenaAlloc(size_t cap) {
        if (cap)
            free_[0] = cap;
    }
    static size_t rounded(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
    bool alloc(size_t bytes, size_t &off) {
        const size_t need = rounded(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need)
                continue;
            off = it->first;
            const size_t rest = it->second - need, at = it->first + need;
            free_.erase(it);
            if (rest)
                free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto it = free_.emplace(off, rounded(bytes)).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
};

struct Req {
    int64_t ticket = 0;
    int gen = 0, kind = 0;
    uint64_t count = 0, skip = 0;
    size_t off = 0, bytes = 0;
    double t_submit = 0, t_launch = 0, t_done = -1;
};

// Batch life: free -> formed (its requests, state slots, stream and
// dependencies fixed by the scheduler, in arrival order) -> planned (a
// planner thread built its work list) -> in flight (launched, in formation
// order) -> free (retired).
enum BatchState { kFree, kFormed, kPlanned, kInflight };

struct Batch {
    mtgp::Plan plan;
    hipEvent_t done = nullptr;
    int stream = -1;
    std::atomic<int> state{kFree};
    uint64_t seq = 0;
    std::vector<int64_t> tickets;
    std::vector<int> gens;
    std::vector<std::pair<int, uint64_t>> deps;  // (batch, its seq) advancing one of our generators
    std::vector<mtgp::GenRef> refs;              // state slots at formation
    std::vector<mtgp_request_t> rq;
    double plan_us = 0;
};

// 64-bit digest of a byte string (mixing 8 bytes at a time); for
// comparing outputs, not for security.
uint64_t digest64(const unsigned char *p, size_t n) {
    uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)n;
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        memcpy(&w, p + i, 8);
        h = (h ^ w) * 0xFF51AFD7ED558CCDull;
        h ^= h >> 29;
    }
    for (; i < n; i++) {
        h = (h ^ p[i]) * 0xC4CEB9FE1A85EC53ull;
        h ^= h >> 29;
    }
    return h;
}

}  // namespace

struct mtgp_server {
    mtgp_genset *set = nullptr;
    mtgp::Context *ctx = nullptr;
    mtgp_server_config_t cfg{};
    bool plain = false;
    char *d_arena = nullptr;
    std::unique_ptr<ArenaAlloc> arena;
    std::shared_ptr<mtgp::JumpCache> cache;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_load;  // in-flight batches per stream
    std::vector<std::unique_ptr<Batch>> pool;
    std::vector<int> last_batch;  // per generator: in-flight batch advancing it, or -1
    std::vector<int> gen_group;   // per generator: 3 * (bits == 64) + class

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::unordered_map<int64_t, Req> reqs;  // live requests (until released)
    std::deque<int64_t> pending;            // submitted, not admitted
    std::vector<int64_t> admitted;          // admitted, not launched (scheduler)
    std::deque<mtgp_completion_t> done_q;
    int64_t next_ticket = 0, n_completed = 0;
    mtgp_server_stats_t stats{};
    std::string failure;
    std::atomic<bool> failed{false};  // failure is set

    std::thread th;
    std::atomic<bool> stop{false};
    Clock::time_point t0 = Clock::now();
    uint64_t seq = 0;

    // Planner threads: batches are formed in order by the scheduler, built
    // (Plan::build) concurrently here, and launched by the scheduler in
    // formation order.
    std::vector<std::thread> planners;
    std::mutex pmu;
    std::condition_variable pcv;
    std::deque<int> jobs;      // formed batches waiting for a planner
    std::deque<int> formed;    // formed or planned, not launched, formation order
    bool planners_stop = false;

    double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }

    void start(int n_planners) {
        for (int k = 0; k < n_planners; k++)
            planners.emplace_back([this] { plan_loop(); });
        th = std::thread([this] { run(); });
    }

    void plan_loop() {
        try {
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            for (;;) {
                int bi;
                {
                    std::unique_lock<std::mutex> lk(pmu);
                    pcv.wait(lk, [&] { return planners_stop || !jobs.empty(); });
                    if (jobs.empty())
                        return;
                    bi = jobs.front();
                    jobs.pop_front();
                }
                Batch &B = *pool[(size_t)bi];
                build(B);
                B.state.store(kPlanned, std::memory_order_release);
            }
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            failure = e.what();
            failed = true;
            cv_done.notify_all();
        }
    }

    void build(Batch &B) {
        mtgp::PlanOptions o;
        o.cache = cache.get();
        o.discard_max = cfg.cell;
        o.stream = streams[(size_t)B.stream];
        o.sync = false;
        o.jump_mode = (cfg.flags & MTGP_SERVER_FUSED_JUMPS) ? 0 : 1;
        // Batches carry few jumps: split each down to one chunk per wave so
        // a jump takes about one chunk's fold instead of several.
        o.max_parts = 64;
        const auto tp = Clock::now();
        B.plan.build(*ctx, B.refs, B.rq, o);
        B.plan_us = std::chrono::duration<double, std::micro>(Clock::now() - tp).count();
    }

    void run() {
        try {
            hip_check(hipSetDevice(ctx->device), "hipSetDevice");
            loop();
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            failure = e.what();
            failed = true;
            cv_done.notify_all();
        }
    }

    int count_state(int st) const {
        int k = 0;
        for (const auto &b : pool)
            k += b->state.load(std::memory_order_acquire) == st;
        return k;
    }
    int inflight() const { return count_state(kInflight); }
    int free_batch() const {
        for (size_t k = 0; k < pool.size(); k++)
            if (pool[k]->state.load(std::memory_order_acquire) == kFree)
                return (int)k;
        return -1;
    }

    void retire(int bi) {
        Batch &B = *pool[(size_t)bi];
        const double t = now_us();
        for (int g : B.gens)
            if (last_batch[(size_t)g] == bi)
                last_batch[(size_t)g] = -1;
        stream_load[(size_t)B.stream]--;
        B.state.store(kFree, std::memory_order_release);
        std::lock_guard<std::mutex> lk(mu);
        for (int64_t tk : B.tickets) {
            Req &r = reqs.at(tk);
            r.t_done = t;
            mtgp_completion_t c{};
            c.ticket = tk;
            c.out = d_arena + r.off;
            c.bytes = r.bytes;
            c.submit_us = r.t_submit;
            c.launch_us = r.t_launch;
            c.done_us = t;
            done_q.push_back(c);
            n_completed++;
        }
        stats.completed = n_completed;
        cv_done.notify_all();
    }

    // Fixes a batch's requests, state slots, stream and dependencies (in
    // arrival order) and hands it to a planner thread (or, with no planner
    // threads, builds it here).
    void form(const std::vector<int64_t> &tickets) {
        const int bi = free_batch();
        Batch &B = *pool[(size_t)bi];
        B.rq.clear();
        B.gens.clear();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t tk : tickets) {
                const Req &r = reqs.at(tk);
                mtgp_request_t q{};
                q.gen = r.gen;
                q.kind = r.kind;
                q.count = r.count;
                q.skip = r.skip;
                q.out = d_arena + r.off;
                B.rq.push_back(q);
                B.gens.push_back(r.gen);
            }
        }
        std::sort(B.gens.begin(), B.gens.end());
        B.gens.erase(std::unique(B.gens.begin(), B.gens.end()), B.gens.end());
        // Dependencies: formed or in-flight batches that advance one of our
        // generators (they launch before this one).
        B.deps.clear();
        for (int g : B.gens) {
            const int d = last_batch[(size_t)g];
            if (d >= 0)
                B.deps.emplace_back(d, pool[(size_t)d]->seq);
        }
        std::sort(B.deps.begin(), B.deps.end());
        B.deps.erase(std::unique(B.deps.begin(), B.deps.end()), B.deps.end());
        int s = 0;
        if (!plain) {
            if (!B.deps.empty()) {
                int best = B.deps[0].first;
                for (const auto &d : B.deps)
                    if (pool[(size_t)d.first]->seq > pool[(size_t)best]->seq)
                        best = d.first;
                s = pool[(size_t)best]->stream;
            } else {
                for (int k = 1; k < (int)streams.size(); k++)
                    if (stream_load[(size_t)k] < stream_load[(size_t)s])
                        s = k;
            }
        }
        // State slots: this batch reads cur[g] and writes cur[g] ^ 1; later
        // batches see the flipped slots (and wait for this one).
        B.refs = set->refs();
        for (int g : B.gens)
            set->set_cur(g, set->cur[(size_t)g] ^ 1);
        for (int g : B.gens)
            last_batch[(size_t)g] = bi;
        B.stream = s;
        stream_load[(size_t)s]++;
        B.seq = ++seq;
        B.tickets = tickets;
        B.state.store(kFormed, std::memory_order_release);
        formed.push_back(bi);
        if (planners.empty()) {
            build(B);
            B.state.store(kPlanned, std::memory_order_release);
        } else {
            std::lock_guard<std::mutex> lk(pmu);
            jobs.push_back(bi);
            pcv.notify_one();
        }
    }

    // Launches planned batches in formation order while in-flight slots
    // last; returns whether one was launched.
    bool launch_ready() {
        bool any = false;
        while (!formed.empty() && inflight() < cfg.max_inflight) {
            Batch &B = *pool[(size_t)formed.front()];
            if (B.state.load(std::memory_order_acquire) != kPlanned)
                break;
            const hipStream_t st = streams[(size_t)B.stream];
            for (const auto &d : B.deps) {
                const Batch &D = *pool[(size_t)d.first];
                // (a dependency already retired, or its slot reused by a
                // newer batch, has finished)
                if (D.seq == d.second && D.state.load(std::memory_order_acquire) == kInflight && D.stream != B.stream)
                    hip_check(hipStreamWaitEvent(st, D.done, 0), "hipStreamWaitEvent");
            }
            B.plan.launch(st);
            hip_check(hipEventRecord(B.done, st), "hipEventRecord");
            B.state.store(kInflight, std::memory_order_release);
            formed.pop_front();
            any = true;
            const double t = now_us();
            std::lock_guard<std::mutex> lk(mu);
            for (int64_t tk : B.tickets)
                reqs.at(tk).t_launch = t;
            stats.batches++;
            stats.plan_us += B.plan_us;
            stats.launches += (int64_t)(B.plan.launches.size() + B.plan.jlaunches.size());
            stats.numbers += (double)B.plan.numbers;
            stats.max_batch_requests = std::max<int64_t>(stats.max_batch_requests, (int64_t)B.tickets.size());
        }
        return any;
    }

    void loop() {
        auto idle_since = Clock::now();
        for (;;) {
            if (failed)  // a planner thread failed: stop scheduling
                return;
            bool progress = false;
            for (size_t k = 0; k < pool.size(); k++) {
                if (pool[k]->state.load(std::memory_order_acquire) != kInflight)
                    continue;
                const hipError_t q = hipEventQuery(pool[k]->done);
                if (q == hipSuccess) {
                    retire((int)k);
                    progress = true;
                } else if (q != hipErrorNotReady) {
                    hip_check(q, "batch");
                }
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                while (!pending.empty()) {
                    Req &r = reqs.at(pending.front());
                    if (r.bytes && !arena->alloc(r.bytes, r.off))
                        break;
                    admitted.push_back(r.ticket);
                    pending.pop_front();
                    progress = true;
                }
            }
            // Batches formed ahead of launch: at most one per planner thread
            // (one when planning inline).
            const int max_formed = std::max<int>(1, (int)planners.size());
            auto can_form = [&] {
                return free_batch() >= 0 && (int)formed.size() < max_formed &&
                       inflight() + (int)formed.size() < cfg.max_inflight + max_formed;
            };
            if (launch_ready())
                progress = true;
            if (!admitted.empty() && can_form()) {
                if (plain) {
                    form({admitted.front()});
                    admitted.erase(admitted.begin());
                } else {
                    // Batches of one state-size class each, in arrival
                    // order, capped in requests and numbers, while slots
                    // last: one batch per class per pass, the class with
                    // the oldest waiting request first, so a busy class
                    // does not take every slot from the others.
                    std::map<int, std::vector<int64_t>> by_group;
                    {
                        std::lock_guard<std::mutex> lk(mu);
                        for (int64_t tk : admitted)
                            by_group[gen_group[(size_t)reqs.at(tk).gen]].push_back(tk);
                    }
                    std::vector<std::pair<int64_t, std::vector<int64_t> *>> order;
                    for (auto &kv : by_group)
                        order.emplace_back(kv.second.front(), &kv.second);
                    std::sort(order.begin(), order.end());
                    std::vector<size_t> next(order.size(), 0);
                    for (bool more = true; more && can_form();) {
                        more = false;
                        for (size_t o = 0; o < order.size() && can_form(); o++) {
                            const auto &v = *order[o].second;
                            size_t &i = next[o];
                            if (i >= v.size())
                                continue;
                            std::vector<int64_t> part;
                            uint64_t numbers = 0;
                            {
                                std::lock_guard<std::mutex> lk(mu);
                                while (i < v.size() && (int)part.size() < cfg.max_batch &&
                                       (part.empty() || numbers < kBatchNumbers)) {
                                    const Req &r = reqs.at(v[i]);
                                    numbers += r.count;
                                    part.push_back(v[i++]);
                                }
                            }
                            form(part);
                            more |= i < v.size();
                        }
                    }
                    std::vector<int64_t> left;
                    for (size_t o = 0; o < order.size(); o++)
                        left.insert(left.end(), order[o].second->begin() + (long)next[o], order[o].second->end());
                    std::sort(left.begin(), left.end());
                    admitted.swap(left);
                }
                progress = true;
            }
            if (launch_ready())
                progress = true;
            if (progress) {
                idle_since = Clock::now();
                continue;
            }
            const bool busy = inflight() > 0 || !formed.empty();
            std::unique_lock<std::mutex> lk(mu);
            if (stop && pending.empty() && admitted.empty() && !busy)
                return;
            if (!busy && pending.empty()) {
                cv_work.wait_for(lk, std::chrono::milliseconds(2));
            } else if (Clock::now() - idle_since < std::chrono::microseconds(300)) {
                lk.unlock();
                std::this_thread::yield();
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }

    void check_failure() {
        if (!failure.empty())
            throw Error("server scheduler failed: " + failure);
    }

    int64_t submit(int gen, int kind, uint64_t count, uint64_t skip) {
        if (gen < 0 || gen >= (int)set->bits.size())
            throw Error("request for generator " + std::to_string(gen) + " outside the set");
        if (kind < 0 || kind > 4)
            throw Error("unknown output kind " + std::to_string(kind));
        const size_t bytes = (size_t)count * (size_t)(set->bits[(size_t)gen] / 8);
        if (ArenaAlloc::rounded(bytes) > cfg.arena_bytes)
            throw Error("request larger than the server's output budget");
        std::lock_guard<std::mutex> lk(mu);
        check_failure();
        if (stop)
            throw Error("server is stopping");
        Req r;
        r.ticket = next_ticket++;
        r.gen = gen;
        r.kind = kind;
        r.count = count;
        r.skip = skip;
        r.bytes = bytes;
        r.t_submit = now_us();
        reqs.emplace(r.ticket, r);
        pending.push_back(r.ticket);
        stats.submitted = next_ticket;
        cv_work.notify_one();
        return r.ticket;
    }

    int poll(mtgp_completion_t *out, int max, int wait_ms) {
        std::unique_lock<std::mutex> lk(mu);
        if (done_q.empty() && wait_ms > 0)
            cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms),
                             [&] { return !done_q.empty() || !failure.empty(); });
        check_failure();
        int k = 0;
        while (k < max && !done_q.empty()) {
            out[k++] = done_q.front();
            done_q.pop_front();
        }
        return k;
    }

    void release(int64_t ticket) {
        std::lock_guard<std::mutex> lk(mu);
        auto it = reqs.find(ticket);
        if (it == reqs.end())
            throw Error("unknown ticket");
        if (it->second.t_done < 0)
            throw Error("request not completed yet");
        if (it->second.bytes)
            arena->release(it->second.off, it->second.bytes);
        reqs.erase(it);
        cv_work.notify_one();
    }

    void drain() {
        std::unique_lock<std::mutex> lk(mu);
        cv_done.wait(lk, [&] { return n_completed == next_ticket || !failure.empty(); });
        check_failure();
    }

    ~mtgp_server() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
            cv_work.notify_all();
        }
        if (th.joinable())
            th.join();
        {
            std::lock_guard<std::mutex> lk(pmu);
            planners_stop = true;
            pcv.notify_all();
        }
        for (auto &t : planners)
            if (t.joinable())
                t.join();
        for (auto &b : pool)
            if (b->done)
                (void)hipEventDestroy(b->done);
        pool.clear();
        for (auto s : streams)
            (void)hipStreamDestroy(s);
        if (d_arena)
            (void)hipFree(d_arena);
    }
};

#define API_BEGIN try {
#define API_END(fail)                   \
    }                                   \
    catch (const std::exception &e) {   \
        mtgp::set_error(e.what());      \
        return fail;                    \
    }

extern "C" {

mtgp_server_t *mtgp_server_create(mtgp_genset_t *set, const mtgp_server_config_t *config) {
    API_BEGIN
    if (!set)
        throw Error("NULL generator set");
    auto s = std::make_unique<mtgp_server>();
    s->set = set;
    s->ctx = set->ctx;
    if (config)
        s->cfg = *config;
    auto &c = s->cfg;
    if (!c.arena_bytes)
        c.arena_bytes = (size_t)4 << 30;
    if (!c.cell)
        c.cell = (uint64_t)1 << 17;
    if (c.cell & (c.cell - 1))
        throw Error("cell size must be a power of two");
    if (!c.cached_cells)
        c.cached_cells = 512;
    if (!c.streams)
        c.streams = 4;
    if (!c.max_inflight)
        c.max_inflight = 4;  // one per stream: more queue behind each other on the
                             // device's hardware queues and lengthen the tail
    if (!c.max_batch)
        c.max_batch = 256;
    if (c.streams < 1 || c.streams > 8 || c.max_inflight < 1 || c.max_inflight > 64 || c.max_batch < 1)
        throw Error("streams must be 1..8, max_inflight 1..64 and max_batch >= 1");
    s->plain = (c.flags & MTGP_SERVER_PLAIN) != 0;
    hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
    hip_check(hipMalloc((void **)&s->d_arena, c.arena_bytes), "hipMalloc server arena");
    s->arena = std::make_unique<ArenaAlloc>(c.arena_bytes);
    s->cache = mtgp::JumpCache::shared(*s->ctx, set->slot, c.cell, c.cached_cells);
    const int ns = s->plain ? 1 : c.streams;
    for (int k = 0; k < ns; k++) {
        hipStream_t st;
        hip_check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        s->streams.push_back(st);
        s->stream_load.push_back(0);
    }
    // Planner threads (MTGP_SERVER_PLANNERS, default 2; 0 = plan on the
    // scheduler thread).  With the classes taking turns for batch slots,
    // two planner threads served the whole offered 368 GB/s of the serving
    // mix at 240,000 req/s (p99 6.6-9 ms) against 311-315 GB/s inline, and
    // 396-415 against 325-335 GB/s at 320,000 req/s, with the same median
    // at 20,000 req/s (0.32 ms).  The plain baseline always plans inline.
    int n_planners = 2;
    if (const char *e = getenv("MTGP_SERVER_PLANNERS"))
        n_planners = std::max(0, std::min(16, atoi(e)));
    if (s->plain)
        n_planners = 0;
    for (int k = 0; k < c.max_inflight + std::max(1, n_planners); k++) {
        auto b = std::make_unique<Batch>();
        hip_check(hipEventCreateWithFlags(&b->done, hipEventDisableTiming), "hipEventCreate");
        b->plan.ctx = s->ctx;
        b->plan.reserve((size_t)8 << 20, (size_t)64 << 20);
        s->pool.push_back(std::move(b));
    }
    s->last_batch.assign(set->bits.size(), -1);
    for (size_t g = 0; g < set->bits.size(); g++)
        s->gen_group.push_back((set->bits[g] == 64 ? 3 : 0) + s->ctx->class_of(set->slot[g]));
    // Finish whatever the set's own calls left in flight before the
    // scheduler starts reading its states.
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    s->start(n_planners);
    return s.release();
    API_END(nullptr)
}

int mtgp_server_destroy(mtgp_server_t *s) {
    API_BEGIN
    if (!s)
        return 0;
    int rc = 0;
    try {
        s->drain();
    } catch (const std::exception &e) {
        mtgp::set_error(e.what());
        rc = -1;
    }
    delete s;
    return rc;
    API_END(-1)
}

int64_t mtgp_server_submit(mtgp_server_t *s, int gen, int kind, uint64_t count, uint64_t skip) {
    API_BEGIN
    return s->submit(gen, kind, count, skip);
    API_END(-1)
}

int mtgp_server_poll(mtgp_server_t *s, mtgp_completion_t *out, int max, int wait_ms) {
    API_BEGIN
    return s->poll(out, max, wait_ms);
    API_END(-1)
}

int mtgp_server_release(mtgp_server_t *s, int64_t ticket) {
    API_BEGIN
    s->release(ticket);
    return 0;
    API_END(-1)
}

int mtgp_server_drain(mtgp_server_t *s) {
    API_BEGIN
    s->drain();
    return 0;
    API_END(-1)
}

int mtgp_server_stats(mtgp_server_t *s, mtgp_server_stats_t *out) {
    API_BEGIN
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return 0;
    API_END(-1)
}

int mtgp_server_replay(mtgp_server_t *s, int n, const double *t_us, const int32_t *gen, const int32_t *kind,
                       const uint64_t *count, const uint64_t *skip, double *latency_us, double *queue_us,
                       uint64_t *digests, double *wall_us) {
    API_BEGIN
    if (n <= 0)
        return 0;
    std::atomic<int64_t> base{-1};
    std::atomic<bool> failed{false};
    std::string consumer_error;
    double last_done = 0;
    const double start = s->now_us();
    std::thread consumer([&] {
        try {
            hip_check(hipSetDevice(s->ctx->device), "hipSetDevice");
            hipStream_t cs = nullptr;
            unsigned char *host = nullptr;
            size_t host_cap = 0;
            if (digests)
                hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags");
            std::vector<mtgp_completion_t> got(256);
            int seen = 0;
            while (seen < n && !failed) {
                const int k = s->poll(got.data(), (int)got.size(), 5);
                for (int j = 0; j < k; j++) {
                    const mtgp_completion_t &c = got[(size_t)j];
                    while (base.load() < 0)
                        std::this_thread::yield();
                    const int64_t i = c.ticket - base.load();
                    if (i < 0 || i >= n)
                        throw Error("completion of a request outside the replay");
                    latency_us[i] = c.done_us - c.submit_us;
                    if (queue_us)
                        queue_us[i] = c.launch_us - c.submit_us;
                    last_done = std::max(last_done, c.done_us);
                    if (digests) {
                        if (c.bytes > host_cap) {
                            if (host)
                                hip_check(hipHostFree(host), "hipHostFree");
                            hip_check(hipHostMalloc((void **)&host, c.bytes, hipHostMallocDefault), "hipHostMalloc");
                            host_cap = c.bytes;
                        }
                        if (c.bytes) {
                            hip_check(hipMemcpyAsync(host, c.out, c.bytes, hipMemcpyDeviceToHost, cs), "copy out");
                            hip_check(hipStreamSynchronize(cs), "copy out sync");
                        }
                        digests[i] = digest64(host, c.bytes);
                    }
                    s->release(c.ticket);
                    seen++;
                }
            }
            if (host)
                (void)hipHostFree(host);
            if (cs)
                (void)hipStreamDestroy(cs);
        } catch (const std::exception &e) {
            consumer_error = e.what();
            failed = true;
        }
    });
    try {
        for (int i = 0; i < n && !failed; i++) {
            const double target = start + t_us[i];
            while (s->now_us() < target)
                std::this_thread::yield();
            const int64_t tk = s->submit(gen[i], kind[i], count[i], skip[i]);
            if (i == 0)
                base = tk;
            else if (tk != base.load() + i)
                throw Error("another thread submitted during the replay");
        }
    } catch (...) {
        failed = true;
        if (base.load() < 0)
            base = 0;
        consumer.join();
        throw;
    }
    consumer.join();
    if (failed)
        throw Error("replay: " + consumer_error);
    if (wall_us)
        *wall_us = last_done - start;
    return 0;
    API_END(-1)
}

uint64_t mtgp_digest64(const void *data, size_t bytes) {
    return digest64(static_cast<const unsigned char *>(data), bytes);
}

}  // extern "C"
