// sf_server.cpp -- request loop over the batch (include/sonic_amd_serve.h).
//
// One loop thread owns the batch.  Each pass it (1) takes the requests
// submitted since the last pass, (2) admits waiting requests in arrival
// order while the arena and the problem slots have room -- a request that
// does not fit is passed over by later, smaller ones for at most
// kMaxBypass passes, after which admission waits for it -- (3) advances the
// batch by the smallest remaining step count of the resident problems
// (capped at steps_per_tick), so a problem finishes exactly at a pass
// boundary, and (4) hands finished problems back and frees their slots.
// Each problem stops at its own step count on the device (sf_batch_add's
// max_steps), so it never takes a step too many.  A finished problem's
// state is copied out on the batch's copy stream into pinned host memory
// while the next pass computes; its slot is freed and its result handed
// back once that copy has completed.
#include "sonic_amd_serve.h"
#include "sf_batch_copy.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace sfamd {
void set_last_error(const std::string& message);
}

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxBypass = 8;

struct Request {
    long long ticket = 0;
    sf_batch_item item{};
    long long steps = 0;
    long long nodes = 0;          // unique nodes
    std::vector<double> state;    // initial state, then the result
    Clock::time_point submit, done;
    int status = 0;               // 0 pending, 1 done, -1 failed
    int slot = -1;
    long long steps_done = 0;
    double time = 0.0;
    int bypassed = 0;
    std::string error;
};

using RequestPtr = std::shared_ptr<Request>;

int fail(int code, const std::string& msg) {
    sfamd::set_last_error(msg);
    return code;
}

std::string last_error() {
    char buf[512] = {0};
    sf_last_error(buf, sizeof buf);
    return buf;
}

}  // namespace

struct sf_server {
    sf_batch* batch = nullptr;
    int max_items = 0;
    int steps_per_tick = 1;
    size_t arena_nodes = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    // guarded by mu
    std::deque<RequestPtr> incoming;
    std::unordered_map<long long, RequestPtr> table;
    long long next_ticket = 1;
    bool stop = false;
    long long submitted = 0, completed = 0, failed = 0, outstanding = 0, resident = 0;
    std::vector<double> latencies_ms;
    double cell_updates = 0.0, busy_s = 0.0, admit_s = 0.0, retire_s = 0.0;
    long long ticks = 0;
    int peak_items = 0;
    size_t peak_nodes = 0;
    bool have_first = false;
    Clock::time_point first_submit, last_done;

    bool shape_pending = false;
    int shape[4] = {255, 32, 16, 0};   // xlen, ylen, blocks per CU, fused
    int launch = 2;                      // sf_batch_step's use_graph

    // loop thread only
    std::deque<RequestPtr> waiting;
    std::vector<RequestPtr> running;
    struct Pinned {
        double* p = nullptr;
        long long cap = 0;      // doubles
        bool busy = false;
    };
    struct Outgoing {
        RequestPtr r;
        int buffer = -1;        // index into pinned
        hipEvent_t done = nullptr;
    };
    std::vector<Pinned> pinned;
    std::vector<hipEvent_t> events;  // free copy-out events
    std::vector<Outgoing> outgoing;  // copy-outs in flight
    std::thread worker;
};

namespace {

void finish(sf_server* s, const RequestPtr& r, int status, const std::string& error) {
    std::lock_guard<std::mutex> lk(s->mu);
    r->status = status;
    r->error = error;
    r->done = Clock::now();
    --s->outstanding;
    if (r->slot >= 0) --s->resident;
    if (status == 1) {
        ++s->completed;
        s->latencies_ms.push_back(
            std::chrono::duration<double, std::milli>(r->done - r->submit).count());
        s->cell_updates += (double)r->nodes * (double)r->steps;
        s->last_done = r->done;
    } else {
        ++s->failed;
    }
    s->cv_done.notify_all();
}

void admit(sf_server* s) {
    for (auto it = s->waiting.begin(); it != s->waiting.end();) {
        const RequestPtr& r = *it;
        int slot = -1;
        const int rc = sf_batch_add(s->batch, &r->item, r->state.data(), r->steps, &slot);
        if (rc == 0) {
            r->slot = slot;
            std::vector<double>().swap(r->state);  // the device holds it now
            {
                std::lock_guard<std::mutex> lk(s->mu);
                ++s->resident;
            }
            s->running.push_back(r);
            it = s->waiting.erase(it);
            continue;
        }
        if (rc == SF_E_MEMORY) {
            if (it == s->waiting.begin() && ++r->bypassed > kMaxBypass) break;
            ++it;
            continue;
        }
        const RequestPtr dead = r;
        it = s->waiting.erase(it);
        finish(s, dead, -1, "admission failed: " + last_error());
    }
    size_t used = 0, free_nodes = 0;
    int items = 0;
    if (sf_batch_memory(s->batch, &used, &free_nodes, &items) == 0) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->peak_items = std::max(s->peak_items, items);
        s->peak_nodes = std::max(s->peak_nodes, used);
    }
}

// A pinned host buffer of at least n doubles (reused; grown on demand).
int pinned_buffer(sf_server* s, long long n) {
    int best = -1;
    for (int i = 0; i < (int)s->pinned.size(); ++i)
        if (!s->pinned[i].busy && s->pinned[i].cap >= n &&
            (best < 0 || s->pinned[i].cap < s->pinned[best].cap))
            best = i;
    if (best >= 0) return best;
    sf_server::Pinned b;
    if (hipHostMalloc((void**)&b.p, sizeof(double) * (size_t)n, hipHostMallocDefault) != hipSuccess)
        return -1;
    b.cap = n;
    s->pinned.push_back(b);
    return (int)s->pinned.size() - 1;
}

// Start the copy-out of a finished problem; on failure the request fails.
void start_copy_out(sf_server* s, const RequestPtr& r) {
    hipEvent_t ev = nullptr;
    if (!s->events.empty()) {
        ev = s->events.back();
        s->events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        sf_batch_remove(s->batch, r->slot);
        finish(s, r, -1, "copy-out: event creation failed");
        return;
    }
    const int buf = pinned_buffer(s, sfamd::batch_copy_doubles(s->batch, r->slot));
    if (buf < 0 || sfamd::batch_copy_out_async(s->batch, r->slot, s->pinned[buf].p, ev) != 0) {
        const std::string why = buf < 0 ? "pinned host allocation failed" : last_error();
        s->events.push_back(ev);
        sf_batch_remove(s->batch, r->slot);
        finish(s, r, -1, "copy-out: " + why);
        return;
    }
    s->pinned[buf].busy = true;
    s->outgoing.push_back({r, buf, ev});
}

// Hand back the problems whose copy-out has completed (all of them when
// `wait`), freeing their slots.
void collect(sf_server* s, bool wait) {
    for (auto it = s->outgoing.begin(); it != s->outgoing.end();) {
        const hipError_t e = wait ? hipEventSynchronize(it->done) : hipEventQuery(it->done);
        if (e == hipErrorNotReady) {
            ++it;
            continue;
        }
        const RequestPtr r = it->r;
        sf_server::Pinned& b = s->pinned[(size_t)it->buffer];
        std::string err;
        if (e == hipSuccess) {
            r->state.resize((size_t)5 * (size_t)r->nodes);
            sfamd::batch_unpack(s->batch, r->slot, b.p, r->state.data(), &r->time);
        } else {
            err = std::string("copy-out: ") + hipGetErrorString(e);
        }
        b.busy = false;
        s->events.push_back(it->done);
        sf_batch_remove(s->batch, r->slot);
        it = s->outgoing.erase(it);
        finish(s, r, err.empty() ? 1 : -1, err);
    }
}

void drop_running(sf_server* s, const std::string& why) {
    for (const RequestPtr& r : s->running) {
        sf_batch_remove(s->batch, r->slot);
        finish(s, r, -1, why);
    }
    s->running.clear();
}

void loop(sf_server* s) {
    std::vector<long long> steps((size_t)s->max_items);
    for (;;) {
        collect(s, false);
        if (s->waiting.empty() && s->running.empty() && !s->outgoing.empty()) {
            bool idle;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                idle = s->incoming.empty() && !s->stop;
            }
            if (idle) collect(s, true);
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->waiting.empty() && s->running.empty() && s->outgoing.empty())
                s->cv_work.wait(lk, [&] { return s->stop || !s->incoming.empty(); });
            if (s->stop) {
                lk.unlock();
                collect(s, true);
                return;
            }
            while (!s->incoming.empty()) {
                s->waiting.push_back(std::move(s->incoming.front()));
                s->incoming.pop_front();
            }
            if (s->shape_pending) {
                sf_batch_set_tiles(s->batch, s->shape[0], s->shape[1]);
                sf_batch_set_occupancy(s->batch, s->shape[2]);
                sf_batch_set_schedule(s->batch, s->shape[3]);
                s->shape_pending = false;
            }
        }
        const auto ta = Clock::now();
        admit(s);
        {
            const double d = std::chrono::duration<double>(Clock::now() - ta).count();
            std::lock_guard<std::mutex> lk(s->mu);
            s->admit_s += d;
        }
        if (s->running.empty()) {
            // nothing runs: either copy-outs still hold the memory the
            // waiting requests need, or nothing fits at all (only possible
            // after failures); wait for the copies or fail the waiting
            if (!s->outgoing.empty()) {
                collect(s, true);
                continue;
            }
            if (!s->waiting.empty()) {
                for (const RequestPtr& r : s->waiting) finish(s, r, -1, "cannot be admitted");
                s->waiting.clear();
            }
            continue;
        }
        long long k = s->steps_per_tick;
        for (const RequestPtr& r : s->running) k = std::min(k, r->steps - r->steps_done);
        k = std::max(1LL, k);
        const auto t0 = Clock::now();
        int rc = sf_batch_step(s->batch, (int)k, s->launch);
        const double dt = std::chrono::duration<double>(Clock::now() - t0).count();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->busy_s += dt;
            ++s->ticks;
        }
        if (rc == SF_E_NUMERIC) {
            // only the problems that raised flags are failed; they are
            // frozen on the device and the others advanced normally
            const std::string why = last_error();
            std::vector<uint32_t> flags((size_t)s->max_items);
            if (sf_batch_faults(s->batch, flags.data()) != 0) {
                drop_running(s, "batch step failed: " + why);
                continue;
            }
            for (auto it = s->running.begin(); it != s->running.end();) {
                const RequestPtr r = *it;
                if (!flags[(size_t)r->slot]) {
                    ++it;
                    continue;
                }
                char msg[96];
                std::snprintf(msg, sizeof msg, "numerical fault flags 0x%x", flags[(size_t)r->slot]);
                sf_batch_remove(s->batch, r->slot);
                it = s->running.erase(it);
                finish(s, r, -1, msg);
            }
        } else if (rc != 0) {
            drop_running(s, "batch step failed: " + last_error());
            continue;
        }
        if (sf_batch_steps(s->batch, steps.data()) != 0) {
            drop_running(s, "batch clock failed: " + last_error());
            continue;
        }
        const auto tr = Clock::now();
        for (auto it = s->running.begin(); it != s->running.end();) {
            const RequestPtr r = *it;
            r->steps_done = steps[(size_t)r->slot];
            if (r->steps_done < r->steps) {
                ++it;
                continue;
            }
            it = s->running.erase(it);
            start_copy_out(s, r);
        }
        const double d = std::chrono::duration<double>(Clock::now() - tr).count();
        std::lock_guard<std::mutex> lk(s->mu);
        s->retire_s += d;
    }
}

RequestPtr find(sf_server* s, long long ticket) {
    auto it = s->table.find(ticket);
    return it == s->table.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

int sf_server_create(int device, double gamma, size_t device_bytes, int max_items,
                     int steps_per_tick, sf_server** out) {
    if (!out) return fail(SF_E_ARG, "sf_server_create: null output");
    *out = nullptr;
    if (max_items < 1 || steps_per_tick < 1)
        return fail(SF_E_ARG, "sf_server_create: max_items and steps_per_tick must be positive");
    const size_t nodes = device_bytes / (4 * 5 * sizeof(double));
    if (nodes < 64) return fail(SF_E_ARG, "sf_server_create: device budget too small");
    auto* s = new (std::nothrow) sf_server;
    if (!s) return fail(SF_E_MEMORY, "host allocation failed");
    if (int rc = sf_batch_create(device, gamma, nodes, max_items, &s->batch)) {
        delete s;
        return rc;
    }
    // the launch shape the autotuner picks for one-variant streams of
    // mid-size problems: one launch per variant, long x tiles, a deep
    // persistent grid
    sf_batch_set_schedule(s->batch, 0);
    sf_batch_set_tiles(s->batch, 255, 32);
    sf_batch_set_occupancy(s->batch, 16);
    s->max_items = max_items;
    s->steps_per_tick = steps_per_tick;
    s->arena_nodes = nodes;
    try {
        s->worker = std::thread(loop, s);
    } catch (...) {
        sf_batch_destroy(s->batch);
        delete s;
        return fail(SF_E_STATE, "sf_server_create: cannot start the loop thread");
    }
    *out = s;
    return 0;
}

int sf_server_destroy(sf_server* s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    sf_batch_destroy(s->batch);
    for (auto& b : s->pinned)
        if (b.p) (void)hipHostFree(b.p);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    delete s;
    return 0;
}

int sf_server_submit(sf_server* s, const sf_batch_item* item, const double* soa, long long steps,
                     long long* ticket) {
    if (!s || !item || !soa || !ticket) return fail(SF_E_ARG, "sf_server_submit: null argument");
    if (steps < 1) return fail(SF_E_ARG, "sf_server_submit: steps must be positive");
    size_t padded = 0;
    if (int rc = sf_batch_item_nodes(item, &padded)) return rc;
    if (padded > s->arena_nodes)
        return fail(SF_E_MEMORY, "sf_server_submit: problem larger than the whole arena");
    auto r = std::make_shared<Request>();
    r->item = *item;
    r->steps = steps;
    r->nodes = (long long)item->n[0] * std::max(1, item->n[1]);
    try {
        r->state.assign(soa, soa + (size_t)5 * (size_t)r->nodes);
    } catch (...) {
        return fail(SF_E_MEMORY, "sf_server_submit: host allocation failed");
    }
    r->submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        r->ticket = s->next_ticket++;
        s->table[r->ticket] = r;
        s->incoming.push_back(r);
        ++s->submitted;
        ++s->outstanding;
        if (!s->have_first) {
            s->have_first = true;
            s->first_submit = r->submit;
        }
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return 0;
}

int sf_server_poll(sf_server* s, long long ticket, int* status) {
    if (!s || !status) return fail(SF_E_ARG, "sf_server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    RequestPtr r = find(s, ticket);
    if (!r) return fail(SF_E_ARG, "sf_server_poll: unknown ticket");
    *status = r->status;
    return 0;
}

int sf_server_wait(sf_server* s, long long ticket, double timeout_s, int* status) {
    if (!s || !status) return fail(SF_E_ARG, "sf_server_wait: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    RequestPtr r = find(s, ticket);
    if (!r) return fail(SF_E_ARG, "sf_server_wait: unknown ticket");
    s->cv_done.wait_for(lk, std::chrono::duration<double>(std::max(0.0, timeout_s)),
                        [&] { return r->status != 0; });
    *status = r->status;
    return 0;
}

int sf_server_result(sf_server* s, long long ticket, double* soa, double* time, long long* steps,
                     double* latency_ms) {
    if (!s) return fail(SF_E_ARG, "sf_server_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    RequestPtr r = find(s, ticket);
    if (!r) return fail(SF_E_ARG, "sf_server_result: unknown ticket");
    if (r->status == 0) return fail(SF_E_STATE, "sf_server_result: request still pending");
    if (r->status < 0) {
        const std::string msg = "request failed: " + r->error;
        s->table.erase(ticket);
        return fail(SF_E_NUMERIC, msg);
    }
    if (soa) std::memcpy(soa, r->state.data(), sizeof(double) * r->state.size());
    if (time) *time = r->time;
    if (steps) *steps = r->steps_done;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->done - r->submit).count();
    s->table.erase(ticket);
    return 0;
}

int sf_server_drain(sf_server* s, double timeout_s) {
    if (!s) return fail(SF_E_ARG, "sf_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(std::max(0.0, timeout_s)),
                                        [&] { return s->outstanding == 0; });
    if (!ok) return fail(SF_E_STATE, "sf_server_drain: requests still outstanding");
    return 0;
}

int sf_server_stats(sf_server* s, sf_serve_stats* st) {
    if (!s || !st) return fail(SF_E_ARG, "sf_server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->running = s->resident;
    st->queued = s->outstanding - s->resident;
    st->cell_updates = s->cell_updates;
    st->busy_s = s->busy_s;
    st->admit_s = s->admit_s;
    st->retire_s = s->retire_s;
    st->ticks = s->ticks;
    st->peak_items = s->peak_items;
    st->peak_nodes = s->peak_nodes;
    st->arena_nodes = s->arena_nodes;
    if (!s->latencies_ms.empty()) {
        std::vector<double> v = s->latencies_ms;
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto rank = [&](double p) {
            size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
            return v[std::min(v.size() - 1, k > 0 ? k - 1 : 0)];
        };
        st->p50_ms = rank(50.0);
        st->p99_ms = rank(99.0);
        double sum = 0.0;
        for (double x : v) sum += x;
        st->mean_ms = sum / (double)v.size();
        st->max_ms = v.back();
        const double span = std::chrono::duration<double>(s->last_done - s->first_submit).count();
        st->solves_per_s = span > 0.0 ? (double)s->completed / span : 0.0;
    }
    return 0;
}

int sf_server_set_shape(sf_server* s, int xlen, int ylen, int blocks_per_cu, int fused) {
    if (!s) return fail(SF_E_ARG, "sf_server_set_shape: null server");
    if ((xlen != 63 && xlen != 127 && xlen != 255) || ylen < 4 || ylen > 4096 ||
        blocks_per_cu < 0 || blocks_per_cu > 64)
        return fail(SF_E_ARG, "sf_server_set_shape: shape out of range (see sf_batch_set_tiles)");
    std::lock_guard<std::mutex> lk(s->mu);
    s->shape[0] = xlen;
    s->shape[1] = ylen;
    s->shape[2] = blocks_per_cu;
    s->shape[3] = fused ? 1 : 0;
    s->shape_pending = true;
    return 0;
}

int sf_server_set_launch(sf_server* s, int use_graph) {
    if (!s) return fail(SF_E_ARG, "sf_server_set_launch: null server");
    if (use_graph < 0 || use_graph > 2)
        return fail(SF_E_ARG, "sf_server_set_launch: 0 direct, 1 graph, 2 automatic");
    std::lock_guard<std::mutex> lk(s->mu);
    s->launch = use_graph;
    return 0;
}

int sf_server_reset_stats(sf_server* s) {
    if (!s) return fail(SF_E_ARG, "sf_server_reset_stats: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted = s->outstanding;
    s->completed = s->failed = 0;
    s->latencies_ms.clear();
    s->cell_updates = s->busy_s = s->admit_s = s->retire_s = 0.0;
    s->ticks = 0;
    s->peak_items = 0;
    s->peak_nodes = 0;
    s->have_first = false;
    return 0;
}

}  // extern "C"
