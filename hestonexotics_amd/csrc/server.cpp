// server.cpp -- serving loop with continuous batching.
//
// Requests are submitted from any thread and queue in arrival order. A scheduler thread admits
// the head of the queue while the device-memory budget (one resident arena) has room, packs the
// admitted requests into a batch and launches it on a free stream; requests that arrive while
// batches are in flight are admitted and launched on the next free stream as soon as memory and
// a stream slot free up, so their work overlaps the work already on the device. Results are the
// same as pricing each request alone: a request's tiles, reduction order and RNG stream depend
// only on the request.
//
// Large and small requests are kept apart, because a batch's results are ready when its
// longest request is. A request above kLargeWork path-steps runs in a batch of its own, one such
// batch at a time, with its grid capped at kLargeWavesPerSimd waves per SIMD (the FP32 kernels
// hold at most 5). The free wave slot on every SIMD lets a small batch launched later run beside
// it at once, instead of waiting for the persistent grid to drain. Small requests are batched in
// arrival order up to kSmallBatchWork path-steps, on the other streams. A large request waiting
// for its turn does not hold up the small ones behind it.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"

namespace {
constexpr double kLargeWork = 1e9;          // FP32 path-steps: about 2 ms alone on the GPU
constexpr double kSmallBatchWork = 5e8;     // work of one batch of small requests
constexpr uint32_t kLargeWavesPerSimd = 4;  // grid cap of a large request's batch
}  // namespace

struct hexo_server_state {
    struct Ticket {
        hexo_request req{};
        std::vector<double> T, K;
        std::vector<uint32_t> ns;
        std::vector<double> prices, ses;
        hexo_ticket_info info{};
        int state = 0;  // 0 queued, 1 in flight, 2 done, 3 failed/cancelled
        size_t bytes = 0;
        uint64_t n_out = 0;
        double work = 0.0;  // request_work
        bool large = false;
    };
    struct Flight {
        std::unique_ptr<hexo::Batch> b;
        std::vector<uint64_t> ids;
        hipEvent_t done = nullptr;
        int slot = 0;
        uint32_t batch_id = 0;
        size_t bytes = 0;
        bool large = false;
    };

    hexo_ctx* ctx = nullptr;
    hexo_server_config cfg{};
    hexo::Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<bool> slot_busy;
    std::vector<std::unique_ptr<hexo::PinnedBuf>> stages;  // one pinned staging buffer per stream slot
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    bool stopping = false;
    uint64_t next_ticket = 1;
    // shared: a waiter keeps its ticket alive across the condition wait, while another waiter on
    // the same ticket may consume and erase it
    std::unordered_map<uint64_t, std::shared_ptr<Ticket>> tickets;
    std::deque<uint64_t> queue;
    std::vector<Flight> flights;
    size_t budget = 0, reserved = 0, peak = 0;
    hexo_server_stats st{};
    uint32_t inject_launch = 0, inject_admit = 0;
    uint32_t batch_counter = 0;
    std::string worker_error;

    void run();
    void complete_ready();
    void launch_ready();
};

namespace {

void fail_ticket(hexo_server_state::Ticket& t, int status) {
    t.state = 3;
    t.info.status = status;
    t.info.done_s = hexo::now_s();
}

}  // namespace

void hexo_server_state::complete_ready() {
    for (size_t i = 0; i < flights.size();) {
        Flight& f = flights[i];
        hipError_t q = hipEventQuery(f.done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        const double t = hexo::now_s();
        const int status = q == hipSuccess ? HEXO_OK : HEXO_ERR_HIP;
        for (size_t k = 0; k < f.ids.size(); ++k) {
            auto it = tickets.find(f.ids[k]);
            if (it == tickets.end()) continue;
            Ticket& tk = *it->second;
            if (status == HEXO_OK) {
                const uint64_t base = f.b->out_base[k];
                tk.prices.assign(f.b->h_prices + base, f.b->h_prices + base + tk.n_out);
                tk.ses.assign(f.b->h_ses + base, f.b->h_ses + base + tk.n_out);
                tk.state = 2;
                tk.info.status = HEXO_OK;
                tk.info.done_s = t;
                st.completed++;
            } else {
                fail_ticket(tk, status);
            }
        }
        st.path_steps_done += f.b->path_steps;
        (void)hipEventDestroy(f.done);
        hexo::batch_destroy(*f.b);  // returns the region to the arena (the batch's work is complete)
        reserved -= f.bytes;
        slot_busy[f.slot] = false;
        flights.erase(flights.begin() + (long)i);
        cv.notify_all();
    }
}

void hexo_server_state::launch_ready() {
    for (;;) {
        if (queue.empty()) return;
        int slot = -1;
        for (size_t s = 0; s < slot_busy.size(); ++s)
            if (!slot_busy[s]) {
                slot = (int)s;
                break;
            }
        if (slot < 0) return;
        std::vector<uint64_t> ids;
        std::vector<hexo_request> reqs;
        size_t bytes = 16 * 256;
        const uint32_t max_batch = cfg.max_batch ? cfg.max_batch : 4096;
        uint32_t large_in_flight = 0;
        for (const Flight& f : flights) large_in_flight += f.large ? 1u : 0u;
        const bool large_slot = large_in_flight == 0;
        bool large = false;
        // a large request at the head runs alone, on a large slot, once the budget has room
        Ticket& head = *tickets[queue.front()];
        if (head.large && large_slot) {
            if (reserved + bytes + head.bytes > budget) return;  // wait for memory to free up
            bytes += head.bytes;
            ids.push_back(queue.front());
            reqs.push_back(head.req);
            queue.pop_front();
            large = true;
        }
        if (ids.empty()) {
            // small requests in arrival order, past large ones waiting for a large slot (they keep
            // their place), while the budget has room
            double work = 0.0;
            for (auto it = queue.begin(); it != queue.end() && ids.size() < max_batch;) {
                Ticket& t = *tickets[*it];
                if (t.large) {
                    ++it;
                    continue;
                }
                if (reserved + bytes + t.bytes > budget) break;
                if (!ids.empty() && work + t.work > kSmallBatchWork) break;
                bytes += t.bytes;
                work += t.work;
                ids.push_back(*it);
                reqs.push_back(t.req);
                it = queue.erase(it);
            }
        }
        if (ids.empty()) return;  // wait for memory or a slot to free up
        const double t_admit = hexo::now_s();
        if (inject_admit) {
            --inject_admit;
            for (uint64_t id : ids) fail_ticket(*tickets[id], HEXO_ERR_INJECTED);
            st.rejected += ids.size();
            cv.notify_all();
            continue;
        }
        std::unique_ptr<hexo::Batch> b;
        std::string err;
        hexo_tuning tun = ctx->tuning_now();
        if (large && (tun.waves_per_simd == 0 || tun.waves_per_simd > kLargeWavesPerSimd))
            tun.waves_per_simd = kLargeWavesPerSimd;
        hexo::BatchMemory mem;
        mem.arena = &arena;
        mem.stage = stages[slot].get();
        mem.events = false;
        int rc = hexo::batch_create(ctx, reqs.data(), (uint32_t)reqs.size(), streams[slot], mem, b, err, nullptr, &tun);
        if (rc == HEXO_ERR_NOMEM && !flights.empty()) {
            // fragmentation: put the requests back in arrival order and retry when a batch completes
            for (uint64_t id : ids) queue.push_back(id);
            std::sort(queue.begin(), queue.end());  // ticket ids grow with arrival
            return;
        }
        if (rc == HEXO_OK && inject_launch) {
            --inject_launch;
            rc = HEXO_ERR_INJECTED;
        }
        hipEvent_t ev = nullptr;
        if (rc == HEXO_OK) rc = hexo::batch_enqueue(*b, hexo::PathMode::Normal, false, nullptr, err);
        if (rc == HEXO_OK && (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                              hipEventRecord(ev, streams[slot]) != hipSuccess))
            rc = HEXO_ERR_HIP;
        const double t_launch = hexo::now_s();
        ++batch_counter;
        if (rc != HEXO_OK) {
            if (b) {
                (void)hipStreamSynchronize(streams[slot]);
                hexo::batch_destroy(*b);
            }
            if (ev) (void)hipEventDestroy(ev);
            worker_error = err;
            for (uint64_t id : ids) fail_ticket(*tickets[id], rc);
            cv.notify_all();
            continue;
        }
        for (uint64_t id : ids) {
            Ticket& t = *tickets[id];
            t.state = 1;
            t.info.admit_s = t_admit;
            t.info.launch_s = t_launch;
            t.info.batch_id = batch_counter;
            t.info.batch_size = (uint32_t)ids.size();
        }
        reserved += bytes;
        peak = std::max(peak, reserved);
        st.batches++;
        slot_busy[slot] = true;
        Flight f;
        f.b = std::move(b);
        f.ids = std::move(ids);
        f.done = ev;
        f.slot = slot;
        f.batch_id = batch_counter;
        f.bytes = bytes;
        f.large = large;
        flights.push_back(std::move(f));
    }
}

void hexo_server_state::run() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        complete_ready();
        if (!stopping || !queue.empty()) launch_ready();
        if (stopping && queue.empty() && flights.empty()) break;
        if (!flights.empty())
            cv.wait_for(lk, std::chrono::microseconds(50));
        else if (queue.empty() || stopping)
            cv.wait_for(lk, std::chrono::milliseconds(5));
        else
            cv.wait_for(lk, std::chrono::microseconds(200));
    }
}

extern "C" {

int hexo_server_start(hexo_ctx* ctx, const hexo_server_config* cfg) {
    if (!ctx) return HEXO_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (ctx->server_ref()) return ctx->set_error(HEXO_ERR_STATE, "server already running");
    if (cfg && (cfg->streams > 4)) return ctx->set_error(HEXO_ERR_INVALID, "streams must be <= 4");
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "hipSetDevice failed");
    auto s = std::make_shared<hexo_server_state>();
    s->ctx = ctx;
    if (cfg) s->cfg = *cfg;
    if (s->cfg.streams == 0) s->cfg.streams = 2;
    size_t budget = s->cfg.memory_budget_bytes;
    if (budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
        budget = std::min<size_t>(fr / 4, (size_t)8 << 30);
    }
    if (!s->arena.init(budget)) return ctx->set_error(HEXO_ERR_NOMEM, "could not reserve the serving memory budget");
    s->budget = s->arena.capacity();
    s->st.budget_bytes = s->budget;
    s->streams.resize(s->cfg.streams);
    s->slot_busy.assign(s->cfg.streams, false);
    for (uint32_t i = 0; i < s->cfg.streams; ++i) s->stages.push_back(std::make_unique<hexo::PinnedBuf>());
    for (auto& st : s->streams)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            st = nullptr;
            for (auto& t : s->streams)
                if (t) (void)hipStreamDestroy(t);
            return ctx->set_error(HEXO_ERR_HIP, "hipStreamCreate failed");
        }
    // the worker never outlives the state: hexo_server_stop joins it before releasing its reference
    hexo_server_state* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    std::lock_guard<std::mutex> gs(ctx->server_mu);
    ctx->server = std::move(s);
    return HEXO_OK;
}

int hexo_server_stop(hexo_ctx* ctx) {
    if (!ctx) return HEXO_ERR_INVALID;
    std::shared_ptr<hexo_server_state> s;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        std::lock_guard<std::mutex> gs(ctx->server_mu);
        s = std::move(ctx->server);
        ctx->server.reset();
    }
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv.notify_all();
    }
    s->worker.join();  // drains the queue and the batches in flight
    for (auto& st : s->streams) (void)hipStreamDestroy(st);
    s->streams.clear();
    s->arena.release();
    // calls still holding a reference (a wait on a finished ticket) end on the state they hold
    return HEXO_OK;
}

int hexo_submit(hexo_ctx* ctx, const hexo_request* req, uint64_t* ticket) {
    if (!ctx || !ticket) return HEXO_ERR_INVALID;
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::string why = hexo::validate(req);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!why.empty()) {
        s->st.rejected++;
        return ctx->set_error(HEXO_ERR_INVALID, why);
    }
    if (s->stopping) return ctx->set_error(HEXO_ERR_STATE, "server stopping");
    auto t = std::make_shared<hexo_server_state::Ticket>();
    t->req = *req;
    const uint64_t n_opts = hexo::n_opts_of(req);
    t->T.assign(req->expiries, req->expiries + req->n_expiries);
    t->ns.assign(req->n_strikes, req->n_strikes + req->n_expiries);
    t->K.assign(req->strikes, req->strikes + n_opts);
    t->req.expiries = t->T.data();
    t->req.n_strikes = t->ns.data();
    t->req.strikes = t->K.data();
    t->n_out = n_opts * hexo::n_payoffs(req->payoffs);
    t->bytes = hexo::request_device_bytes(req) + 4 * 256;
    t->work = hexo::request_work(req);
    t->large = t->work > kLargeWork;
    if (t->bytes + 16 * 256 > s->budget) {
        s->st.rejected++;
        return ctx->set_error(HEXO_ERR_ADMISSION, "request needs " + std::to_string(t->bytes) +
                                                      " device bytes, more than the serving budget");
    }
    t->info.arrival_s = hexo::now_s();
    t->info.status = 1;
    const uint64_t id = s->next_ticket++;
    s->tickets[id] = std::move(t);
    s->queue.push_back(id);
    s->st.submitted++;
    s->st.max_queue_depth = std::max<uint64_t>(s->st.max_queue_depth, s->queue.size());
    *ticket = id;
    s->cv.notify_all();
    return HEXO_OK;
}

int hexo_poll(hexo_ctx* ctx, uint64_t ticket, int* done) {
    if (!ctx || !done) return HEXO_ERR_INVALID;
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return HEXO_ERR_NOT_FOUND;
    *done = it->second->state >= 2 ? 1 : 0;
    return HEXO_OK;
}

int hexo_wait(hexo_ctx* ctx, uint64_t ticket, double timeout_s, double* prices, double* std_errs,
              hexo_ticket_info* info) {
    if (!ctx) return HEXO_ERR_INVALID;
    if (std::isnan(timeout_s)) return ctx->set_error(HEXO_ERR_INVALID, "timeout_s is NaN");
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return HEXO_ERR_NOT_FOUND;
    // Several threads may wait on one ticket: the first to wake up with the result consumes it.
    // The others find it gone from the map on their next wake-up and report HEXO_ERR_NOT_FOUND;
    // their reference keeps the ticket's memory valid until then.
    const std::shared_ptr<hexo_server_state::Ticket> t = it->second;
    auto gone = [&] {
        auto j = s->tickets.find(ticket);
        return j == s->tickets.end() || j->second != t;
    };
    auto ready = [&] { return t->state >= 2 || gone(); };
    // beyond 1e7 s (about 116 days) the deadline would overflow the clock's range: wait forever
    if (timeout_s < 0 || timeout_s > 1e7)
        s->cv.wait(lk, ready);
    else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
        return HEXO_ERR_TIMEOUT;
    if (gone()) return HEXO_ERR_NOT_FOUND;
    const int status = t->state == 2 ? HEXO_OK : t->info.status;
    if (info) *info = t->info;
    if (status == HEXO_OK) {
        if (prices) std::memcpy(prices, t->prices.data(), sizeof(double) * t->n_out);
        if (std_errs) std::memcpy(std_errs, t->ses.data(), sizeof(double) * t->n_out);
    }
    s->tickets.erase(ticket);
    s->cv.notify_all();  // other waiters on this ticket see it gone
    return status;
}

int hexo_cancel(hexo_ctx* ctx, uint64_t ticket) {
    if (!ctx) return HEXO_ERR_INVALID;
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return HEXO_ERR_NOT_FOUND;
    if (it->second->state != 0) return ctx->set_error(HEXO_ERR_STATE, "ticket already launched or finished");
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), ticket), s->queue.end());
    fail_ticket(*it->second, HEXO_ERR_CANCELLED);
    s->st.cancelled++;
    s->cv.notify_all();
    return HEXO_OK;
}

int hexo_server_stats_get(hexo_ctx* ctx, hexo_server_stats* out) {
    if (!ctx || !out) return HEXO_ERR_INVALID;
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->bytes_in_use = s->reserved;
    out->bytes_peak = s->peak;
    out->budget_bytes = s->budget;
    return HEXO_OK;
}

int hexo_inject_fault(hexo_ctx* ctx, uint32_t kind, uint32_t count) {
    if (!ctx || kind > 1) return HEXO_ERR_INVALID;
    const std::shared_ptr<hexo_server_state> s = ctx->server_ref();
    if (!s) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    std::lock_guard<std::mutex> lk(s->mu);
    if (kind == 0)
        s->inject_launch = count;
    else
        s->inject_admit = count;
    return HEXO_OK;
}

int hexo_server_run_trace(hexo_ctx* ctx, const hexo_request* reqs, const double* arrival_s, uint32_t n,
                          double* prices, double* std_errs, hexo_ticket_info* info) {
    if (!ctx || (!reqs && n) || (!arrival_s && n)) return HEXO_ERR_INVALID;
    if (!ctx->server_ref()) return ctx->set_error(HEXO_ERR_STATE, "server not running");
    for (uint32_t i = 0; i < n; ++i) {
        std::string why = hexo::validate(&reqs[i]);
        if (!why.empty()) return ctx->set_error(HEXO_ERR_INVALID, "request " + std::to_string(i) + ": " + why);
    }
    std::vector<uint64_t> out_base(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i)
        out_base[i + 1] = out_base[i] + hexo::n_opts_of(&reqs[i]) * hexo::n_payoffs(reqs[i].payoffs);
    std::vector<uint64_t> ids(n, 0);
    std::vector<int> sub_rc(n, HEXO_OK);
    std::vector<std::atomic<int>> submitted(n);
    for (auto& a : submitted) a.store(0);
    const double t0 = hexo::now_s();
    std::thread submitter([&] {
        for (uint32_t i = 0; i < n; ++i) {
            const double due = t0 + arrival_s[i];
            for (;;) {
                const double left = due - hexo::now_s();
                if (left <= 0) break;
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(left, 0.002)));
            }
            sub_rc[i] = hexo_submit(ctx, &reqs[i], &ids[i]);
            submitted[i].store(1, std::memory_order_release);
        }
    });
    int rc = HEXO_OK;
    for (uint32_t i = 0; i < n; ++i) {
        while (!submitted[i].load(std::memory_order_acquire)) std::this_thread::sleep_for(std::chrono::microseconds(100));
        if (sub_rc[i] != HEXO_OK) {
            if (info) {
                std::memset(&info[i], 0, sizeof(info[i]));
                info[i].status = sub_rc[i];
            }
            rc = rc == HEXO_OK ? sub_rc[i] : rc;
            continue;
        }
        hexo_ticket_info ti{};
        int w = hexo_wait(ctx, ids[i], -1.0, prices ? prices + out_base[i] : nullptr,
                          std_errs ? std_errs + out_base[i] : nullptr, &ti);
        ti.arrival_s -= t0;
        ti.admit_s -= t0;
        ti.launch_s -= t0;
        ti.done_s -= t0;
        ti.status = w;
        if (info) info[i] = ti;
        if (w != HEXO_OK && rc == HEXO_OK) rc = w;
    }
    submitter.join();
    return rc;
}

}  // extern "C"
