// abi.cpp -- extern "C" entry points declared in include/hexo.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "engine.hpp"

using hexo::PathMode;

namespace {

int no_ctx() { return HEXO_ERR_INVALID; }

struct DeviceGuard {
    explicit DeviceGuard(int dev) { ok = hipSetDevice(dev) == hipSuccess; }
    bool ok;
};

// RAII device buffer for the component entry points.
struct DBuf {
    void* p = nullptr;
    explicit DBuf(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) p = nullptr;
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
};

// The synchronous path: the batch's device region comes from the context's resident arena and its
// staging from the context's pinned buffer. Both grow (once, to the largest call so far) when a
// call does not fit; a call that fits allocates nothing.
int run_single(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, double* prices, double* ses) {
    std::string err;
    std::unique_ptr<hexo::Batch> b;
    size_t need = 0;
    hexo::BatchMemory mem;
    mem.arena = &ctx->sync_arena;
    mem.stage = &ctx->sync_stage;
    mem.events = false;
    mem.need = &need;
    // an arena too small (empty at the first call) reports the size the batch needs
    int rc = hexo::batch_create(ctx, reqs, n, ctx->stream, mem, b, err);
    if (rc == HEXO_ERR_NOMEM && need) {
        // grow; the arena is empty, every earlier call's batch being complete and released
        const size_t grown = std::max<size_t>({need, (size_t)16 << 20, 2 * ctx->sync_arena.capacity()});
        if (!ctx->sync_arena.init(grown) && !ctx->sync_arena.init(need))
            return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc of " + std::to_string(need) + " device bytes failed");
        rc = hexo::batch_create(ctx, reqs, n, ctx->stream, mem, b, err);
    }
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    rc = hexo::batch_enqueue(*b, PathMode::Normal, false, nullptr, err);
    if (rc == HEXO_OK) rc = hexo::batch_finish(*b, err);
    if (rc != HEXO_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return ctx->set_error(rc, err);
    }
    hexo::batch_copy_results(*b, prices, ses);
    return HEXO_OK;
}

}  // namespace

extern "C" {

const char* hexo_strerror(int status) {
    switch (status) {
        case HEXO_OK: return "ok";
        case HEXO_ERR_INVALID: return "invalid request or argument";
        case HEXO_ERR_NO_DEVICE: return "no usable HIP device";
        case HEXO_ERR_HIP: return "HIP runtime error";
        case HEXO_ERR_NOMEM: return "out of memory";
        case HEXO_ERR_ADMISSION: return "request exceeds the serving memory budget";
        case HEXO_ERR_NOT_FOUND: return "unknown ticket";
        case HEXO_ERR_CANCELLED: return "ticket cancelled";
        case HEXO_ERR_TIMEOUT: return "wait timed out";
        case HEXO_ERR_STATE: return "call not valid in the current state";
        case HEXO_ERR_INJECTED: return "injected failure";
        default: return "unknown status";
    }
}

const char* hexo_version(void) { return "hexo-mi355x 0.1 (gfx950; abi 1)"; }

int hexo_validate_request(const hexo_request* req) { return hexo::validate(req).empty() ? HEXO_OK : HEXO_ERR_INVALID; }

uint64_t hexo_request_n_outputs(const hexo_request* req) {
    if (!hexo::validate(req).empty()) return 0;
    return hexo::n_opts_of(req) * hexo::n_payoffs(req->payoffs);
}

int hexo_time_grid(const double* expiries, uint32_t n_expiries, const uint32_t* n_strikes, uint32_t steps,
                   uint64_t* fin_step, double* fin_w, uint64_t* total_steps) {
    if (!expiries || !n_strikes || n_expiries == 0) return HEXO_ERR_INVALID;
    for (uint32_t c = 0; c < n_expiries; ++c)
        if (!std::isfinite(expiries[c]) || !(expiries[c] > 0.0) || (c && !(expiries[c] > expiries[c - 1])))
            return HEXO_ERR_INVALID;
    std::vector<hexo::HostSeg> segs;
    uint64_t total = 0;
    if (!hexo::build_schedule(expiries, n_expiries, n_strikes, steps, segs, total)) return HEXO_ERR_INVALID;
    if (fin_step)
        for (uint32_t c = 0; c < n_expiries; ++c) fin_step[c] = 0;
    if (fin_w)
        for (uint32_t c = 0; c < n_expiries; ++c) fin_w[c] = 0.0;
    uint64_t k = 0;
    for (const auto& s : segs) {
        k += s.n_steps;
        for (const auto& f : s.fins) {
            if (fin_step) fin_step[f.chain] = k;
            if (fin_w) fin_w[f.chain] = f.w;
        }
    }
    if (total_steps) *total_steps = total;
    return HEXO_OK;
}

int hexo_ctx_create(int device, hexo_ctx** out) {
    if (!out) return HEXO_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return HEXO_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return HEXO_ERR_INVALID;
    hexo_ctx* c = new (std::nothrow) hexo_ctx();
    if (!c) return HEXO_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&c->prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return HEXO_ERR_HIP;
    }
    // hexo_price_noalloc's fixed region on this device (its address, looked up with the device
    // current); on failure that entry point reports it and everything else works
    if (hexo::scratch_region(device, &c->scratch, &c->scratch_bytes) != HEXO_OK) {
        c->scratch = nullptr;
        c->scratch_bytes = 0;
    }
    *out = c;
    return HEXO_OK;
}

int hexo_ctx_destroy(hexo_ctx* ctx) {
    if (!ctx) return HEXO_ERR_INVALID;
    if (ctx->server_ref()) (void)hexo_server_stop(ctx);
    (void)hipSetDevice(ctx->device);
    {
        // batches still alive lose their device state here; their handles stay valid for
        // hexo_batch_destroy, which then only frees the handle
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (hexo_batch* hb : ctx->batches) hb->b.reset();
        ctx->batches.clear();
    }
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->swift_dev) (void)hipFree(ctx->swift_dev);
    delete ctx;
    return HEXO_OK;
}

const char* hexo_ctx_last_error(hexo_ctx* ctx) {
    if (!ctx) return "no context";
    // a per-thread copy: another thread's failure may replace the context's message meanwhile
    thread_local std::string copy;
    std::lock_guard<std::mutex> g(ctx->err_mu);
    copy = ctx->last_error;
    return copy.c_str();
}

int hexo_device_info(hexo_ctx* ctx, char* name, size_t name_len, int* n_cu, int* clock_khz) {
    if (!ctx) return no_ctx();
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", ctx->prop.gcnArchName);
    }
    if (n_cu) *n_cu = ctx->prop.multiProcessorCount;
    if (clock_khz) *clock_khz = ctx->prop.clockRate;
    return HEXO_OK;
}

int hexo_price(hexo_ctx* ctx, const hexo_request* req, double* prices, double* std_errs) {
    if (!ctx) return no_ctx();
    if (!req || !prices) return ctx->set_error(HEXO_ERR_INVALID, "request and prices must be non-NULL");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    return run_single(ctx, req, 1, prices, std_errs);
}

int hexo_price_batch(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, double* prices, double* std_errs) {
    if (!ctx) return no_ctx();
    if (!reqs || !prices || n == 0) return ctx->set_error(HEXO_ERR_INVALID, "requests and prices must be non-NULL");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    return run_single(ctx, reqs, n, prices, std_errs);
}

int hexo_batch_create(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, hexo_batch** out) {
    if (!ctx) return no_ctx();
    if (!reqs || !out || n == 0) return ctx->set_error(HEXO_ERR_INVALID, "requests and out must be non-NULL");
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    auto* hb = new (std::nothrow) hexo_batch();
    if (!hb) return ctx->set_error(HEXO_ERR_NOMEM, "host allocation failed");
    std::string err;
    int rc = hexo::batch_create(ctx, reqs, n, ctx->stream, hexo::BatchMemory(), hb->b, err);
    if (rc == HEXO_OK) rc = hexo::batch_capture_graph(*hb->b, err);
    if (rc == HEXO_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
        rc = HEXO_ERR_HIP;
        err = "synchronise after upload failed";
    }
    if (rc != HEXO_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        delete hb;
        return ctx->set_error(rc, err);
    }
    ctx->batches.insert(hb);
    *out = hb;
    return HEXO_OK;
}

int hexo_batch_run(hexo_batch* hb) {
    if (!hb || !hb->b) return HEXO_ERR_INVALID;
    hexo_ctx* ctx = hb->b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    std::string err;
    int rc = hexo::batch_launch_graph(*hb->b, err);
    return rc == HEXO_OK ? rc : ctx->set_error(rc, err);
}

int hexo_batch_run_eager(hexo_batch* hb) {
    if (!hb || !hb->b) return HEXO_ERR_INVALID;
    hexo_ctx* ctx = hb->b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    std::string err;
    int rc = hexo::batch_enqueue(*hb->b, PathMode::Normal, true, nullptr, err);
    return rc == HEXO_OK ? rc : ctx->set_error(rc, err);
}

int hexo_batch_sync(hexo_batch* hb) {
    if (!hb || !hb->b) return HEXO_ERR_INVALID;
    hexo_ctx* ctx = hb->b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    std::string err;
    int rc = hexo::batch_finish(*hb->b, err);
    return rc == HEXO_OK ? rc : ctx->set_error(rc, err);
}

int hexo_batch_results(hexo_batch* hb, double* prices, double* std_errs) {
    if (!hb || !hb->b || !prices) return HEXO_ERR_INVALID;
    hexo_ctx* ctx = hb->b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    // the outputs land in pinned host buffers at the end of each run: wait for a run in flight
    if (hipStreamSynchronize(hb->b->stream) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "wait for the batch's run failed");
    hexo::batch_copy_results(*hb->b, prices, std_errs);
    return HEXO_OK;
}

int hexo_batch_stats_get(hexo_batch* hb, hexo_batch_stats* st) {
    if (!hb || !hb->b || !st) return HEXO_ERR_INVALID;
    const hexo::Batch& b = *hb->b;
    st->n_requests = b.n_req;
    st->path_steps = b.path_steps;
    st->nominal_path_steps = b.nominal_path_steps;
    st->last_run_ms = b.last_run_ms;
    st->path_kernel_ms = b.timed ? b.path_ms : 0.0;
    st->reduce_kernel_ms = b.timed ? b.reduce_ms : 0.0;
    st->path_kernel_launches = b.path_launches;
    return HEXO_OK;
}

int hexo_batch_branch_counts(hexo_batch* hb, uint64_t counts[4]) {
    if (!hb || !hb->b || !counts) return HEXO_ERR_INVALID;
    hexo::Batch& b = *hb->b;
    hexo_ctx* ctx = b.ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    std::string err;
    int rc = hexo::batch_enqueue(b, PathMode::Count, false, nullptr, err);
    if (rc == HEXO_OK) rc = hexo::batch_finish(b, err);
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    unsigned long long h[4];
    if (hipMemcpy(h, b.d_counts, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "copy of branch counts failed");
    for (int i = 0; i < 4; ++i) counts[i] = h[i];
    return HEXO_OK;
}

int hexo_batch_destroy(hexo_batch* hb) {
    if (!hb) return HEXO_ERR_INVALID;
    if (hb->b) {
        hexo_ctx* ctx = hb->b->ctx;
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        DeviceGuard dg(ctx->device);
        (void)hipStreamSynchronize(hb->b->stream);
        hb->b.reset();
        ctx->batches.erase(hb);
    }
    delete hb;
    return HEXO_OK;
}

int hexo_paths(hexo_ctx* ctx, const hexo_request* req, uint64_t first, uint64_t n, double* finals) {
    if (!ctx) return no_ctx();
    if (!req || !finals || n == 0) return ctx->set_error(HEXO_ERR_INVALID, "request and finals must be non-NULL");
    std::string why = hexo::validate(req);
    if (!why.empty()) return ctx->set_error(HEXO_ERR_INVALID, why);
    if (first >= req->n_paths || n > req->n_paths - first)
        return ctx->set_error(HEXO_ERR_INVALID, "path range outside the request");
    if (n > (1ull << 26)) return ctx->set_error(HEXO_ERR_INVALID, "at most 2^26 paths per hexo_paths call");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    // simulate from the tile boundary at or below `first` (path offsets are tile-aligned) and
    // return the requested window
    const uint64_t lead = first % hexo::kTile;
    hexo_request sub = *req;
    sub.path_offset = req->path_offset + (first - lead);
    sub.n_paths = n + lead;
    const uint64_t n_fin = sub.n_paths * (uint64_t)req->n_expiries * 2;
    // the NaN fill below counts 32-bit words in a uint32_t
    if (n_fin > (1ull << 30))
        return ctx->set_error(HEXO_ERR_INVALID, "at most 2^30 finals (paths x expiries x 2) per hexo_paths call");
    std::string err;
    std::unique_ptr<hexo::Batch> b;
    int rc = hexo::batch_create(ctx, &sub, 1, ctx->stream, hexo::BatchMemory(), b, err, &n_fin);
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    double* d_fin = b->pa[0].finals ? b->pa[0].finals : b->pa[1].finals;
    // expiries that are never finalised (trailing chains without strikes) read as NaN
    if (hexo::launch_fill_u32(reinterpret_cast<uint32_t*>(d_fin), (uint32_t)(n_fin * 2), 0xFFFFFFFFu, ctx->stream) !=
        hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "fill of the finals buffer failed");
    rc = hexo::batch_enqueue(*b, PathMode::Debug, false, d_fin, err);
    if (rc == HEXO_OK) rc = hexo::batch_finish(*b, err);
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    if (hipMemcpy(finals, d_fin + lead * (uint64_t)req->n_expiries * 2, sizeof(double) * n * req->n_expiries * 2,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "copy of finals failed");
    return HEXO_OK;
}

int hexo_geo_asian_price(hexo_ctx* ctx, const hexo_request* req, double* prices) {
    if (!ctx) return no_ctx();
    if (!req || !prices) return ctx->set_error(HEXO_ERR_INVALID, "request and prices must be non-NULL");
    // a batch of the request's control-variate block alone: creating it computes the exact prices
    hexo_request q = *req;
    q.payoffs = HEXO_ASIAN_GEO_CV;
    q.n_paths = hexo::kTile;
    q.path_offset = 0;
    std::string why = hexo::validate(&q);
    if (!why.empty()) return ctx->set_error(HEXO_ERR_INVALID, why);
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    std::string err;
    std::unique_ptr<hexo::Batch> b;
    int rc = hexo::batch_create(ctx, &q, 1, ctx->stream, hexo::BatchMemory(), b, err);
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    const uint64_t n_opts = hexo::n_opts_of(&q);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(prices, b->ga.geo, sizeof(double) * n_opts, hipMemcpyDeviceToHost) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "exact geometric prices failed");
    if (q.flags & HEXO_DISCOUNT) {
        uint64_t o = 0;
        for (uint32_t c = 0; c < q.n_expiries; ++c) {
            const double disc = std::exp(-q.r * q.expiries[c]);
            for (uint32_t j = 0; j < q.n_strikes[c]; ++j) prices[o++] *= disc;
        }
    }
    return HEXO_OK;
}

int hexo_rng_fill(hexo_ctx* ctx, uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths, uint64_t word0,
                  uint64_t n_words, uint32_t* out) {
    if (!ctx) return no_ctx();
    if (!out || n_paths == 0 || n_words == 0) return ctx->set_error(HEXO_ERR_INVALID, "empty request");
    if (n_paths > (1ull << 32) / n_words || n_paths * n_words > (1ull << 28))
        return ctx->set_error(HEXO_ERR_INVALID, "at most 2^28 words per call");
    if (word0 + n_words - 1 < word0 || ((word0 + n_words - 1) >> 2) > 0xFFFFFFFFull)
        return ctx->set_error(HEXO_ERR_INVALID, "word index beyond the 2^34-word stream");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    DBuf d(sizeof(uint32_t) * n_paths * n_words);
    if (!d.p) return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    if (hexo::launch_rng_fill(d.as<uint32_t>(), seed, stream, path0, n_paths, word0, n_words, ctx->stream) !=
            hipSuccess ||
        hipMemcpyAsync(out, d.p, sizeof(uint32_t) * n_paths * n_words, hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "rng fill failed");
    return HEXO_OK;
}

int hexo_philox4x32_10(hexo_ctx* ctx, const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out) {
    if (!ctx) return no_ctx();
    if (!ctr || !key || !out || n == 0 || n > (1ull << 26)) return ctx->set_error(HEXO_ERR_INVALID, "bad arguments");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    DBuf dc(16 * n), dk(8 * n), dout(16 * n);
    if (!dc.p || !dk.p || !dout.p) return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    if (hipMemcpyAsync(dc.p, ctr, 16 * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dk.p, key, 8 * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hexo::launch_philox(dc.as<uint32_t>(), dk.as<uint32_t>(), n, dout.as<uint32_t>(), ctx->stream) != hipSuccess ||
        hipMemcpyAsync(out, dout.p, 16 * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "philox failed");
    return HEXO_OK;
}

int hexo_normals(hexo_ctx* ctx, const uint32_t* words, uint64_t n, uint32_t precision, double* out) {
    if (!ctx) return no_ctx();
    if (!words || !out || n == 0 || n > (1ull << 28) || precision > 1u)
        return ctx->set_error(HEXO_ERR_INVALID, "bad arguments");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    DBuf dw(4 * n), dout(8 * n);
    if (!dw.p || !dout.p) return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    if (hipMemcpyAsync(dw.p, words, 4 * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hexo::launch_normals(dw.as<uint32_t>(), n, precision, dout.as<double>(), ctx->stream) != hipSuccess ||
        hipMemcpyAsync(out, dout.p, 8 * n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        return ctx->set_error(HEXO_ERR_HIP, "normals failed");
    return HEXO_OK;
}

int hexo_get_tuning(hexo_ctx* ctx, hexo_tuning* t) {
    if (!ctx || !t) return HEXO_ERR_INVALID;
    *t = ctx->tuning_now();
    return HEXO_OK;
}

int hexo_set_tuning(hexo_ctx* ctx, const hexo_tuning* t) {
    if (!ctx || !t) return HEXO_ERR_INVALID;
    const uint32_t p = t->paths_per_thread, b = t->block_threads;
    if (!(p >= 1 && p <= 4)) return ctx->set_error(HEXO_ERR_INVALID, "paths_per_thread must be 1, 2, 3 or 4");
    if (!(b == 64 || b == 128 || b == 256)) return ctx->set_error(HEXO_ERR_INVALID, "block_threads must be 64, 128 or 256");
    if (t->waves_per_simd > 8) return ctx->set_error(HEXO_ERR_INVALID, "waves_per_simd must be <= 8");
    ctx->set_tuning(*t);
    return HEXO_OK;
}

int hexo_autotune(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, uint32_t reps, hexo_tuning* best) {
    if (!ctx) return no_ctx();
    if (!reqs || n == 0) return ctx->set_error(HEXO_ERR_INVALID, "no requests to tune on");
    if (reps == 0) reps = 1;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    // trial shapes are passed to the batches directly: the context's shape (which the serving
    // thread may be reading) changes once, to the winner
    hexo_tuning winner = ctx->tuning_now();
    double best_ms = 1e300;
    const uint32_t ppts[] = {1, 2, 3, 4};
    const uint32_t blocks[] = {64, 128, 256};
    const uint32_t waves[] = {0, 2, 4};
    for (uint32_t p : ppts)
        for (uint32_t b : blocks)
            for (uint32_t w : waves) {
                const hexo_tuning trial{p, b, w, 0};
                std::string err;
                std::unique_ptr<hexo::Batch> bt;
                int rc = hexo::batch_create(ctx, reqs, n, ctx->stream, hexo::BatchMemory(), bt, err, nullptr, &trial);
                if (rc != HEXO_OK) return ctx->set_error(rc, err);
                double total = 0.0;
                for (uint32_t r = 0; r <= reps && rc == HEXO_OK; ++r) {  // first run warms up
                    rc = hexo::batch_enqueue(*bt, PathMode::Normal, true, nullptr, err);
                    if (rc == HEXO_OK) rc = hexo::batch_finish(*bt, err);
                    if (r > 0) total += bt->path_ms;
                }
                (void)hipStreamSynchronize(ctx->stream);
                if (rc != HEXO_OK) return ctx->set_error(rc, err);
                const double ms = total / reps;
                if (ms < best_ms) {
                    best_ms = ms;
                    winner = trial;
                }
            }
    ctx->set_tuning(winner);
    if (best) *best = winner;
    return HEXO_OK;
}

int hexo_copy_probe(hexo_ctx* ctx, uint64_t bytes, uint32_t reps, double* ms_per_copy) {
    if (!ctx) return no_ctx();
    if (bytes < 16 || bytes % 16 || bytes > (1ull << 34) || reps == 0)
        return ctx->set_error(HEXO_ERR_INVALID, "bytes must be a multiple of 16 in [16, 2^34]");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    DeviceGuard dg(ctx->device);
    DBuf src(bytes), dst(bytes);
    if (!src.p || !dst.p) return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "event");
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return ctx->set_error(HEXO_ERR_HIP, "event");
    }
    const uint32_t grid = (uint32_t)std::max(1, ctx->prop.multiProcessorCount * 8);
    int rc = HEXO_OK;
    if (hexo::launch_fill_u32(src.as<uint32_t>(), (uint32_t)std::min<uint64_t>(bytes / 4, 0xFFFFFFFFull), 1u,
                              ctx->stream) != hipSuccess ||
        hexo::launch_copy(src.p, dst.p, bytes, grid, ctx->stream) != hipSuccess ||
        hipEventRecord(e0, ctx->stream) != hipSuccess)
        rc = HEXO_ERR_HIP;
    for (uint32_t r = 0; r < reps && rc == HEXO_OK; ++r)
        if (hexo::launch_copy(src.p, dst.p, bytes, grid, ctx->stream) != hipSuccess) rc = HEXO_ERR_HIP;
    float ms = 0;
    if (rc == HEXO_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = HEXO_ERR_HIP;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HEXO_OK) return ctx->set_error(rc, "copy probe failed");
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return HEXO_OK;
}

}  // extern "C"
