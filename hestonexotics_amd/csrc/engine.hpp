// engine.hpp -- host engine of the Heston QE pricer: contexts, schedules, batches.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/hexo.h"
#include "kernels.hpp"
#include "layout.hpp"

namespace hexo {

// ---- the reference's time grid, as runs of equal steps ------------------------------------
struct HostFin {
    uint32_t chain;
    double T, tp, dt, w;
};
struct HostSeg {
    uint32_t n_steps = 0;
    double delta = 0.0;
    double dnext = 0.0;
    std::vector<HostFin> fins;
};
// Replays src/HSimulation.tpp:30-45 (with src/inc/AsianContract.h:35-38) in double and cuts the
// step sequence at every expiry finalisation. Returns false on bad input.
bool build_schedule(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps,
                    std::vector<HostSeg>& segs, uint64_t& total_steps);

// ---- geometric-average control variate (geo_host.cpp) -------------------------------------
constexpr uint32_t kGeoMaxPanels = 20000;  // 64-node panels per exact price before it reads NaN
constexpr uint32_t kGeoMaxPoints = 65536;  // grid steps of one control-variate expiry
constexpr uint64_t kGeoMaxRequestPoints = 1u << 22;  // grid points over a request's control-variate expiries
constexpr uint32_t kGeoMaxP0 = 128;        // panels per item evaluated in parallel
struct HostGeoItem {
    uint32_t pts_begin = 0;  // request-relative (t_j, v_j) pair index
    uint32_t n_pts = 0;
    uint32_t n_strikes = 0;
    uint32_t opt = 0;  // request-relative first option of the block
    uint32_t max_panels = 0;
    uint32_t p0 = 0;  // panels evaluated in parallel
    double c = 0.0, h = 0.0;
};
struct GeoPlan {
    std::vector<double> fin_W;  // per finalisation (schedule order): W of its Asian functional, 1 if unused
    std::vector<double> pts;    // (t_j, v_j) pairs of every control-variate expiry
    std::vector<HostGeoItem> items;
};
// Weights w[0..n] (and times t[0..n]) of chain `chain`'s Asian functional A = sum_j w_j X_{t_j}
// as the path kernels accumulate it over the run schedule; returns n, or -1 if never finalised.
int64_t functional_weights(const std::vector<HostSeg>& segs, uint32_t flags, uint32_t chain, std::vector<double>& t,
                           std::vector<double>& w);
bool plan_geo(const hexo_request* r, const std::vector<HostSeg>& segs, GeoPlan& g, std::string& why);
void gauss_legendre64(double* x, double* w);

std::string validate(const hexo_request* r);  // "" if valid, else the reason
uint32_t n_payoffs(uint32_t payoffs);
uint64_t n_opts_of(const hexo_request* r);

// ---- device memory ------------------------------------------------------------------------
// First-fit sub-allocator over one device allocation (serving memory budget).
class Arena {
   public:
    bool init(size_t bytes);
    void release();
    void* alloc(size_t bytes);
    void free(void* p);
    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }
    ~Arena() { release(); }

   private:
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::vector<std::pair<size_t, size_t>> free_;  // (offset, size), sorted by offset
    std::unordered_map<size_t, size_t> live_;      // offset -> size
};

// Reusable pinned host buffer (serving: one per stream slot, so no pinned allocation or free --
// which may synchronise the device -- happens per batch).
struct PinnedBuf {
    char* p = nullptr;
    size_t cap = 0;
    // hipHostMallocCoherent for buffers that kernels read or write in place (fine-grained: no
    // stale GPU-cached copy between calls)
    unsigned flags = hipHostMallocDefault;
    bool reserve(size_t bytes);
    ~PinnedBuf();
};

// ---- a batch of requests resident on the device --------------------------------------------
// kernel classes: 0 = FP64 kernel; FP32 kernels 1 + (martingale-corrected step) + 2 (geometric
// control variate's accumulation)
constexpr uint32_t kPathClasses = 5;
inline uint32_t class_precision(uint32_t cls) { return cls == 0 ? HEXO_FP64 : HEXO_FP32; }
inline bool class_martingale(uint32_t cls) { return cls == 2 || cls == 4; }
inline bool class_geo(uint32_t cls) { return cls >= 3; }

struct Batch {
    hexo_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    uint32_t n_req = 0;
    std::vector<uint64_t> out_base;  // per request, into the concatenated outputs
    uint64_t n_out_total = 0;
    double path_steps = 0.0, nominal_path_steps = 0.0;
    // device memory (one region)
    char* d_mem = nullptr;
    size_t d_bytes = 0;
    Arena* arena = nullptr;  // owner of d_mem when not hipMalloc'ed
    bool d_external = false;  // d_mem is a caller's fixed region (freed by nobody here)
    // host pinned staging
    char* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    bool h_external = false;  // h_stage belongs to a PinnedBuf
    size_t upload_bytes = 0;
    double* h_prices = nullptr;
    double* h_ses = nullptr;
    PathArgs pa[kPathClasses]{};
    uint32_t grid[kPathClasses]{}, block[kPathClasses]{}, ppt[kPathClasses]{};
    bool has_class[kPathClasses]{};
    ReduceArgs ra{};
    GeoArgs ga{};  // exact geometric prices, computed once after the upload (n_items 0: none)
    uint64_t n_group_threads = 0, n_out_threads = 0;
    uint32_t* d_counters = nullptr;
    unsigned long long* d_counts = nullptr;
    // graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // timing
    hipEvent_t ev[3]{};
    double last_run_ms = 0, path_ms = 0, reduce_ms = 0;
    uint64_t path_launches = 0;
    std::chrono::steady_clock::time_point t_run;
    bool timed = false;
    ~Batch();
};

}  // namespace hexo

struct hexo_ctx {
    int device = 0;
    hipDeviceProp_t prop{};
    hipStream_t stream = nullptr;
    hexo_tuning tuning{2, 256, 0, 0};  // guarded by tuning_mu: the serving thread reads it too
    std::mutex tuning_mu;
    std::recursive_mutex mu;
    std::mutex err_mu;  // guards last_error: failures may be reported from several threads at once
    std::string last_error;
    // the serving loop (server.cpp); every server call holds a reference while it runs, so a
    // concurrent hexo_server_stop never frees the state under it. Guarded by server_mu.
    std::mutex server_mu;
    std::shared_ptr<struct hexo_server_state> server;
    std::unordered_set<struct hexo_batch*> batches;  // live batches; released by hexo_ctx_destroy
    // SWIFT pricer workspace, grown on demand and kept for the context's lifetime (no per-call
    // allocation, which would also synchronise the device). The kernels read their inputs from
    // and write their results to the two fine-grained pinned buffers in place: the grid kernel's
    // inputs, strikes and outputs in swift_in_host; the pricing kernels' tables (which a first
    // kernel copies to the device) and results in swift_host. No copy-engine transfer, whose
    // hand-offs with the kernels would cost more than the transfers themselves.
    void* swift_dev = nullptr;
    size_t swift_dev_bytes = 0;
    hexo::PinnedBuf swift_host{nullptr, 0, hipHostMallocCoherent};
    hexo::PinnedBuf swift_in_host{nullptr, 0, hipHostMallocCoherent};
    // the synchronous pricing calls' device arena and staging buffer, likewise grown on demand and
    // resident (guarded by mu): hexo_price and hexo_price_batch allocate nothing once they fit
    hexo::Arena sync_arena;
    hexo::PinnedBuf sync_stage;
    // the allocation-free entry point's fixed region on this device (a file-scope device array)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    int set_error(int code, const std::string& msg) {
        std::lock_guard<std::mutex> g(err_mu);
        last_error = msg;
        return code;
    }
    hexo_tuning tuning_now() {
        std::lock_guard<std::mutex> g(tuning_mu);
        return tuning;
    }
    void set_tuning(const hexo_tuning& t) {
        std::lock_guard<std::mutex> g(tuning_mu);
        tuning = t;
        tuning.reserved = 0;
    }
    std::shared_ptr<struct hexo_server_state> server_ref() {
        std::lock_guard<std::mutex> g(server_mu);
        return server;
    }
};

struct hexo_batch {
    std::unique_ptr<hexo::Batch> b;
};

namespace hexo {

// Where a batch's memory comes from. By default the device region is hipMalloc'ed and the host
// staging buffer hipHostMalloc'ed per batch (long-lived batches); the serving loop and the
// synchronous entry points bring a resident arena and staging buffer instead, and the
// allocation-free entry point a fixed device region and host buffer.
struct BatchMemory {
    Arena* arena = nullptr;     // device region sub-allocated from this arena
    PinnedBuf* stage = nullptr;  // host staging reused from this buffer
    void* region = nullptr;     // fixed device region of region_bytes (nothing allocated)
    size_t region_bytes = 0;
    void* host = nullptr;       // fixed host staging buffer of host_bytes (nothing allocated)
    size_t host_bytes = 0;
    bool events = true;         // create the timing events (hexo_batch_run_eager, the autotuner)
    size_t* need = nullptr;     // out: the device bytes the batch needs, set when they do not fit
};

// Creates a batch: validates, builds the schedules, packs and uploads the tables, plans the
// launches. tuning: the launch shape to plan with (null: the context's, read once under tuning_mu).
int batch_create(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, hipStream_t stream, const BatchMemory& mem,
                 std::unique_ptr<Batch>& out, std::string& err, const uint64_t* finals_n = nullptr,
                 const hexo_tuning* tuning = nullptr);
// Device bytes a set of requests needs (admission accounting).
size_t batch_device_bytes(const hexo_request* reqs, uint32_t n);
size_t request_device_bytes(const hexo_request* r);
// Simulated path-steps of a request in FP32-kernel units (scheduling weight; 0 if invalid).
double request_work(const hexo_request* r);
// Enqueue the batch's launches (eager) on its stream; timed = record HIP events around them.
int batch_enqueue(Batch& b, PathMode mode, bool timed, double* d_finals, std::string& err);
int batch_capture_graph(Batch& b, std::string& err);
int batch_launch_graph(Batch& b, std::string& err);
int batch_finish(Batch& b, std::string& err);  // synchronise + timing readout
void batch_copy_results(const Batch& b, double* prices, double* ses);
void batch_destroy(Batch& b);

double now_s();

// The allocation-free entry point's fixed device region on `device` (current), looked up once per
// device (noalloc.hip). HEXO_OK or an error code.
int scratch_region(int device, void** ptr, size_t* bytes);

}  // namespace hexo
