// noalloc.hip -- hexo_price_noalloc: hexo_price without any device or pinned-host allocation.
//
// For callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc). The batch's device region (request tables, tile partials, group sums,
// prices) lives in `scratch`, a file-scope device array of kScratchBytes that every GPU the
// library runs on holds until the process ends; the host staging of the tables and results lives
// in a file-scope host array. In host code `scratch` names a host-side stand-in, not a device
// address: its address on each device is looked up once per device (hexo_ctx_create, with the
// device current) and only that address is used.
//
// Calls are serialised by the library (the array of a device and the host array are shared by
// every call); each call overwrites the array's previous contents. What must start at zero (the
// work-queue counters) is cleared by a fill kernel at every call. Requests whose region needs more
// than kScratchBytes, or whose tables and results need more than kStageBytes of host staging, are
// refused with HEXO_ERR_INVALID. Host-side planning (the time grid, the launch plan) uses the C++
// heap, as in hexo_price.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.hpp"

namespace hexo {
namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr size_t kStageBytes = (size_t)16 << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];
alignas(256) unsigned char host_stage[kStageBytes];

std::mutex g_mu;                       // serialises calls and the per-device lookup
void* g_scratch[kMaxDevices] = {};     // this process's address of `scratch` on each device

}  // namespace

// The address of `scratch` on the current device (looked up once per device). HEXO_OK or
// HEXO_ERR_HIP.
int scratch_region(int device, void** ptr, size_t* bytes) {
    std::lock_guard<std::mutex> g(g_mu);
    if (device < 0 || device >= kMaxDevices) return HEXO_ERR_INVALID;
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HEXO_ERR_HIP;
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    *bytes = kScratchBytes;
    return HEXO_OK;
}

}  // namespace hexo

extern "C" int hexo_price_noalloc(hexo_ctx* ctx, const hexo_request* req, double* prices, double* std_errs) {
    if (!ctx) return HEXO_ERR_INVALID;
    if (!req || !prices) return ctx->set_error(HEXO_ERR_INVALID, "request and prices must be non-NULL");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    std::lock_guard<std::mutex> gs(hexo::g_mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "hipSetDevice failed");
    if (!ctx->scratch)
        return ctx->set_error(HEXO_ERR_STATE, "the allocation-free region of this device could not be looked up");
    hexo::BatchMemory mem;
    mem.region = ctx->scratch;
    mem.region_bytes = ctx->scratch_bytes;
    mem.host = hexo::host_stage;
    mem.host_bytes = hexo::kStageBytes;
    mem.events = false;
    size_t need = 0;
    mem.need = &need;
    std::string err;
    std::unique_ptr<hexo::Batch> b;
    int rc = hexo::batch_create(ctx, req, 1, ctx->stream, mem, b, err);
    if (rc == HEXO_ERR_NOMEM)
        return ctx->set_error(HEXO_ERR_INVALID, "hexo_price_noalloc: " + err + " (device region " +
                                                    std::to_string(hexo::kScratchBytes) + " bytes, host staging " +
                                                    std::to_string(hexo::kStageBytes) + " bytes); use hexo_price");
    if (rc != HEXO_OK) return ctx->set_error(rc, err);
    rc = hexo::batch_enqueue(*b, hexo::PathMode::Normal, false, nullptr, err);
    if (rc == HEXO_OK) rc = hexo::batch_finish(*b, err);
    if (rc != HEXO_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        return ctx->set_error(rc, err);
    }
    hexo::batch_copy_results(*b, prices, std_errs);
    return HEXO_OK;
}
