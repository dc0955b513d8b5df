// engine.cpp -- host engine: request validation, the reference time grid, batch packing into
// one device region, launch planning, HIP-graph capture and timing.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

namespace hexo {

#define HIP_TRY(expr)                                                               \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            err = std::string(#expr) + " failed: " + hipGetErrorString(_e);         \
            return HEXO_ERR_HIP;                                                    \
        }                                                                           \
    } while (0)

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

uint32_t n_payoffs(uint32_t payoffs) { return (uint32_t)__builtin_popcount(payoffs & 15u); }

uint64_t n_opts_of(const hexo_request* r) {
    uint64_t n = 0;
    for (uint32_t c = 0; c < r->n_expiries; ++c) n += r->n_strikes[c];
    return n;
}

// ------------------------------------------------------------------------------------------
// Time grid. The reference steps with dt = T_c/steps while expiry c is the earliest unpriced one,
// accumulates time in double (state.cur_time += delta, src/HSimulation.tpp:84) and finalises
// expiry c after the first step with cur_time >= T_c (src/HSimulation.tpp:36, src/inc/SDE.h:30).
// The grid does not depend on the path, so it is replayed once here.
// ------------------------------------------------------------------------------------------
bool build_schedule(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps,
                    std::vector<HostSeg>& segs, uint64_t& total_steps) {
    segs.clear();
    total_steps = 0;
    if (n_exp == 0 || steps == 0) return false;
    uint64_t n_opts = 0;
    for (uint32_t c = 0; c < n_exp; ++c) n_opts += n_strikes[c];
    if (n_opts == 0) return false;
    const uint64_t guard = (uint64_t)n_exp * ((uint64_t)steps + 2) + 16;
    double t = 0.0, tp = 0.0, dt = T[0] / steps;
    uint32_t c = 0;
    uint64_t priced = 0;
    HostSeg cur;
    cur.delta = dt;
    tp = t;
    t += dt;
    cur.n_steps = 1;
    total_steps = 1;
    for (;;) {
        bool fin = false;
        while (priced < n_opts && t >= T[c]) {
            cur.fins.push_back(HostFin{c, T[c], tp, dt, (T[c] - tp) / dt});
            priced += n_strikes[c];
            fin = true;
            if (priced < n_opts) {
                ++c;
                dt = T[c] / steps;
            }
        }
        if (fin) {
            cur.dnext = priced < n_opts ? dt : 0.0;
            segs.push_back(cur);
            if (priced >= n_opts) break;
            cur = HostSeg();
            cur.delta = dt;
        }
        tp = t;
        t += dt;
        ++cur.n_steps;
        if (++total_steps > guard) return false;
    }
    return true;
}

std::string validate(const hexo_request* r) {
    if (!r) return "request is NULL";
    if (r->reserved != 0) return "reserved field must be 0";
    auto fin = [](double x) { return std::isfinite(x); };
    if (!fin(r->v0) || r->v0 < 0.0) return "v0 must be finite and >= 0";
    if (!fin(r->theta) || !(r->theta > 0.0)) return "theta must be finite and > 0";
    if (!fin(r->kappa) || !(r->kappa > 0.0)) return "kappa must be finite and > 0";
    if (!fin(r->sigma) || !(r->sigma > 0.0)) return "sigma must be finite and > 0";
    if (!fin(r->rho) || r->rho < -1.0 || r->rho > 1.0) return "rho must lie in [-1, 1]";
    if (!fin(r->S0) || !(r->S0 > 0.0)) return "S0 must be finite and > 0";
    if (!fin(r->r)) return "r must be finite";
    if (r->n_paths == 0) return "n_paths must be >= 1";
    // tile indices and work items are 32-bit and a batch holds fewer than 2^31 work items: at most
    // 2^31 - 1 tiles of 64 paths
    if (r->n_paths > kMaxPaths) return "n_paths must be <= 2^37 - 64 (at most 2^31 - 1 tiles of 64 paths)";
    if (r->path_offset > ~0ull - r->n_paths) return "path_offset + n_paths overflows";
    if (r->path_offset % kTile) return "path_offset must be a multiple of 64 (one tile)";
    if (r->steps == 0 || r->steps > (1u << 22)) return "steps must lie in [1, 2^22]";
    if (r->n_expiries == 0 || r->n_expiries > 4096) return "n_expiries must lie in [1, 4096]";
    if (!r->expiries || !r->n_strikes) return "expiries and n_strikes must be non-NULL";
    for (uint32_t c = 0; c < r->n_expiries; ++c) {
        if (!fin(r->expiries[c]) || !(r->expiries[c] > 0.0)) return "expiries must be finite and > 0";
        if (c > 0 && !(r->expiries[c] > r->expiries[c - 1])) return "expiries must be strictly increasing";
    }
    uint64_t n_opts = n_opts_of(r);
    if (n_opts == 0) return "at least one strike is required";
    if (n_opts > (1u << 20)) return "at most 2^20 strikes per request";
    if (!r->strikes) return "strikes must be non-NULL";
    for (uint64_t i = 0; i < n_opts; ++i)
        if (!fin(r->strikes[i])) return "strikes must be finite";
    if (r->payoffs == 0 || (r->payoffs & ~15u))
        return "payoffs must be a non-empty subset of ASIAN|EUROPEAN|ASIAN_GEO_CV|GEO_ASIAN";
    if (r->flags & ~31u) return "unknown flag bits";
    if ((r->flags & HEXO_CV_REGRESSION) &&
        (r->payoffs & (HEXO_ASIAN | HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN)) != (HEXO_ASIAN | HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN))
        return "HEXO_CV_REGRESSION needs the ASIAN, ASIAN_GEO_CV and GEO_ASIAN payoffs";
    if (r->precision > 2u) return "precision must be HEXO_FP64, HEXO_FP32 or HEXO_FP32_COMPENSATED";
    const uint64_t n_tiles = (r->n_paths + kTile - 1) / kTile;
    if (n_tiles * n_opts * n_payoffs(r->payoffs) > (1ull << 33)) return "n_paths * outputs too large for one request";
    std::vector<HostSeg> segs;
    uint64_t total = 0;
    if (!build_schedule(r->expiries, r->n_expiries, r->n_strikes, r->steps, segs, total))
        return "time grid could not be built";
    if (r->precision != HEXO_FP64) {
        // the FP32 step's moment ratio 2 m^2 / s2 ~ 2 V / (sigma^2 dt) must stay below sqrt(FLT_MAX)
        // (1.8e19). This bound leaves a factor of 9 for V above max(v0, theta); it only binds for
        // tiny sigma, where V hardly leaves its deterministic path between v0 and theta.
        double dmin = 1e300;
        for (const HostSeg& S : segs) dmin = std::min(dmin, S.delta);
        if (r->sigma * r->sigma * dmin < 1e-18 * std::max(r->v0, r->theta))
            return "sigma^2 dt below 1e-18 max(v0, theta) overflows the FP32 variance step; use HEXO_FP64";
    }
    // At rho = +-1 the geometric average's log is a functional of the variance path alone: its
    // characteristic function does not decay and the exact price's Fourier integral never
    // converges (it would read NaN)
    if ((r->payoffs & HEXO_ASIAN_GEO_CV) && std::fabs(r->rho) >= 1.0)
        return "the geometric control variate's exact price does not converge at rho = +-1";
    if (r->payoffs & (HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN)) {
        GeoPlan g;
        std::string why;
        if (!plan_geo(r, segs, g, why)) return why;
    }
    return "";
}

// ------------------------------------------------------------------------------------------
// Arena
// ------------------------------------------------------------------------------------------
bool Arena::init(size_t bytes) {
    release();
    bytes = (bytes + 255) & ~(size_t)255;
    if (hipMalloc(reinterpret_cast<void**>(&base_), bytes) != hipSuccess) {
        base_ = nullptr;
        return false;
    }
    cap_ = bytes;
    used_ = 0;
    free_.assign(1, {0, bytes});
    live_.clear();
    return true;
}

void Arena::release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = used_ = 0;
    free_.clear();
    live_.clear();
}

void* Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (size_t i = 0; i < free_.size(); ++i) {
        if (free_[i].second >= bytes) {
            const size_t off = free_[i].first;
            free_[i].first += bytes;
            free_[i].second -= bytes;
            if (free_[i].second == 0) free_.erase(free_.begin() + (long)i);
            live_[off] = bytes;
            used_ += bytes;
            return base_ + off;
        }
    }
    return nullptr;
}

void Arena::free(void* p) {
    if (!p) return;
    const size_t off = (size_t)(static_cast<char*>(p) - base_);
    auto it = live_.find(off);
    if (it == live_.end()) return;
    size_t sz = it->second;
    live_.erase(it);
    used_ -= sz;
    auto pos = std::lower_bound(free_.begin(), free_.end(), std::make_pair(off, (size_t)0));
    pos = free_.insert(pos, {off, sz});
    // coalesce with the next and previous blocks
    auto nx = pos + 1;
    if (nx != free_.end() && pos->first + pos->second == nx->first) {
        pos->second += nx->second;
        free_.erase(nx);
    }
    if (pos != free_.begin()) {
        auto pv = pos - 1;
        if (pv->first + pv->second == pos->first) {
            pv->second += pos->second;
            free_.erase(pos);
        }
    }
}

// ------------------------------------------------------------------------------------------
// Batch packing
// ------------------------------------------------------------------------------------------
namespace {

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct ReqPlan {
    std::vector<HostSeg> segs;
    GeoPlan geo;
    std::string geo_err;
    uint64_t total_steps = 0, n_opts = 0, n_out = 0;
    uint32_t n_pay = 0, n_tiles = 0, n_groups = 0, cls = 0;
    uint32_t n_fins = 0;
};

struct Layout {
    size_t reqs, segs, fins, opts, order[kPathClasses], item_end[kPathClasses], group_end, out_end;  // uploaded
    size_t upload_end;
    size_t geo_items, geo_pts, gl, geo_wave_end;                                      // uploaded
    size_t counters, counts, partials, groups, prices, ses, finals, geo, geo_panel, geo_pmax, geo_state;  // device
    size_t total;
};

Layout make_layout(uint32_t n_req, uint64_t n_seg, uint64_t n_fin, uint64_t n_opt, const uint32_t* n_cls,
                   uint64_t n_partials, uint64_t n_groups, uint64_t n_out, uint64_t n_finals, uint64_t n_geo_items,
                   uint64_t n_geo_pts, uint64_t n_geo_waves) {
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.reqs = take(sizeof(DevReq) * n_req);
    L.segs = take(sizeof(DevSeg) * std::max<uint64_t>(n_seg, 1));
    L.fins = take(sizeof(DevFin) * std::max<uint64_t>(n_fin, 1));
    L.opts = take(sizeof(DevOpt) * std::max<uint64_t>(n_opt, 1));
    for (uint32_t c = 0; c < kPathClasses; ++c) {
        L.order[c] = take(sizeof(uint32_t) * std::max<uint32_t>(n_cls[c], 1));
        L.item_end[c] = take(sizeof(uint32_t) * std::max<uint32_t>(n_cls[c], 1));
    }
    L.group_end = take(sizeof(uint64_t) * n_req);
    L.out_end = take(sizeof(uint64_t) * n_req);
    L.geo_items = take(sizeof(DevGeo) * std::max<uint64_t>(n_geo_items, 1));
    L.geo_pts = take(2 * sizeof(double) * std::max<uint64_t>(n_geo_pts, 1));
    L.gl = take(128 * sizeof(double));
    L.geo_wave_end = take(sizeof(uint32_t) * std::max<uint64_t>(n_geo_items, 1));
    L.upload_end = o;
    L.counters = take(256);
    L.counts = take(256);
    L.partials = take((size_t)16 * std::max<uint64_t>(n_partials, 1));
    L.groups = take((size_t)16 * std::max<uint64_t>(n_groups, 1));
    L.prices = take(sizeof(double) * n_out);
    L.ses = take(sizeof(double) * n_out);
    L.finals = take(sizeof(double) * std::max<uint64_t>(n_finals, 1));
    L.geo = take(sizeof(double) * std::max<uint64_t>(n_out, 1));
    L.geo_panel = take(64 * sizeof(double) * std::max<uint64_t>(n_geo_waves, 1));
    L.geo_pmax = take(sizeof(double) * std::max<uint64_t>(n_geo_waves, 1));
    L.geo_state = take(sizeof(double) * std::max<size_t>(geo_state_doubles((uint32_t)n_geo_items), 1));
    L.total = o;
    return L;
}

bool plan_request(const hexo_request* r, ReqPlan& p) {
    if (!build_schedule(r->expiries, r->n_expiries, r->n_strikes, r->steps, p.segs, p.total_steps)) return false;
    p.n_opts = n_opts_of(r);
    p.n_pay = n_payoffs(r->payoffs);
    p.n_out = p.n_opts * p.n_pay;
    p.n_tiles = (uint32_t)((r->n_paths + kTile - 1) / kTile);
    p.n_groups = (p.n_tiles + kGroupTiles - 1) / kGroupTiles;
    p.cls = r->precision == HEXO_FP64
                ? 0u
                : 1u + ((r->flags & HEXO_MARTINGALE) ? 1u : 0u) +
                      ((r->payoffs & (HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN)) ? 2u : 0u);
    if (!plan_geo(r, p.segs, p.geo, p.geo_err) && p.geo_err.empty()) p.geo_err = "geometric control variate";
    p.n_fins = 0;
    for (auto& s : p.segs) p.n_fins += (uint32_t)s.fins.size();
    return true;
}

}  // namespace

size_t request_device_bytes(const hexo_request* r) {
    ReqPlan p;
    if (!plan_request(r, p)) return 0;
    return sizeof(DevReq) + p.segs.size() * sizeof(DevSeg) + p.n_fins * sizeof(DevFin) + p.n_opts * sizeof(DevOpt) +
           32 + 16 + (size_t)16 * p.n_tiles * p.n_out + (size_t)16 * p.n_groups * p.n_out + 24 * p.n_out +
           p.geo.items.size() * (sizeof(DevGeo) + 4 + (kGeoMaxP0 + 1) * 520 + sizeof(double) * geo_state_doubles(1)) + p.geo.pts.size() * sizeof(double);
}

double request_work(const hexo_request* r) {
    ReqPlan p;
    if (!plan_request(r, p)) return 0.0;
    // the plain FP64 kernel takes about seven times as long per path-step as the FP32 ones
    return (double)r->n_paths * (double)p.total_steps * (r->precision == HEXO_FP64 ? 7.0 : 1.0);
}

size_t batch_device_bytes(const hexo_request* reqs, uint32_t n) {
    size_t s = 40 * 256 + 128 * sizeof(double);  // alignment slack of the region table, quadrature table
    for (uint32_t i = 0; i < n; ++i) s += request_device_bytes(&reqs[i]);
    return s;
}

bool PinnedBuf::reserve(size_t bytes) {
    if (bytes <= cap) return true;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
    if (hipHostMalloc(reinterpret_cast<void**>(&p), want, flags) != hipSuccess) {
        p = nullptr;
        return false;
    }
    cap = want;
    return true;
}

PinnedBuf::~PinnedBuf() {
    if (p) (void)hipHostFree(p);
}

int batch_create(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, hipStream_t stream, const BatchMemory& mem,
                 std::unique_ptr<Batch>& out, std::string& err, const uint64_t* finals_n,
                 const hexo_tuning* tuning) {
    if (n == 0 || !reqs) {
        err = "empty batch";
        return HEXO_ERR_INVALID;
    }
    const hexo_tuning tun = tuning ? *tuning : ctx->tuning_now();
    // tiles per work item of each kernel class: 1 for the FP64 kernel, the FP32 kernels' paths per
    // thread; a wave stores one partial per item and output (layout.hpp, kChunkTiles), or one per
    // tile when its tiles do not divide a chunk (3 paths per thread)
    uint32_t cls_ppt[kPathClasses], cls_store[kPathClasses];
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) {
        cls_ppt[cls] = cls == 0 ? 1u : tun.paths_per_thread;
        cls_store[cls] = kChunkTiles % cls_ppt[cls] == 0 ? cls_ppt[cls] : 1u;
    }
    std::vector<ReqPlan> plans(n);
    uint64_t n_seg = 0, n_fin = 0, n_opt = 0, n_partials = 0, n_groups = 0, n_out = 0;
    uint64_t n_geo_items = 0, n_geo_pts = 0, n_geo_waves = 0;
    uint32_t n_cls[kPathClasses] = {};
    for (uint32_t i = 0; i < n; ++i) {
        std::string why = validate(&reqs[i]);
        if (!why.empty()) {
            err = "request " + std::to_string(i) + ": " + why;
            return HEXO_ERR_INVALID;
        }
        if (!plan_request(&reqs[i], plans[i])) {
            err = "request " + std::to_string(i) + ": time grid could not be built";
            return HEXO_ERR_INVALID;
        }
        const ReqPlan& p = plans[i];
        if (!p.geo_err.empty()) {
            err = "request " + std::to_string(i) + ": " + p.geo_err;
            return HEXO_ERR_INVALID;
        }
        n_geo_items += p.geo.items.size();
        n_geo_pts += p.geo.pts.size() / 2;
        for (const HostGeoItem& it : p.geo.items) n_geo_waves += it.p0 + 1;
        n_seg += p.segs.size();
        n_fin += p.n_fins;
        n_opt += p.n_opts;
        n_partials += (uint64_t)((p.n_tiles + cls_store[p.cls] - 1) / cls_store[p.cls]) * p.n_out;
        n_groups += (uint64_t)p.n_groups * p.n_out;
        n_out += p.n_out;
        n_cls[p.cls]++;
    }
    auto b = std::make_unique<Batch>();
    b->ctx = ctx;
    b->stream = stream;
    b->n_req = n;
    b->n_out_total = n_out;
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) b->ppt[cls] = cls_ppt[cls];
    const uint64_t n_finals = finals_n ? *finals_n : 0;
    const Layout L =
        make_layout(n, n_seg, n_fin, n_opt, n_cls, n_partials, n_groups, n_out, n_finals, n_geo_items, n_geo_pts,
                    n_geo_waves);

    // the device region must fit a fixed region or the arena before anything is staged
    if ((mem.region && L.total > mem.region_bytes) ||
        (!mem.region && mem.arena && L.total > mem.arena->capacity() - mem.arena->in_use())) {
        if (mem.need) *mem.need = L.total;
        err = "the batch needs " + std::to_string(L.total) + " device bytes; " +
              (mem.region ? "the fixed region holds " + std::to_string(mem.region_bytes)
                          : "the arena has " + std::to_string(mem.arena->capacity() - mem.arena->in_use()) + " free");
        return HEXO_ERR_NOMEM;
    }

    // host staging: the uploaded tables, then the result buffers
    b->upload_bytes = L.upload_end;
    b->h_stage_bytes = L.upload_end + 2 * sizeof(double) * n_out;
    if (mem.host) {
        if (b->h_stage_bytes > mem.host_bytes) {
            err = "the batch needs " + std::to_string(b->h_stage_bytes) + " host staging bytes; the fixed buffer holds " +
                  std::to_string(mem.host_bytes);
            return HEXO_ERR_NOMEM;
        }
        b->h_stage = static_cast<char*>(mem.host);
        b->h_external = true;
    } else if (mem.stage) {
        if (!mem.stage->reserve(b->h_stage_bytes)) {
            err = "hipHostMalloc of the staging buffer failed";
            return HEXO_ERR_NOMEM;
        }
        b->h_stage = mem.stage->p;
        b->h_external = true;
    } else if (hipHostMalloc(reinterpret_cast<void**>(&b->h_stage), b->h_stage_bytes, hipHostMallocDefault) !=
               hipSuccess) {
        b->h_stage = nullptr;
        err = "hipHostMalloc of the staging buffer failed";
        return HEXO_ERR_NOMEM;
    }
    std::memset(b->h_stage, 0, L.upload_end);
    b->h_prices = reinterpret_cast<double*>(b->h_stage + L.upload_end);
    b->h_ses = b->h_prices + n_out;
    char* H = b->h_stage;
    DevReq* hr = reinterpret_cast<DevReq*>(H + L.reqs);
    DevSeg* hs = reinterpret_cast<DevSeg*>(H + L.segs);
    DevFin* hf = reinterpret_cast<DevFin*>(H + L.fins);
    DevOpt* ho = reinterpret_cast<DevOpt*>(H + L.opts);
    uint64_t* hge = reinterpret_cast<uint64_t*>(H + L.group_end);
    uint64_t* hoe = reinterpret_cast<uint64_t*>(H + L.out_end);
    DevGeo* hgi = reinterpret_cast<DevGeo*>(H + L.geo_items);
    double* hgp = reinterpret_cast<double*>(H + L.geo_pts);
    gauss_legendre64(reinterpret_cast<double*>(H + L.gl), reinterpret_cast<double*>(H + L.gl) + 64);
    uint32_t* hgw = reinterpret_cast<uint32_t*>(H + L.geo_wave_end);
    uint64_t geo_i = 0, geo_p = 0, geo_w = 0;

    const double L2E = 1.4426950408889634074;  // log2(e)
    uint64_t seg_i = 0, fin_i = 0, opt_i = 0, part_i = 0, grp_i = 0, out_i = 0, ge = 0;
    b->out_base.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const hexo_request& q = reqs[i];
        const ReqPlan& p = plans[i];
        DevReq& d = hr[i];
        d.v0 = q.v0;
        d.theta = q.theta;
        d.rho = q.rho;
        d.kappa = q.kappa;
        d.sigma = q.sigma;
        d.S0 = q.S0;
        d.r_drift = (q.flags & HEXO_DRIFT) ? q.r : 0.0;
        d.log_S0 = std::log(q.S0);
        d.n_paths = q.n_paths;
        d.path_offset = q.path_offset;
        d.partial_base = part_i;
        d.group_base = grp_i;
        d.seed_lo = (uint32_t)q.seed;
        d.seed_hi = (uint32_t)(q.seed >> 32);
        d.seg_begin = (uint32_t)seg_i;
        d.n_segs = (uint32_t)p.segs.size();
        d.n_opts = (uint32_t)p.n_opts;
        d.n_pay = p.n_pay;
        d.payoffs = q.payoffs;
        d.flags = q.flags;
        d.precision = q.precision;
        d.n_tiles = p.n_tiles;
        d.opt_begin = (uint32_t)opt_i;
        d.out_base = (uint32_t)out_i;
        d.n_exp = q.n_expiries;
        d.n_groups = p.n_groups;
        d.v0_f = (float)q.v0;
        d.store_tiles = cls_store[p.cls];
        // options (chain-major) with their expiry's discount factor
        uint64_t oo = 0;
        std::vector<uint32_t> chain_opt_begin(q.n_expiries);
        for (uint32_t c = 0; c < q.n_expiries; ++c) {
            chain_opt_begin[c] = (uint32_t)oo;
            const double disc = (q.flags & HEXO_DISCOUNT) ? std::exp(-q.r * q.expiries[c]) : 1.0;
            for (uint32_t j = 0; j < q.n_strikes[c]; ++j, ++oo) {
                DevOpt& O = ho[opt_i + oo];
                O.k_abs = q.strikes[oo];
                O.k_rel = (float)(q.strikes[oo] / q.S0);
                O.disc = disc;
            }
        }
        // runs of equal steps with their FP32 constants (computed in FP64)
        size_t fin_local = 0;  // index into the request's per-finalisation W
        for (const HostSeg& S : p.segs) {
            DevSeg& g = hs[seg_i++];
            const double dl = S.delta, k = q.kappa, th = q.theta, e = q.sigma, rho = q.rho;
            const double dd = std::exp(-k * dl), omd = -std::expm1(-k * dl);
            g.omd = (float)omd;
            g.th = (float)th;
            // half the conditional variance's coefficients (the FP32 step forms s2 / 2)
            g.c1h = (float)(0.5 * e * e * dd * omd / k);
            g.c2h = (float)(th * e * e * omd * omd / (4.0 * k));
            const double kc = 0.5 * dl * (k * rho / e - 0.5);
            const double kd = rho / e;
            const double K3 = 0.5 * dl * (1.0 - rho * rho);
            // g0 = K0 + 2 th kc (src/HSimulation.tpp:75-80 with the drift extension), taken with the
            // FP32 level th_f in both terms since the FP32 variance reverts to th_f: the rho/sigma
            // parts cancel exactly, leaving (r - th_f/2) dt. With the exact theta in K0 they would
            // leave rho kappa dt (th_f - theta) / sigma per step, 1e-5 at sigma = 1e-6.
            g.g0 = (float)((d.r_drift * dl - 0.5 * (double)g.th * dl) * L2E);
            // K2 = kc + kd, K1 = kc - kd; kb = -(K1 + K2 e^-k dl) = kd omd - kc (1 + e^-k dl)
            g.ka = (float)((kc + kd) * L2E);
            g.kb = (float)((kd * omd - kc * (1.0 + dd)) * L2E);
            g.k3 = (float)(K3 * L2E * L2E);
            const double A = kc + kd + 0.5 * K3;  // K2 + K4/2
            g.mA = (float)A;
            g.ma = (float)(A * L2E);
            g.mh = (float)(0.5 * K3 * L2E);
            g.mr = (float)(d.r_drift * dl * L2E);
            g.delta_f = (float)dl;
            g.dnext_f = (float)S.dnext;
            g.n_steps = S.n_steps;
            g.fin_begin = (uint32_t)fin_i;
            g.fin_count = (uint32_t)S.fins.size();
            g.delta = dl;
            g.dnext = S.dnext;
            for (const HostFin& F : S.fins) {
                DevFin& f = hf[fin_i++];
                f.chain = F.chain;
                f.opt_begin = chain_opt_begin[F.chain];
                f.n_strikes = q.n_strikes[F.chain];
                f.w_f = (float)F.w;
                f.invT_f = (float)(1.0 / F.T);
                f.hwdt_f = (float)(0.5 * (F.T - F.tp));
                f.T = F.T;
                f.tp = F.tp;
                f.dt = F.dt;
                const double W = p.geo.fin_W[fin_local++];
                f.gW = W;
                f.gW_f = (float)W;
                f.gInvW_f = (float)(1.0 / W);
            }
        }
        // exact-price items of the geometric control variate: points rebased into the batch's
        // list; outputs follow the Asian and European blocks
        if (!p.geo.items.empty()) {
            const uint64_t cvpay = n_payoffs(q.payoffs & (HEXO_ASIAN | HEXO_EUROPEAN));
            for (const HostGeoItem& it : p.geo.items) {
                DevGeo& g = hgi[geo_i++];
                g.req = i;
                g.pts_begin = (uint32_t)(geo_p + it.pts_begin);
                g.n_pts = it.n_pts;
                g.n_strikes = it.n_strikes;
                g.opt = (uint32_t)(opt_i + it.opt);
                g.max_panels = it.max_panels;
                g.out = out_i + cvpay * p.n_opts + it.opt;
                g.c = it.c;
                g.h = it.h;
                g.wave_begin = (uint32_t)geo_w;
                g.p0 = it.p0;
                geo_w += it.p0 + 1;
                hgw[geo_i - 1] = (uint32_t)geo_w;
            }
            std::memcpy(hgp + 2 * geo_p, p.geo.pts.data(), sizeof(double) * p.geo.pts.size());
            geo_p += p.geo.pts.size() / 2;
        }
        b->out_base[i] = out_i;
        opt_i += p.n_opts;
        part_i += (uint64_t)((p.n_tiles + cls_store[p.cls] - 1) / cls_store[p.cls]) * p.n_out;
        grp_i += (uint64_t)p.n_groups * p.n_out;
        out_i += p.n_out;
        ge += (uint64_t)p.n_groups * p.n_out;
        hge[i] = ge;
        hoe[i] = out_i;
        b->path_steps += (double)q.n_paths * (double)p.total_steps;
        b->nominal_path_steps += (double)q.n_paths * (double)q.steps * (double)q.n_expiries;
    }
    b->n_group_threads = ge;
    b->n_out_threads = out_i;

    // work queues: per kernel class, requests longest first
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) {
        std::vector<uint32_t> idx;
        for (uint32_t i = 0; i < n; ++i)
            if (plans[i].cls == cls) idx.push_back(i);
        b->has_class[cls] = !idx.empty();
        if (idx.empty()) continue;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
            if (plans[x].total_steps != plans[y].total_steps) return plans[x].total_steps > plans[y].total_steps;
            return reqs[x].n_paths > reqs[y].n_paths;
        });
        uint32_t* ord = reinterpret_cast<uint32_t*>(H + L.order[cls]);
        uint32_t* ie = reinterpret_cast<uint32_t*>(H + L.item_end[cls]);
        uint64_t acc = 0;
        const uint32_t ppt = b->ppt[cls];
        for (size_t s = 0; s < idx.size(); ++s) {
            ord[s] = idx[s];
            acc += (plans[idx[s]].n_tiles + ppt - 1) / ppt;
            if (acc >= (1ull << 31)) {
                batch_destroy(*b);
                err = "too many work items in one batch";
                return HEXO_ERR_INVALID;
            }
            ie[s] = (uint32_t)acc;
        }
        b->pa[cls].n_sorted = (uint32_t)idx.size();
        b->pa[cls].total_items = (uint32_t)acc;
    }

    // device region
    b->d_bytes = L.total;
    if (mem.region) {
        b->d_mem = static_cast<char*>(mem.region);
        b->d_external = true;
    } else if (mem.arena) {
        b->d_mem = static_cast<char*>(mem.arena->alloc(L.total));
        if (!b->d_mem) {  // free space, but fragmented
            if (mem.need) *mem.need = L.total;
            batch_destroy(*b);
            err = "device arena exhausted";
            return HEXO_ERR_NOMEM;
        }
        b->arena = mem.arena;
    } else if (hipMalloc(reinterpret_cast<void**>(&b->d_mem), L.total) != hipSuccess) {
        b->d_mem = nullptr;
        batch_destroy(*b);
        err = "hipMalloc of " + std::to_string(L.total) + " bytes failed";
        return HEXO_ERR_NOMEM;
    }
    char* D = b->d_mem;
    b->d_counters = reinterpret_cast<uint32_t*>(D + L.counters);
    b->d_counts = reinterpret_cast<unsigned long long*>(D + L.counts);
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) {
        PathArgs& a = b->pa[cls];
        a.reqs = reinterpret_cast<const DevReq*>(D + L.reqs);
        a.segs = reinterpret_cast<const DevSeg*>(D + L.segs);
        a.fins = reinterpret_cast<const DevFin*>(D + L.fins);
        a.opts = reinterpret_cast<const DevOpt*>(D + L.opts);
        a.order = reinterpret_cast<const uint32_t*>(D + L.order[cls]);
        a.item_end = reinterpret_cast<const uint32_t*>(D + L.item_end[cls]);
        a.counter = b->d_counters + cls;
        a.partials = D + L.partials;
        a.finals = n_finals ? reinterpret_cast<double*>(D + L.finals) : nullptr;
        a.counts = nullptr;
    }
    ReduceArgs& ra = b->ra;
    ra.reqs = reinterpret_cast<const DevReq*>(D + L.reqs);
    ra.opts = reinterpret_cast<const DevOpt*>(D + L.opts);
    ra.geo = geo_i ? reinterpret_cast<const double*>(D + L.geo) : nullptr;
    GeoArgs& ga = b->ga;
    ga.reqs = ra.reqs;
    ga.opts = ra.opts;
    ga.items = reinterpret_cast<const DevGeo*>(D + L.geo_items);
    ga.pts = reinterpret_cast<const double*>(D + L.geo_pts);
    ga.gl = reinterpret_cast<const double*>(D + L.gl);
    ga.geo = reinterpret_cast<double*>(D + L.geo);
    ga.n_items = (uint32_t)geo_i;
    ga.n_waves = (uint32_t)geo_w;
    ga.wave_end = reinterpret_cast<const uint32_t*>(D + L.geo_wave_end);
    ga.panel = reinterpret_cast<double*>(D + L.geo_panel);
    ga.panel_max = reinterpret_cast<double*>(D + L.geo_pmax);
    ga.state = reinterpret_cast<double*>(D + L.geo_state);
    ra.req_of = nullptr;
    ra.group_end = reinterpret_cast<const uint64_t*>(D + L.group_end);
    ra.out_end = reinterpret_cast<const uint64_t*>(D + L.out_end);
    ra.n_req = n;
    ra.partials = D + L.partials;
    ra.groups = D + L.groups;
    ra.prices = reinterpret_cast<double*>(D + L.prices);
    ra.std_errs = reinterpret_cast<double*>(D + L.ses);

    // launch shapes
    const int n_cu = ctx->prop.multiProcessorCount > 0 ? ctx->prop.multiProcessorCount : 256;
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) {
        if (!b->has_class[cls]) continue;
        const uint32_t prec = class_precision(cls);
        const uint32_t block = cls == 0 ? 256u : tun.block_threads;
        const size_t lds = path_lds_bytes(prec, b->ppt[cls], block);
        int bpc = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc,
                                                          path_kernel_symbol(prec, b->ppt[cls], PathMode::Normal,
                                                                             class_martingale(cls), class_geo(cls)),
                                                          (int)block, lds) != hipSuccess ||
            bpc <= 0)
            bpc = std::max(1, 1024 / (int)block);
        if (cls != 0 && tun.waves_per_simd)
            bpc = std::max(1, std::min(bpc, (int)(tun.waves_per_simd * 4 * 64 / block)));
        const uint64_t waves_per_block = block / kTile;
        const uint64_t need = (b->pa[cls].total_items + waves_per_block - 1) / waves_per_block;
        b->grid[cls] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)n_cu * (uint64_t)bpc));
        b->block[cls] = block;
    }
    for (auto& e : b->ev)
        if (mem.events && hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            batch_destroy(*b);
            err = "hipEventCreate failed";
            return HEXO_ERR_HIP;
        }
    if (hipMemcpyAsync(b->d_mem, b->h_stage, b->upload_bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
        (void)hipStreamSynchronize(stream);  // a copy may still be writing the region or reading the staging buffer
        batch_destroy(*b);
        err = "upload of the batch tables failed";
        return HEXO_ERR_HIP;
    }
    // the exact geometric prices depend on the requests only: computed once, before any run
    if (ga.n_items && launch_geo(ga, stream) != hipSuccess) {
        // the upload is in flight: the region and the (possibly shared) staging buffer are handed
        // back only once it is done
        (void)hipStreamSynchronize(stream);
        batch_destroy(*b);
        err = "launch of the geometric control-variate prices failed";
        return HEXO_ERR_HIP;
    }
    if (ga.n_items) {
        // A request whose exact price did not converge (NaN) is refused rather than priced NaN: its
        // variance process stays near 0 (v0 = 0 with a Feller ratio 2 kappa theta / sigma^2 near 0),
        // so the characteristic function of the geometric average's log hardly decays. The result
        // buffer is free until the first run.
        if (hipMemcpyAsync(b->h_prices, ga.geo, sizeof(double) * n_out, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            (void)hipStreamSynchronize(stream);
            batch_destroy(*b);
            err = "read-back of the geometric control-variate prices failed";
            return HEXO_ERR_HIP;
        }
        for (uint32_t i = 0; i < n; ++i) {
            if (!(reqs[i].payoffs & HEXO_ASIAN_GEO_CV)) continue;
            const uint64_t first = b->out_base[i] + n_payoffs(reqs[i].payoffs & (HEXO_ASIAN | HEXO_EUROPEAN)) * plans[i].n_opts;
            for (uint64_t o = first; o < first + plans[i].n_opts; ++o)
                if (std::isnan(b->h_prices[o])) {
                    const hexo_request& q = reqs[i];
                    err = "request " + std::to_string(i) +
                          ": the geometric control variate's exact price did not converge (v0 " + std::to_string(q.v0) +
                          ", Feller ratio 2 kappa theta / sigma^2 = " +
                          std::to_string(2.0 * q.kappa * q.theta / (q.sigma * q.sigma)) +
                          ": the variance stays near 0); price it without HEXO_ASIAN_GEO_CV";
                    batch_destroy(*b);
                    return HEXO_ERR_INVALID;
                }
        }
    }
    out = std::move(b);
    return HEXO_OK;
}

static int enqueue_body(Batch& b, PathMode mode, bool timed, double* d_finals, std::string& err) {
    HIP_TRY(launch_fill_u32(b.d_counters, kPathClasses, 0u, b.stream));
    if (mode == PathMode::Count) HIP_TRY(launch_zero_u64(b.d_counts, 4, b.stream));
    if (timed) HIP_TRY(hipEventRecord(b.ev[0], b.stream));
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) {
        if (!b.has_class[cls]) continue;
        PathArgs a = b.pa[cls];
        a.finals = d_finals;
        a.counts = mode == PathMode::Count ? b.d_counts : nullptr;
        HIP_TRY(launch_paths(class_precision(cls), b.ppt[cls], mode, b.grid[cls], b.block[cls], b.stream, a,
                             class_martingale(cls), class_geo(cls)));
        ++b.path_launches;
    }
    if (timed) HIP_TRY(hipEventRecord(b.ev[1], b.stream));
    HIP_TRY(launch_reduce(b.ra, b.n_group_threads, b.n_out_threads, b.stream));
    if (timed) HIP_TRY(hipEventRecord(b.ev[2], b.stream));
    HIP_TRY(hipMemcpyAsync(b.h_prices, b.ra.prices, sizeof(double) * b.n_out_total, hipMemcpyDeviceToHost, b.stream));
    HIP_TRY(hipMemcpyAsync(b.h_ses, b.ra.std_errs, sizeof(double) * b.n_out_total, hipMemcpyDeviceToHost, b.stream));
    return HEXO_OK;
}

int batch_enqueue(Batch& b, PathMode mode, bool timed, double* d_finals, std::string& err) {
    if (timed && !b.ev[0]) {
        err = "the batch was created without timing events";
        return HEXO_ERR_STATE;
    }
    b.timed = timed;
    b.t_run = std::chrono::steady_clock::now();
    return enqueue_body(b, mode, timed, d_finals, err);
}

int batch_capture_graph(Batch& b, std::string& err) {
    if (b.exec) return HEXO_OK;
    HIP_TRY(hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal));
    std::string e2;
    const uint64_t launches = b.path_launches;
    int rc = enqueue_body(b, PathMode::Normal, false, nullptr, e2);
    b.path_launches = launches;
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(b.stream, &g);
    if (rc != HEXO_OK) {
        if (g) (void)hipGraphDestroy(g);
        err = "graph capture: " + e2;
        return rc;
    }
    if (ee != hipSuccess) {
        err = std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ee);
        return HEXO_ERR_HIP;
    }
    b.graph = g;
    HIP_TRY(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    return HEXO_OK;
}

int batch_launch_graph(Batch& b, std::string& err) {
    if (!b.exec) {
        int rc = batch_capture_graph(b, err);
        if (rc != HEXO_OK) return rc;
    }
    b.timed = false;
    b.t_run = std::chrono::steady_clock::now();
    HIP_TRY(hipGraphLaunch(b.exec, b.stream));
    for (uint32_t cls = 0; cls < kPathClasses; ++cls) b.path_launches += b.has_class[cls] ? 1 : 0;
    return HEXO_OK;
}

int batch_finish(Batch& b, std::string& err) {
    HIP_TRY(hipStreamSynchronize(b.stream));
    b.last_run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b.t_run).count();
    if (b.timed) {
        float a = 0, c = 0;
        HIP_TRY(hipEventElapsedTime(&a, b.ev[0], b.ev[1]));
        HIP_TRY(hipEventElapsedTime(&c, b.ev[1], b.ev[2]));
        b.path_ms = a;
        b.reduce_ms = c;
    }
    return HEXO_OK;
}

void batch_copy_results(const Batch& b, double* prices, double* ses) {
    if (prices) std::memcpy(prices, b.h_prices, sizeof(double) * b.n_out_total);
    if (ses) std::memcpy(ses, b.h_ses, sizeof(double) * b.n_out_total);
}

Batch::~Batch() { batch_destroy(*this); }

void batch_destroy(Batch& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
    for (auto& e : b.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (b.d_mem && !b.d_external) {
        if (b.arena)
            b.arena->free(b.d_mem);
        else
            (void)hipFree(b.d_mem);
    }
    b.d_mem = nullptr;
    if (b.h_stage && !b.h_external) (void)hipHostFree(b.h_stage);
    b.h_stage = nullptr;
}

}  // namespace hexo
