// swift_host.cpp -- host side of the SWIFT European pricer: validation, the dynamic grid (the
// reference's SwiftParameters constructor, src/SWIFT.cpp:21-35) for hexo_swift_grid_for (a call's
// grids are chosen on the GPU, swift_grid.hpp), the layout of one call's chains in the context's
// pinned stages and device workspace, and the C entry points hexo_swift_grid_for /
// hexo_swift_price.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hexo.h"
#include "engine.hpp"
#include "swift.hpp"
#include "swift_grid.hpp"

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

// The reference's characteristic function (src/HDistribution.cpp:12-38), for the truncation test.
cd heston_chf(const double p[5], double tau, double u) {
    const double v0 = p[0], theta = p[1], rho = p[2], kappa = p[3], sigma = p[4];
    const cd iu(0.0, u);
    const cd xi = kappa + sigma * rho * iu;
    const cd fac = u * u - iu;
    const cd d = std::sqrt(xi * xi + sigma * sigma * fac);
    const cd sh = std::sinh(d * (tau * 0.5)), ch = std::cosh(d * (tau * 0.5));
    const cd A = fac * sh / (d / v0 * ch + xi / v0 * sh);
    const cd D = std::log(2.0 * d / (d + xi + (d - xi) * std::exp(-d * tau))) + (kappa - d) * (tau * 0.5);
    return std::exp(kappa * theta * rho * tau * iu / sigma - A + 2.0 * D * kappa * theta / (sigma * sigma));
}

// Truncation error of scale m (HDistribution::int_error, src/HDistribution.cpp:8-11).
double int_error(const double p[5], double tau, unsigned m) {
    const double e2m = std::exp2((double)m);
    return std::abs(heston_chf(p, tau, e2m * kPi) + heston_chf(p, tau, -e2m * kPi)) / (4.0 * e2m * kPi * kPi * tau);
}

bool finite_pos(double x) { return std::isfinite(x) && x > 0.0; }

// Everything but the strikes' values.
int check_chain(const hexo_swift_request& q, std::string& err) {
    if (!(finite_pos(q.v0) && finite_pos(q.theta) && finite_pos(q.kappa) && finite_pos(q.sigma) &&
          std::isfinite(q.rho) && q.rho > -1.0 && q.rho < 1.0)) {
        err = "Heston parameters out of range (v0, theta, kappa, sigma > 0, |rho| < 1)";
        return HEXO_ERR_INVALID;
    }
    if (!finite_pos(q.S0) || !finite_pos(q.T) || !std::isfinite(q.r)) {
        err = "S0 and T must be positive, r finite";
        return HEXO_ERR_INVALID;
    }
    if (q.n_strikes == 0 || !q.strikes) {
        err = "a chain needs at least one strike";
        return HEXO_ERR_INVALID;
    }
    if (q.flags & ~HEXO_SWIFT_FIXED_GRID) {
        err = "unknown flags";
        return HEXO_ERR_INVALID;
    }
    if (!(q.flags & HEXO_SWIFT_FIXED_GRID) && !(q.trunc > 0.0 && q.trunc < 1.0)) {
        err = "trunc must be in (0, 1)";
        return HEXO_ERR_INVALID;
    }
    return HEXO_OK;
}

int check_strikes(const hexo_swift_request& q, std::string& err) {
    for (uint32_t i = 0; i < q.n_strikes; ++i)
        if (!finite_pos(q.strikes[i])) {
            err = "strikes must be positive";
            return HEXO_ERR_INVALID;
        }
    return HEXO_OK;
}

int check_request(const hexo_swift_request& q, std::string& err) {
    const int rc = check_chain(q, err);
    return rc ? rc : check_strikes(q, err);
}

int check_grid(const hexo_swift_grid& g, std::string& err) {
    const uint32_t J = g.J;
    if (J == 0 || (J & (J - 1)) || J > HEXO_SWIFT_MAX_J) {
        err = "J must be a power of two <= HEXO_SWIFT_MAX_J";
        return HEXO_ERR_INVALID;
    }
    if (g.m > 30 || g.exp2_m != (1u << g.m) || g.reserved != 0 || !(g.sqrt_exp2_m > 0.0) ||
        !std::isfinite(g.lower) || !std::isfinite(g.upper) || !(g.lower < g.upper)) {
        err = "inconsistent SWIFT grid";
        return HEXO_ERR_INVALID;
    }
    // the density input of length 4J must hold every wavelet index once (no aliasing)
    if (g.k2 < g.k1 || (int64_t)g.k2 - (int64_t)g.k1 + 1 > 4 * (int64_t)J) {
        err = "wavelet index range must satisfy k1 <= k2 < k1 + 4J";
        return HEXO_ERR_INVALID;
    }
    return HEXO_OK;
}

// SwiftParameters constructor (src/SWIFT.cpp:21-35).
// The extreme strikes and the rest of a chain's grid-choice inputs.
hexo::SwiftGridIn grid_input(const hexo_swift_request& q) {
    hexo::SwiftGridIn in{{q.v0, q.theta, q.rho, q.kappa, q.sigma}, q.T, q.r, q.S0, q.strikes[0], q.strikes[0], q.trunc,
                         0, q.n_strikes, 0};
    for (uint32_t i = 1; i < q.n_strikes; ++i) {
        in.kmin = std::min(in.kmin, q.strikes[i]);
        in.kmax = std::max(in.kmax, q.strikes[i]);
    }
    return in;
}

int grid_status(int32_t status, double J_wanted, std::string& err) {
    switch (status) {
        case hexo::kSwiftGridOk: return HEXO_OK;
        case hexo::kSwiftGridNoScale: err = "no wavelet scale reaches the truncation precision"; break;
        case hexo::kSwiftGridOverflow: err = "SWIFT index range overflows"; break;
        case hexo::kSwiftGridBadStrike: err = "strikes must be positive"; break;
        default: err = "SWIFT grid needs J = " + std::to_string(J_wanted) + " (outside [1, HEXO_SWIFT_MAX_J])"; break;
    }
    return HEXO_ERR_INVALID;
}

// SwiftParameters constructor (src/SWIFT.cpp:21-35) on the host.
int dynamic_grid(const hexo_swift_request& q, hexo_swift_grid& g, std::string& err) {
    const hexo::SwiftGridIn in = grid_input(q);
    unsigned m = 0;
    // the reference loops `while (int_error(m++) > trunc)`: m ends one past the first scale
    // whose error bound is below trunc
    while (int_error(in.p, q.T, m++) > q.trunc)
        if (m > 30) return grid_status(hexo::kSwiftGridNoScale, 0.0, err);
    double J = 0.0;
    return grid_status(hexo::swift_grid_finish(in, m, g, J), J, err);
}

// Grows one of the context's SWIFT device buffers to at least `bytes` (kept for the context's
// lifetime, so calls of a calibration loop allocate nothing).
int grow_device(hexo_ctx* ctx, void*& buf, size_t& cap, size_t bytes) {
    if (bytes <= cap) return HEXO_OK;
    // the previous buffer may still be read by work queued on the stream
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "stream sync failed");
    if (buf) (void)hipFree(buf);
    buf = nullptr;
    cap = 0;
    const size_t want = bytes + bytes / 2;
    if (hipMalloc(&buf, want) != hipSuccess) {
        buf = nullptr;
        return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    }
    cap = want;
    return HEXO_OK;
}

int grid_of(const hexo_swift_request& q, hexo_swift_grid& g, std::string& err) {
    int rc = check_request(q, err);
    if (rc) return rc;
    if (q.flags & HEXO_SWIFT_FIXED_GRID)
        g = q.grid;
    else if ((rc = dynamic_grid(q, g, err)))
        return rc;
    return check_grid(g, err);
}

}  // namespace

extern "C" int hexo_swift_grid_for(const hexo_swift_request* req, hexo_swift_grid* out) {
    if (!req || !out) return HEXO_ERR_INVALID;
    std::string err;
    return grid_of(*req, *out, err);
}

extern "C" int hexo_swift_price(hexo_ctx* ctx, const hexo_swift_request* reqs, uint32_t n, double* prices, double* grads,
                                hexo_swift_grid* grids_used) {
    if (!ctx) return HEXO_ERR_INVALID;
    if (!reqs || !prices || n == 0) return ctx->set_error(HEXO_ERR_INVALID, "requests and prices must be non-NULL");
    std::string err;
    std::vector<uint64_t> strike_off(n);
    std::vector<uint32_t> dyn;  // chains whose grid is chosen here (not HEXO_SWIFT_FIXED_GRID)
    uint64_t n_strk = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (int rc = check_chain(reqs[i], err))
            return ctx->set_error(rc, "swift request " + std::to_string(i) + ": " + err);
        strike_off[i] = n_strk;
        n_strk += reqs[i].n_strikes;
        if (!(reqs[i].flags & HEXO_SWIFT_FIXED_GRID)) dyn.push_back(i);
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "hipSetDevice failed");
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    hipStream_t st = ctx->stream;
    // ---- first phase: the grid kernel reads the strikes and grid inputs from the fine-grained
    // pinned stage and writes the grids back into it (a wavefront per dynamic chain, which also
    // finds the chain's extreme strikes and checks them). Stage layout:
    // [strikes | grid inputs | grid outputs].
    const size_t nd = dyn.size();
    const size_t b_strk = align(8 * n_strk), b_gin = align(sizeof(hexo::SwiftGridIn) * nd),
                 b_gout = sizeof(hexo::SwiftGridOut) * nd;
    if (!ctx->swift_in_host.reserve(b_strk + b_gin + b_gout)) return ctx->set_error(HEXO_ERR_NOMEM, "hipHostMalloc failed");
    unsigned char* in_host = reinterpret_cast<unsigned char*>(ctx->swift_in_host.p);
    const double* h_strikes = reinterpret_cast<const double*>(in_host);
    {
        double* ks = reinterpret_cast<double*>(in_host);
        for (uint32_t i = 0; i < n; ++i) std::memcpy(ks + strike_off[i], reqs[i].strikes, 8 * reqs[i].n_strikes);
    }
    hexo::SwiftGridOut* gout = reinterpret_cast<hexo::SwiftGridOut*>(in_host + b_strk + b_gin);
    bool ok = true;
    if (nd) {
        hexo::SwiftGridIn* gin = reinterpret_cast<hexo::SwiftGridIn*>(in_host + b_strk);
        for (size_t k = 0; k < nd; ++k) {
            const hexo_swift_request& q = reqs[dyn[k]];
            gin[k] = hexo::SwiftGridIn{{q.v0, q.theta, q.rho, q.kappa, q.sigma}, q.T, q.r, q.S0, 0.0, 0.0, q.trunc,
                                       strike_off[dyn[k]], q.n_strikes, 0};
        }
        ok = hexo::launch_swift_grid(gin, h_strikes, gout, (uint32_t)nd, st) == hipSuccess;
    }
    auto fail = [&](int code, const std::string& msg) {
        // drain what was queued before the pinned stages are reused
        (void)hipStreamSynchronize(st);
        return ctx->set_error(code, msg);
    };
    if (!ok) return fail(HEXO_ERR_HIP, "SWIFT grid launch failed");
    // while the GPU chooses the grids: the strikes of the fixed-grid chains (the grid kernel checks
    // the others') and the records' grid-independent fields
    for (uint32_t i = 0; i < n; ++i)
        if (reqs[i].flags & HEXO_SWIFT_FIXED_GRID)
            if (int rc = check_strikes(reqs[i], err)) return fail(rc, "swift request " + std::to_string(i) + ": " + err);
    const uint32_t rows = grads ? 6u : 1u;
    std::vector<hexo::DevSwift> recs(n);
    for (uint32_t i = 0; i < n; ++i) {
        const hexo_swift_request& q = reqs[i];
        hexo::DevSwift& R = recs[i];
        std::memset(&R, 0, sizeof(R));
        R.p[0] = q.v0;
        R.p[1] = q.theta;
        R.p[2] = q.rho;
        R.p[3] = q.kappa;
        R.p[4] = q.sigma;
        R.S0 = q.S0;
        R.r = q.r;
        R.T = q.T;
        R.disc = std::exp(-q.r * q.T);
        R.n_strikes = q.n_strikes;
        R.rows = rows;
        R.strike_off = strike_off[i];
        R.out_off = strike_off[i];
    }
    if (nd && hipStreamSynchronize(st) != hipSuccess) return fail(HEXO_ERR_HIP, "SWIFT grid launch failed");
    std::vector<hexo_swift_grid> grids(n);
    {
        size_t k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (reqs[i].flags & HEXO_SWIFT_FIXED_GRID) {
                grids[i] = reqs[i].grid;
            } else {
                if (int rc = grid_status(gout[k].status, gout[k].J_wanted, err))
                    return fail(rc, "swift request " + std::to_string(i) + ": " + err);
                grids[i] = gout[k++].g;
            }
            if (int rc = check_grid(grids[i], err)) return fail(rc, "swift request " + std::to_string(i) + ": " + err);
        }
    }
    // ---- second phase: the grid-dependent fields, tables, FFT chain list and prefix sums, then
    // the work arrays
    std::vector<hexo::SwiftTable> tables;
    std::vector<uint32_t> fft_chains;
    std::vector<uint64_t> ends(4 * (size_t)n);  // j, dens, dcoef, strike blocks
    uint64_t n_pay = 0, n_dens = 0, n_pm = 0;
    hexo::SwiftTotals tot{0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < n; ++i) {
        const hexo_swift_grid& g = grids[i];
        hexo::DevSwift& R = recs[i];
        R.exp2_m = (double)g.exp2_m;
        R.sqrt_exp2_m = g.sqrt_exp2_m;
        R.lower = g.lower;
        R.upper = g.upper;
        R.k1 = g.k1;
        R.k2 = g.k2;
        R.J = g.J;
        R.log2_2J = 1u + (uint32_t)__builtin_ctz(g.J);
        // one twiddle table per distinct J
        size_t t = 0;
        while (t < tables.size() && tables[t].J != g.J) ++t;
        if (t == tables.size()) {
            tables.push_back(hexo::SwiftTable{g.J, 0, tot.tw});
            tot.tw += 4ull * g.J;
        }
        R.tw_off = tables[t].off;
        R.pay_off = n_pay;
        R.coeff_off = n_pay;
        R.dens_off = n_dens;
        R.pm_off = n_pm;
        const uint64_t nk = (uint64_t)((int64_t)g.k2 - (int64_t)g.k1 + 1);
        const bool fft = g.J <= hexo::kSwiftFftMaxJ;
        if (fft) {
            fft_chains.push_back(i);
            tot.fft_lds_bytes = std::max(tot.fft_lds_bytes, (uint32_t)(sizeof(hexo::Cplx) * 2 * g.J));
        } else {
            tot.dens += nk;
            tot.dcoef += g.J;
        }
        n_pay += g.J;
        n_dens += nk;
        n_pm += (uint64_t)rows * g.J;
        tot.j += g.J;
        tot.sblocks += (reqs[i].n_strikes + hexo::kSwiftStrikeBlock - 1) / hexo::kSwiftStrikeBlock;
        ends[i] = tot.j;
        ends[n + i] = tot.dens;
        ends[2 * (size_t)n + i] = tot.dcoef;
        ends[3 * (size_t)n + i] = tot.sblocks;
    }
    tot.n_fft = (uint32_t)fft_chains.size();
    std::vector<uint64_t> tab_end(tables.size());
    for (size_t t = 0; t < tables.size(); ++t) tab_end[t] = tables[t].off + 4ull * tables[t].J;
    const uint32_t n_tab = (uint32_t)tables.size();
    const size_t b_recs = align(sizeof(hexo::DevSwift) * n), b_tabs = align(sizeof(hexo::SwiftTable) * n_tab),
                 b_tabend = align(8 * (size_t)n_tab), b_fft = align(4 * fft_chains.size()),
                 b_ends = align(sizeof(uint64_t) * 4 * (size_t)n), b_tw = align(16 * tot.tw),
                 b_pay = align(16 * n_pay), b_dens = align(8 * n_dens), b_coeff = align(16 * n_pay),
                 b_pm = align(16 * n_pm), b_prices = align(8 * n_strk), b_grads = grads ? align(40 * n_strk) : 0;
    const size_t in_bytes = b_recs + b_tabs + b_tabend + b_fft + b_ends;
    // device: [tables | strikes | work arrays]; pinned stage: [tables | prices | grads]
    const size_t total = in_bytes + b_strk + b_tw + b_pay + b_dens + b_coeff + b_pm;
    if (!ctx->swift_host.reserve(in_bytes + b_prices + b_grads)) return fail(HEXO_ERR_NOMEM, "hipHostMalloc failed");
    if (int rc = grow_device(ctx, ctx->swift_dev, ctx->swift_dev_bytes, total)) return rc;
    unsigned char* host = reinterpret_cast<unsigned char*>(ctx->swift_host.p);
    std::memset(host, 0, in_bytes);
    size_t off = 0;
    std::memcpy(host + off, recs.data(), sizeof(hexo::DevSwift) * n);
    off += b_recs;
    std::memcpy(host + off, tables.data(), sizeof(hexo::SwiftTable) * n_tab);
    off += b_tabs;
    std::memcpy(host + off, tab_end.data(), 8 * (size_t)n_tab);
    off += b_tabend;
    if (!fft_chains.empty()) std::memcpy(host + off, fft_chains.data(), 4 * fft_chains.size());
    off += b_fft;
    std::memcpy(host + off, ends.data(), sizeof(uint64_t) * 4 * (size_t)n);
    unsigned char* dev = reinterpret_cast<unsigned char*>(ctx->swift_dev);
    hexo::SwiftArgs a{};
    off = 0;
    a.reqs = reinterpret_cast<const hexo::DevSwift*>(dev + off);
    a.n_req = n;
    off += b_recs;
    a.tables = reinterpret_cast<const hexo::SwiftTable*>(dev + off);
    a.n_tables = n_tab;
    off += b_tabs;
    a.tab_end = reinterpret_cast<const uint64_t*>(dev + off);
    off += b_tabend;
    a.fft_chains = reinterpret_cast<const uint32_t*>(dev + off);
    off += b_fft;
    const uint64_t* d_ends = reinterpret_cast<const uint64_t*>(dev + off);
    a.j_end = d_ends;
    a.dens_end = d_ends + n;
    a.dcoef_end = d_ends + 2 * (size_t)n;
    a.sblk_end = d_ends + 3 * (size_t)n;
    off += b_ends;
    double* d_strikes = reinterpret_cast<double*>(dev + off);
    a.strikes = d_strikes;
    off += b_strk;
    a.tw = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_tw;
    a.pay = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_pay;
    a.dens = reinterpret_cast<double*>(dev + off);
    off += b_dens;
    a.coeff = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_coeff;
    a.pm = reinterpret_cast<hexo::Cplx*>(dev + off);
    // the contraction writes the results straight into the pinned stage
    double* h_prices = reinterpret_cast<double*>(host + in_bytes);
    double* h_grads = reinterpret_cast<double*>(host + in_bytes + b_prices);
    a.prices = h_prices;
    a.grads = grads ? h_grads : nullptr;
    ok = hexo::launch_swift_stage(host, dev, in_bytes, h_strikes, d_strikes, b_strk, st) == hipSuccess &&
         hexo::launch_swift(a, tot, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    if (!ok) return fail(HEXO_ERR_HIP, "SWIFT launch failed");
    std::memcpy(prices, h_prices, 8 * n_strk);
    if (grads) std::memcpy(grads, h_grads, 40 * n_strk);
    if (grids_used) std::memcpy(grids_used, grids.data(), sizeof(hexo_swift_grid) * n);
    return HEXO_OK;
}
