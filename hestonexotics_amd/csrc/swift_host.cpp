// swift_host.cpp -- host side of the SWIFT European pricer: the dynamic grid (the reference's
// SwiftParameters constructor, src/SWIFT.cpp:21-35), validation, packing of one call's chains
// into one device region, and the C entry points hexo_swift_grid_for / hexo_swift_price.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hexo.h"
#include "engine.hpp"
#include "swift.hpp"

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;

// The reference's characteristic function (src/HDistribution.cpp:12-38), for the truncation test.
cd heston_chf(const double p[5], double tau, double u) {
    const double v0 = p[0], vm = p[1], rho = p[2], kappa = p[3], sigma = p[4];
    const cd iu(0.0, u);
    const cd xi = kappa + sigma * rho * iu;
    const cd fac = u * u - iu;
    const cd d = std::sqrt(xi * xi + sigma * sigma * fac);
    const cd sh = std::sinh(d * (tau * 0.5)), ch = std::cosh(d * (tau * 0.5));
    const cd A = fac * sh / (d / v0 * ch + xi / v0 * sh);
    const cd D = std::log(2.0 * d / (d + xi + (d - xi) * std::exp(-d * tau))) + (kappa - d) * (tau * 0.5);
    return std::exp(kappa * vm * rho * tau * iu / sigma - A + 2.0 * D * kappa * vm / (sigma * sigma));
}

// Truncation error of scale m (HDistribution::int_error, src/HDistribution.cpp:8-11).
double int_error(const double p[5], double tau, unsigned m) {
    const double e2m = std::exp2((double)m);
    return std::abs(heston_chf(p, tau, e2m * kPi) + heston_chf(p, tau, -e2m * kPi)) / (4.0 * e2m * kPi * kPi * tau);
}

// Cumulants that size the truncation range (HDistribution::first/second/fourth_order_moment,
// src/HDistribution.cpp:86-113). The reference binds p = (v0, theta, rho, kappa, sigma) to the
// names (v, s2, r, kappa, k) there; the formulas keep that binding.
double cumulant1(const double p[5], double t) { return -0.5 * p[1] * t; }
double cumulant2(const double p[5], double t) {
    const double s2 = p[1], r = p[2], a = p[3], k = p[4];
    const double a2 = a * a, a3 = a2 * a, k2 = k * k;
    return s2 / (8 * a3) *
           (-k2 * std::exp(-2 * a * t) + 4 * k * std::exp(-a * t) * (k - 2 * a * r) +
            2 * a * t * (4 * a2 + k2 - 4 * a * k * r) + k * (8 * a * r - 3 * k));
}
double cumulant4(const double p[5], double t) {
    const double s2 = p[1], r = p[2], a = p[3], k = p[4];
    const double a2 = a * a, a3 = a2 * a, a4 = a3 * a, k2 = k * k, k3 = k2 * k, k4 = k3 * k;
    const double t2 = t * t, r2 = r * r, q = k - 2 * a * r;
    const double e1 = std::exp(-a * t), e2 = std::exp(-2 * a * t), e3 = std::exp(-3 * a * t), e4 = std::exp(-4 * a * t);
    const double b4 = -3 * k4 * e4;
    const double b3 = -8 * k2 * e3 * (2 * a * k * t * q + 4 * a2 + k2 - 6 * a * k * r);
    const double b2 = -4 * e2 * (4 * a2 * k2 * t2 * q * q + 2 * a * k * t * (k3 - 16 * a3 * r - 12 * a * k2 * r + 4 * a2 * k * (3 + 4 * r2)) +
                                 8 * a4 - 3 * k4 - 32 * a3 * k * r + 8 * a * k3 * r + 16 * a2 * k2 * r2);
    const double b1 = -8 * e1 * (-2 * a2 * k * t2 * q * q * q -
                                 8 * a * t * (k4 - 7 * a * k3 * r + 4 * a4 * r2 - 8 * a3 * k * r * (1 + r2) + a2 * k2 * (3 + 14 * r2)) -
                                 9 * k4 + 70 * a * k3 * r + 32 * a3 * k * r * (4 + 3 * r2) - 16 * a4 * (1 + 4 * r2) -
                                 4 * a2 * k2 * (9 + 40 * r2));
    const double b0 = 4 * a * t * (5 * k4 - 40 * a * k3 * r - 32 * a3 * k * r * (3 + 2 * r2) + 16 * a4 * (1 + 4 * r2) +
                                   24 * a2 * k2 * (1 + 4 * r2)) -
                      73 * k4 + 544 * a * k3 * r + 128 * a3 * k * r * (7 + 6 * r2) - 32 * a4 * (3 + 16 * r2) -
                      64 * a2 * k2 * (4 + 19 * r2);
    return (3 * k2 * s2) / (64 * std::pow(a, 7)) * (b4 + b3 + b2 + b1 + b0);
}

bool finite_pos(double x) { return std::isfinite(x) && x > 0.0; }

int check_request(const hexo_swift_request& q, std::string& err) {
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
    for (uint32_t i = 0; i < q.n_strikes; ++i)
        if (!finite_pos(q.strikes[i])) {
            err = "strikes must be positive";
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
int dynamic_grid(const hexo_swift_request& q, hexo_swift_grid& g, std::string& err) {
    const double p[5] = {q.v0, q.theta, q.rho, q.kappa, q.sigma};
    unsigned m = 0;
    // the reference loops `while (int_error(m++) > trunc)`: m ends one past the first scale
    // whose error bound is below trunc
    while (int_error(p, q.T, m++) > q.trunc)
        if (m > 30) {
            err = "no wavelet scale reaches the truncation precision";
            return HEXO_ERR_INVALID;
        }
    double kmin = q.strikes[0], kmax = q.strikes[0];
    for (uint32_t i = 1; i < q.n_strikes; ++i) {
        kmin = std::min(kmin, q.strikes[i]);
        kmax = std::max(kmax, q.strikes[i]);
    }
    const double hi = q.r * q.T + std::log(q.S0 / kmin);
    const double lo = q.r * q.T + std::log(q.S0 / kmax);
    const double c = std::fabs(cumulant1(p, q.T)) +
                     10.0 * std::sqrt(std::fabs(cumulant2(p, q.T)) + std::sqrt(std::fabs(cumulant4(p, q.T))));
    g.m = m;
    g.exp2_m = 1u << m;
    g.sqrt_exp2_m = std::sqrt((double)g.exp2_m);
    g.lower = lo - c;
    g.upper = hi + c;
    const double k1 = std::ceil(g.exp2_m * g.lower), k2 = std::floor(g.exp2_m * g.upper);
    if (!(std::fabs(k1) < 2e9 && std::fabs(k2) < 2e9)) {
        err = "SWIFT index range overflows";
        return HEXO_ERR_INVALID;
    }
    g.k1 = (int32_t)k1;
    g.k2 = (int32_t)k2;
    const double iota = std::ceil(std::log2(kPi * std::fabs(k1 - k2))) - 1.0;
    const double J = std::exp2(iota - 1.0);
    if (!(J >= 1.0 && J <= (double)HEXO_SWIFT_MAX_J)) {
        err = "SWIFT grid needs J = " + std::to_string(J) + " (outside [1, HEXO_SWIFT_MAX_J])";
        return HEXO_ERR_INVALID;
    }
    g.J = (uint32_t)J;
    g.reserved = 0;
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
    std::vector<hexo_swift_grid> grids(n);
    for (uint32_t i = 0; i < n; ++i) {
        const int rc = grid_of(reqs[i], grids[i], err);
        if (rc) return ctx->set_error(rc, "swift request " + std::to_string(i) + ": " + err);
    }
    // host image of the device region: records, strikes, prefix sums; then the work arrays
    std::vector<hexo::DevSwift> recs(n);
    std::vector<uint64_t> ends(4 * (size_t)n);
    uint64_t n_strk = 0, n_tw = 0, n_pay = 0, n_dens = 0, n_pm = 0;
    hexo::SwiftTotals tot{0, 0, 0, 0};
    const uint32_t rows = grads ? 6u : 1u;
    for (uint32_t i = 0; i < n; ++i) {
        const hexo_swift_request& q = reqs[i];
        const hexo_swift_grid& g = grids[i];
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
        R.exp2_m = (double)g.exp2_m;
        R.sqrt_exp2_m = g.sqrt_exp2_m;
        R.lower = g.lower;
        R.upper = g.upper;
        R.k1 = g.k1;
        R.k2 = g.k2;
        R.J = g.J;
        R.n_strikes = q.n_strikes;
        R.rows = rows;
        R.strike_off = n_strk;
        R.out_off = n_strk;
        R.tw_off = n_tw;
        R.pay_off = n_pay;
        R.coeff_off = n_pay;
        R.dens_off = n_dens;
        R.pm_off = n_pm;
        const uint64_t nk = (uint64_t)((int64_t)g.k2 - (int64_t)g.k1 + 1);
        n_strk += q.n_strikes;
        n_tw += 4ull * g.J;
        n_pay += g.J;
        n_dens += nk;
        n_pm += (uint64_t)rows * g.J;
        tot.tw += 4ull * g.J;
        tot.dens += nk;
        tot.j += g.J;
        tot.strikes += q.n_strikes;
        ends[i] = tot.tw;
        ends[n + i] = tot.dens;
        ends[2 * (size_t)n + i] = tot.j;
        ends[3 * (size_t)n + i] = tot.strikes;
    }
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_recs = align(sizeof(hexo::DevSwift) * n), b_ends = align(sizeof(uint64_t) * 4 * (size_t)n),
                 b_strk = align(8 * n_strk), b_tw = align(16 * n_tw), b_pay = align(16 * n_pay),
                 b_dens = align(8 * n_dens), b_coeff = align(16 * n_pay), b_pm = align(16 * n_pm),
                 b_prices = align(8 * n_strk), b_grads = grads ? align(40 * n_strk) : 0;
    const size_t in_bytes = b_recs + b_ends + b_strk;
    const size_t total = in_bytes + b_tw + b_pay + b_dens + b_coeff + b_pm + b_prices + b_grads;
    std::vector<unsigned char> host(in_bytes, 0);
    std::memcpy(host.data(), recs.data(), sizeof(hexo::DevSwift) * n);
    std::memcpy(host.data() + b_recs, ends.data(), sizeof(uint64_t) * 4 * (size_t)n);
    {
        double* ks = reinterpret_cast<double*>(host.data() + b_recs + b_ends);
        for (uint32_t i = 0; i < n; ++i) std::memcpy(ks + recs[i].strike_off, reqs[i].strikes, 8 * reqs[i].n_strikes);
    }
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->set_error(HEXO_ERR_HIP, "hipSetDevice failed");
    unsigned char* dev = nullptr;
    if (hipMalloc(&dev, total) != hipSuccess) return ctx->set_error(HEXO_ERR_NOMEM, "hipMalloc failed");
    hexo::SwiftArgs a{};
    a.reqs = reinterpret_cast<const hexo::DevSwift*>(dev);
    a.n_req = n;
    const uint64_t* d_ends = reinterpret_cast<const uint64_t*>(dev + b_recs);
    a.tw_end = d_ends;
    a.dens_end = d_ends + n;
    a.j_end = d_ends + 2 * (size_t)n;
    a.strike_end = d_ends + 3 * (size_t)n;
    a.strikes = reinterpret_cast<const double*>(dev + b_recs + b_ends);
    size_t off = in_bytes;
    a.tw = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_tw;
    a.pay = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_pay;
    a.dens = reinterpret_cast<double*>(dev + off);
    off += b_dens;
    a.coeff = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_coeff;
    a.pm = reinterpret_cast<hexo::Cplx*>(dev + off);
    off += b_pm;
    a.prices = reinterpret_cast<double*>(dev + off);
    off += b_prices;
    a.grads = grads ? reinterpret_cast<double*>(dev + off) : nullptr;
    hipStream_t st = ctx->stream;
    bool ok = hipMemcpyAsync(dev, host.data(), in_bytes, hipMemcpyHostToDevice, st) == hipSuccess &&
              hexo::launch_swift(a, tot, st) == hipSuccess &&
              hipMemcpyAsync(prices, a.prices, 8 * n_strk, hipMemcpyDeviceToHost, st) == hipSuccess &&
              (!grads || hipMemcpyAsync(grads, a.grads, 40 * n_strk, hipMemcpyDeviceToHost, st) == hipSuccess) &&
              hipStreamSynchronize(st) == hipSuccess;
    (void)hipFree(dev);
    if (!ok) return ctx->set_error(HEXO_ERR_HIP, "SWIFT launch failed");
    if (grids_used) std::memcpy(grids_used, grids.data(), sizeof(hexo_swift_grid) * n);
    return HEXO_OK;
}
