// swift_grid.hpp -- the dynamic SWIFT grid (the reference's SwiftParameters constructor,
// src/SWIFT.cpp:21-35), shared by the host (hexo_swift_grid_for) and the GPU grid kernel that
// hexo_swift_price runs for all of a call's chains at once.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/hexo.h"

namespace hexo {

// One chain's inputs to the grid choice: model, expiry, rate, spot, the extreme strikes and the
// truncation precision. The GPU grid kernel finds the extreme strikes itself, from the chain's
// n_strikes strikes at strike_off of the call's strike array (checking them on the way).
struct SwiftGridIn {
    double p[5];  // v0, theta, rho, kappa, sigma
    double T, r, S0, kmin, kmax, trunc;
    uint64_t strike_off;
    uint32_t n_strikes;
    uint32_t pad;
};

// One chain's grid and its status: kSwiftGridOk, or which check of the choice failed.
struct SwiftGridOut {
    hexo_swift_grid g;
    int32_t status;
    int32_t pad;
    double J_wanted;  // the Fourier size the index range asks for (kSwiftGridJRange)
};
constexpr int32_t kSwiftGridOk = 0;
constexpr int32_t kSwiftGridNoScale = 1;   // no scale m <= 30 reaches the truncation precision
constexpr int32_t kSwiftGridOverflow = 2;  // the wavelet index range overflows 32 bits
constexpr int32_t kSwiftGridJRange = 3;    // J outside [1, HEXO_SWIFT_MAX_J]
constexpr int32_t kSwiftGridBadStrike = 4; // a strike that is not finite and positive

// Cumulants that size the truncation range (HDistribution::first/second/fourth_order_moment,
// src/HDistribution.cpp:86-113), in the model's parameter names: p = (v0, theta, rho, kappa, sigma).
__host__ __device__ inline double swift_cumulant1(const double p[5], double t) { return -0.5 * p[1] * t; }
__host__ __device__ inline double swift_cumulant2(const double p[5], double t) {
    const double theta = p[1], rho = p[2], kappa = p[3], sigma = p[4];
    const double kappa2 = kappa * kappa, kappa3 = kappa2 * kappa, sigma2 = sigma * sigma;
    return theta / (8 * kappa3) *
           (-sigma2 * ::exp(-2 * kappa * t) + 4 * sigma * ::exp(-kappa * t) * (sigma - 2 * kappa * rho) +
            2 * kappa * t * (4 * kappa2 + sigma2 - 4 * kappa * sigma * rho) + sigma * (8 * kappa * rho - 3 * sigma));
}
__host__ __device__ inline double swift_cumulant4(const double p[5], double t) {
    const double theta = p[1], rho = p[2], kappa = p[3], sigma = p[4];
    const double kappa2 = kappa * kappa, kappa3 = kappa2 * kappa, kappa4 = kappa3 * kappa;
    const double sigma2 = sigma * sigma, sigma3 = sigma2 * sigma, sigma4 = sigma3 * sigma;
    const double t2 = t * t, rho2 = rho * rho, q = sigma - 2 * kappa * rho;
    const double e1 = ::exp(-kappa * t), e2 = ::exp(-2 * kappa * t), e3 = ::exp(-3 * kappa * t),
                 e4 = ::exp(-4 * kappa * t);
    const double b4 = -3 * sigma4 * e4;
    const double b3 = -8 * sigma2 * e3 * (2 * kappa * sigma * t * q + 4 * kappa2 + sigma2 - 6 * kappa * sigma * rho);
    const double b2 =
        -4 * e2 *
        (4 * kappa2 * sigma2 * t2 * q * q +
         2 * kappa * sigma * t *
             (sigma3 - 16 * kappa3 * rho - 12 * kappa * sigma2 * rho + 4 * kappa2 * sigma * (3 + 4 * rho2)) +
         8 * kappa4 - 3 * sigma4 - 32 * kappa3 * sigma * rho + 8 * kappa * sigma3 * rho + 16 * kappa2 * sigma2 * rho2);
    const double b1 = -8 * e1 *
                      (-2 * kappa2 * sigma * t2 * q * q * q -
                       8 * kappa * t *
                           (sigma4 - 7 * kappa * sigma3 * rho + 4 * kappa4 * rho2 -
                            8 * kappa3 * sigma * rho * (1 + rho2) + kappa2 * sigma2 * (3 + 14 * rho2)) -
                       9 * sigma4 + 70 * kappa * sigma3 * rho + 32 * kappa3 * sigma * rho * (4 + 3 * rho2) -
                       16 * kappa4 * (1 + 4 * rho2) - 4 * kappa2 * sigma2 * (9 + 40 * rho2));
    const double b0 = 4 * kappa * t *
                          (5 * sigma4 - 40 * kappa * sigma3 * rho - 32 * kappa3 * sigma * rho * (3 + 2 * rho2) +
                           16 * kappa4 * (1 + 4 * rho2) + 24 * kappa2 * sigma2 * (1 + 4 * rho2)) -
                      73 * sigma4 + 544 * kappa * sigma3 * rho + 128 * kappa3 * sigma * rho * (7 + 6 * rho2) -
                      32 * kappa4 * (3 + 16 * rho2) - 64 * kappa2 * sigma2 * (4 + 19 * rho2);
    return (3 * sigma2 * theta) / (64 * ::pow(kappa, 7)) * (b4 + b3 + b2 + b1 + b0);
}

// The rest of the grid once the scale m is known (src/SWIFT.cpp:24-34): the truncation range from
// the strikes and the cumulants, the wavelet index range k1..k2 and J = 2^(ceil(log2(pi |k2 - k1|)) - 2).
__host__ __device__ inline int32_t swift_grid_finish(const SwiftGridIn& q, uint32_t m, hexo_swift_grid& g,
                                                     double& J_wanted) {
    const double pi = 3.14159265358979323846;
    const double hi = q.r * q.T + ::log(q.S0 / q.kmin);
    const double lo = q.r * q.T + ::log(q.S0 / q.kmax);
    const double c = ::fabs(swift_cumulant1(q.p, q.T)) +
                     10.0 * ::sqrt(::fabs(swift_cumulant2(q.p, q.T)) + ::sqrt(::fabs(swift_cumulant4(q.p, q.T))));
    g.m = m;
    g.exp2_m = 1u << m;
    g.sqrt_exp2_m = ::sqrt((double)g.exp2_m);
    g.lower = lo - c;
    g.upper = hi + c;
    const double k1 = ::ceil(g.exp2_m * g.lower), k2 = ::floor(g.exp2_m * g.upper);
    if (!(::fabs(k1) < 2e9 && ::fabs(k2) < 2e9)) return kSwiftGridOverflow;
    g.k1 = (int32_t)k1;
    g.k2 = (int32_t)k2;
    const double iota = ::ceil(::log2(pi * ::fabs(k1 - k2))) - 1.0;
    const double J = ::exp2(iota - 1.0);
    J_wanted = J;
    if (!(J >= 1.0 && J <= (double)HEXO_SWIFT_MAX_J)) return kSwiftGridJRange;
    g.J = (uint32_t)J;
    g.reserved = 0;
    return kSwiftGridOk;
}

// The grid of chains in[0..n) on the GPU (one wavefront per chain; the scales m = 0..30 are tried
// in parallel and the first that reaches the truncation precision is taken, as the reference's
// loop takes it), with each chain's extreme strikes from `strikes`.
hipError_t launch_swift_grid(const SwiftGridIn* in, const double* strikes, SwiftGridOut* out, uint32_t n,
                             hipStream_t st);

}  // namespace hexo
