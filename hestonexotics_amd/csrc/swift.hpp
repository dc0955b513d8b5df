// swift.hpp -- device layout and launch wrappers of the SWIFT European pricer (FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hexo {

struct Cplx {
    double re, im;
};

// One chain (expiry) of a SWIFT call: model, grid and the offsets of its slices in the call's
// device region. Built on the host by swift_host.cpp.
struct DevSwift {
    double p[5];  // v0, theta, rho, kappa, sigma
    double S0, r, T, disc;
    double exp2_m, sqrt_exp2_m, lower, upper;
    int32_t k1, k2;
    uint32_t J, n_strikes;
    uint32_t rows;  // 1 (prices) or 6 (prices and the five partials)
    uint32_t pad0;
    uint64_t strike_off;  // doubles
    uint64_t tw_off;      // Cplx, 4J: e^{2 pi i k / 4J}
    uint64_t pay_off;     // Cplx, J: payoff coefficients
    uint64_t dens_off;    // doubles, k2 - k1 + 1
    uint64_t coeff_off;   // Cplx, J: density coefficients
    uint64_t pm_off;      // Cplx, rows x J: pricing matrix
    uint64_t out_off;     // doubles: prices; grads at 5x that offset in the grads array
};

struct SwiftArgs {
    const DevSwift* reqs;
    uint32_t n_req;
    const double* strikes;
    Cplx* tw;
    Cplx* pay;
    double* dens;
    Cplx* coeff;
    Cplx* pm;
    double* prices;
    double* grads;  // null when no request asks for gradients
    // inclusive prefix sums over requests of the per-kernel thread (or wave) counts
    const uint64_t* tw_end;      // 4J
    const uint64_t* dens_end;    // k2 - k1 + 1
    const uint64_t* j_end;       // J
    const uint64_t* strike_end;  // n_strikes (one wavefront each)
};

// Totals of the four index spaces (the last entries of the prefix sums).
struct SwiftTotals {
    uint64_t tw, dens, j, strikes;
};

hipError_t launch_swift(const SwiftArgs& a, const SwiftTotals& t, hipStream_t st);

}  // namespace hexo
