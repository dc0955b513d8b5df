// swift.hpp -- device layout and launch wrappers of the SWIFT European pricer (FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cplx.hpp"

namespace hexo {

// Chains with J up to this size run both transforms as radix-2 FFTs of length 2J in one
// workgroup's LDS (16 B x 2J = 128 KiB); larger ones use the direct-sum kernels.
constexpr uint32_t kSwiftFftMaxJ = 4096;
// Strikes of one chain summed by one workgroup of the final contraction.
constexpr int kSwiftStrikeBlock = 16;

// One chain (expiry) of a SWIFT call: model, grid and the offsets of its slices in the call's
// device region. Built on the host by swift_host.cpp.
struct DevSwift {
    double p[5];  // v0, theta, rho, kappa, sigma
    double S0, r, T, disc;
    double exp2_m, sqrt_exp2_m, lower, upper;
    int32_t k1, k2;
    uint32_t J, n_strikes;
    uint32_t rows;  // 1 (prices) or 6 (prices and the five partials)
    uint32_t log2_2J;
    uint64_t strike_off;  // doubles
    uint64_t tw_off;      // Cplx, 4J: e^{2 pi i k / 4J}; shared by the chains with the same J
    uint64_t pay_off;     // Cplx, J: payoff coefficients
    uint64_t dens_off;    // doubles, k2 - k1 + 1
    uint64_t coeff_off;   // Cplx, J: density coefficients
    uint64_t pm_off;      // Cplx, rows x J: pricing matrix
    uint64_t out_off;     // doubles: prices; grads at 5x that offset in the grads array
};

// One twiddle table (per distinct J of the call).
struct SwiftTable {
    uint32_t J;
    uint32_t pad0;
    uint64_t off;  // Cplx
};

struct SwiftArgs {
    const DevSwift* reqs;
    uint32_t n_req;
    uint32_t n_tables;
    const SwiftTable* tables;
    const uint32_t* fft_chains;  // chains whose transforms run as LDS FFTs
    const double* strikes;
    Cplx* tw;
    Cplx* pay;
    double* dens;
    Cplx* coeff;
    Cplx* pm;
    double* prices;
    double* grads;  // null when no request asks for gradients
    // inclusive prefix sums of the per-kernel thread (or wavefront) counts
    const uint64_t* tab_end;     // over tables: 4J
    const uint64_t* j_end;       // over chains: J
    const uint64_t* dens_end;    // over chains: k2 - k1 + 1 for direct-sum chains, else 0
    const uint64_t* dcoef_end;   // over chains: J for direct-sum chains, else 0
    const uint64_t* sblk_end;    // over chains: groups of kSwiftStrikeBlock strikes (one workgroup each)
};

struct SwiftTotals {
    uint64_t tw, j, dens, dcoef, sblocks;
    uint32_t n_fft;         // chains in fft_chains
    uint32_t fft_lds_bytes; // 16 x 2 x (largest J among them)
};

hipError_t launch_swift(const SwiftArgs& a, const SwiftTotals& t, hipStream_t st);
// Copies bytes0 from s0 to d0 and bytes1 from s1 to d1 (multiples of 16; the sources may be
// fine-grained pinned host memory) in one kernel.
hipError_t launch_swift_stage(const void* s0, void* d0, size_t bytes0, const void* s1, void* d1, size_t bytes1,
                              hipStream_t st);

}  // namespace hexo
