// layout.hpp -- HBM layout shared by the host engine and the path kernels.
//
// A batch is uploaded as one contiguous device arena:
//   DevReq[n_req]        one record per pricing request (parameters, RNG key, table offsets)
//   DevSeg[n_seg]        the request's time grid as runs of equal steps (host-built, path independent)
//   DevFin[n_fin]        expiry finalisations at the end of a run
//   DevOpt[n_opt]        strikes (relative and absolute) and discount factor per option
//   uint32 order[n_req]  requests sorted longest-first; item_end[n_req] their work-item prefix sums
//   partials             16 bytes per (stored tile sum, output): FP64 (sum, sum of squares) or
//                        compensated FP32 (sum, carry, sum of squares, carry). A path-kernel wave
//                        that holds several tiles of one 4-tile chunk combines them first (below)
//   groups, results      per (64-tile group, output) partials; per output (price, std error)
#pragma once
#include <stdint.h>

namespace hexo {

constexpr int kTile = 64;         // paths per tile = one wavefront
constexpr int kGroupTiles = 64;   // tiles per first-level reduction group
// Tile sums are combined in chunks of kChunkTiles consecutive tiles by one fixed pairwise tree,
// ((t0 + t1) + (t2 + t3)), with absent tiles (beyond the request's last) taken as zero sums. A
// path-kernel wave holding s = 1, 2 or 4 tiles of a chunk forms the tree's level for s tiles and
// stores that (the request's store_tiles); the group reduction completes the tree. The sums are
// therefore the same, bit for bit, whatever the paths per thread of the launch.
constexpr int kChunkTiles = 4;
constexpr int kMaxPPT = 4;        // paths per thread (FP32 kernel)
constexpr unsigned long long kMaxPaths = (1ull << 37) - kTile;  // per request: 2^31 - 1 tiles

// One run of steps with a constant step width (the reference's dt = T_c / steps between two
// expiry finalisations). FP32 constants are computed on the host in FP64 (per request, per run):
//   m  = V + omd (th - V)            (conditional mean, src/HSimulation.tpp:58-59; omd = 1 - e^-kappa dt)
//   s2 = 2 (V c1h + c2h)             (conditional variance, src/HSimulation.tpp:60; halved constants)
//   y += g0 + ka (V' - m) + kb (th - V) + sqrt(k3 (V + V')) Z              (log2 of X/S0)
// with ka = K2, kb = -(K1 + K2 e^-kappa dt), k3 = K3 = K4 and g0 = K0 + (K1 + K2) th
// (src/HSimulation.tpp:75-80: K1 V + K2 V' = (K1 + K2) th + K2 (V' - m) - (K1 + K2 e^-kappa dt)(th - V)),
// all scaled by log2(e) (k3 by log2(e)^2); th is theta rounded to FP32 and g0 uses that same
// value. This is the reference's step re-associated so that rounding the constants to FP32 moves
// neither fixed point: in the form V e^-kappa dt + theta(1 - e^-kappa dt) the rounded e^-kappa dt
// shifts the mean-reversion level by O(ulp / (kappa dt)), and K0 and K1 V + K2 V' nearly cancel,
// so their separate roundings add a drift of O(ulp K0) per step. kb is O(dt) although K1 and K2
// are O(1/sigma): it is formed in FP64 and rounded once, so no FP32 product of size 1/sigma
// cancels in the step.
//
// With the martingale correction the log-price step is written without K0 and K1:
//   y += ma (V' - m) - mh (V + V') + (A m - ln E[exp(A V')]) log2(e) + mr + sqrt(k3 (V + V')) Z
// with A = K2 + K4/2 (mA), ma = A log2(e), mh = K3/2 log2(e), mr = r dt log2(e). E[V'] = m holds
// by construction, so the first term has mean zero whatever the rounding of ma, and
// A m - ln E[exp(A V')] is O((A a)^2), evaluated in forms without cancellation.
struct DevSeg {
    float omd, th, c1h, c2h;
    float g0, ka, kb, k3;
    float mA, ma, mh, mr;  // martingale-corrected step (HEXO_MARTINGALE)
    float delta_f;  // Asian weight of accumulations inside the run
    float dnext_f;  // weight of the accumulation after the run's last step (new dt), 0 if none
    uint32_t n_steps;
    uint32_t fin_begin, fin_count;
    double delta;  // FP64 kernel: step width
    double dnext;  // FP64 kernel: dt after the finalisations
};

// Finalisation of one expiry (AAsianCallNonAdaptive / EuropeanCallNonAdaptive
// accumulate_final_value, src/inc/AsianContract.h:29-34, src/inc/VanillaContract.h:28-31).
struct DevFin {
    uint32_t chain, opt_begin, n_strikes;
    float gInvW_f;  // geometric control variate: 1 / W
    float w_f;     // (T - t_prev) / dt
    float invT_f;  // 1 / T
    float hwdt_f;  // (T - t_prev) / 2, corrected-trapezoid extension
    float gW_f;    // geometric control variate: W, the sum of the Asian functional's weights
    double T, tp, dt;  // FP64 kernel: reference-form interpolation (X - Xp)(T - tp)/dt
    double gW;         // FP64 kernel: W
};

struct DevOpt {
    float k_rel;  // strike / S0 (FP32 kernel works in units of S0)
    float pad0;
    double k_abs;
    double disc;  // exp(-r T) with HEXO_DISCOUNT, else 1
    double pad1;
};

struct DevReq {
    double v0, theta, rho, kappa, sigma, S0, r_drift, log_S0;
    uint64_t n_paths, path_offset, partial_base, group_base;
    uint32_t seed_lo, seed_hi;
    uint32_t seg_begin, n_segs;
    uint32_t n_opts, n_pay, payoffs, flags;
    uint32_t precision, n_tiles, opt_begin, out_base;
    uint32_t n_exp, n_groups;
    float v0_f;
    uint32_t store_tiles;  // tiles per stored partial: 1, 2 or kChunkTiles (the kernel's paths per thread)
};

struct PathArgs {
    const DevReq* reqs;
    const DevSeg* segs;
    const DevFin* fins;
    const DevOpt* opts;
    const uint32_t* order;     // sorted position -> request index
    const uint32_t* item_end;  // sorted position -> exclusive prefix end of work items
    uint32_t n_sorted;
    uint32_t total_items;
    uint32_t* counter;  // work-queue head, zeroed before each launch
    void* partials;
    double* finals;                 // optional per-path (A, X_T) per expiry
    unsigned long long* counts;     // optional branch counters [4]
};

// Geometric-average control variate (payoff HEXO_ASIAN_GEO_CV). The Asian functional of an
// expiry, A = sum_j w_j X_{t_j} over the grid points, has weight sum W; its geometric counterpart
// is G = W S0 exp(sum_j v_j ln(X_{t_j}/S0)) with v = w / W, whose exact price under the Heston
// model is computed once per batch: one wavefront per (request, expiry, block of <= 64 strikes).
struct DevGeo {
    uint32_t req;         // request index
    uint32_t pts_begin;   // first point (t_0 = 0) of the expiry's (t_j, v_j) list
    uint32_t n_pts;       // intervals: points 0..n_pts
    uint32_t n_strikes;   // strikes of this block (<= 64)
    uint32_t opt;         // first option of the block (batch-global DevOpt index)
    uint32_t max_panels;  // bound on 64-node panels (the price is NaN if not converged by then)
    uint32_t wave_begin;  // first wavefront of the item in the parallel panel pass
    uint32_t p0;          // panels evaluated in parallel (that pass has p0 + 1 waves: the last one E[e^L])
    uint64_t out;         // geo[] index of the block's first control-variate output
    double c;             // W S0
    double h;             // width in u of one panel
};

struct GeoArgs {
    const DevReq* reqs;
    const DevOpt* opts;
    const DevGeo* items;
    const double* pts;  // (t_j, v_j) pairs
    const double* gl;   // 64 Gauss-Legendre nodes on [-1, 1], then their 64 weights
    double* geo;        // exact geometric prices (undiscounted, absolute units), per output
    uint32_t n_items;
    uint32_t n_waves;       // waves of the parallel panel pass (sum over items of p0 + 1)
    const uint32_t* wave_end;  // inclusive prefix over items of p0 + 1
    double* panel;          // per panel wave: each strike's panel sum (64 doubles)
    double* panel_max;      // per panel wave: max |phi| over its nodes; E[e^L] in the item's last wave
    double* state;          // per item between rounds: strike sums, E[e^L], panels done, converged
};

struct ReduceArgs {
    const DevReq* reqs;
    const DevOpt* opts;
    const double* geo;  // exact geometric prices added to the control-variate outputs (or null)
    const uint32_t* req_of;       // request index for sorted reduction slices
    const uint64_t* group_end;    // prefix over requests of n_groups * n_out (pass A threads)
    const uint64_t* out_end;      // prefix over requests of n_out (pass B threads)
    uint32_t n_req;
    const void* partials;
    void* groups;
    double* prices;
    double* std_errs;
};

}  // namespace hexo
