/*
 * hexo.h -- C ABI of the MI355X-native Heston QE Monte-Carlo pricer.
 *
 * This is the drop-in boundary for the reference's pricing path
 *   template<class Scheme> std::vector<ffloat> HSimulation::price(const HParams&, const ffloat S,
 *       const std::list<options_chain>&, unsigned n_simulations, unsigned n_opts, unsigned steps)
 * (src/inc/HSimulation.h:61-63, src/HSimulation.tpp:10-51), with
 *   Scheme = HSimulation::HQEAnderson<ffloat, AAsianCallNonAdaptive | EuropeanCallNonAdaptive>.
 * include/hexo/HSimulation.h wraps these entry points in the reference's template spelling.
 *
 * Conventions: plain C types only; every entry point returns HEXO_OK (0) or a negative
 * hexo_status; no exception crosses the ABI; hexo_ctx_last_error() holds the message of the
 * last failure on a context. One context owns one GPU. Calls on one context are serialised
 * internally. The caller owns every host array passed in; the library copies what it keeps.
 */
#ifndef HEXO_H
#define HEXO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEXO_ABI_VERSION 1

typedef enum hexo_status {
    HEXO_OK = 0,
    HEXO_ERR_INVALID = -1,     /* malformed request or argument */
    HEXO_ERR_NO_DEVICE = -2,   /* no usable HIP device */
    HEXO_ERR_HIP = -3,         /* a HIP runtime call failed */
    HEXO_ERR_NOMEM = -4,       /* device or host allocation failed */
    HEXO_ERR_ADMISSION = -5,   /* request exceeds the serving memory budget */
    HEXO_ERR_NOT_FOUND = -6,   /* unknown ticket */
    HEXO_ERR_CANCELLED = -7,   /* ticket cancelled before it ran */
    HEXO_ERR_TIMEOUT = -8,     /* wait timed out (the ticket stays valid) */
    HEXO_ERR_STATE = -9,       /* call not valid in the current state (e.g. server not running) */
    HEXO_ERR_INJECTED = -10    /* failure forced by hexo_inject_fault (tests of failure handling) */
} hexo_status;

/* payoffs: which option policies are priced from the same paths */
#define HEXO_ASIAN 1u    /* AAsianCallNonAdaptive   (src/inc/AsianContract.h:14-48) */
#define HEXO_EUROPEAN 2u /* EuropeanCallNonAdaptive (src/inc/VanillaContract.h:14-40) */
/* AAsianCallNonAdaptive priced with the geometric-average control variate the reference suggests
 * (src/inc/HSimulation.h:48-60): per path max(A - K, 0) - max(G - K, 0), where G applies the same
 * Asian functional to ln(X/S0) (normalised by its weight sum W: G = W S0 exp(sum_j w_j ln(X_j/S0) / W)),
 * plus the exact price of the G call under the Heston model. Same expectation as the plain Asian
 * estimator up to the scheme's discretisation of G; far smaller standard error. Needs every
 * expiry's weight sum W >= 0.25 (at least 2 steps). Outputs are payoff-major in bit order. */
#define HEXO_ASIAN_GEO_CV 4u
/* The geometric-average call of the same functional, max(G - K, 0), by plain Monte Carlo. With
 * HEXO_ASIAN and HEXO_ASIAN_GEO_CV it gives the variances and covariance for the control variate
 * with the regression coefficient instead of 1 (flag HEXO_CV_REGRESSION). */
#define HEXO_GEO_ASIAN 8u

/* flags: documented extensions; 0 reproduces the reference exactly */
#define HEXO_DRIFT 1u     /* add r*delta to the log-price step (the reference has no drift term) */
#define HEXO_DISCOUNT 2u  /* multiply each price by exp(-r T) (the reference does not discount) */
#define HEXO_TRAPEZOID 4u /* corrected trapezoid for the Asian average's last interval */
/* Andersen's martingale correction (variance reduction, suggested beside src/inc/HSimulation.h:48-60):
 * K0 of each step is replaced by K0* = -ln E[exp(A V') | V] - (K1 + K3/2) V with A = K2 + K4/2,
 * so that E[X' | X, V] = X exp(r delta) exactly (the no-drift reference form: E[X'] = X). Where
 * the expectation does not exist (A >= 1/(2a) in the quadratic branch, A >= beta in the
 * exponential one) that step keeps the reference's K0. */
#define HEXO_MARTINGALE 8u
/* The control variate with the regression coefficient instead of 1: the HEXO_ASIAN_GEO_CV block
 * holds A - beta (G - E[G]) with beta = Cov(A, G) / Var G estimated from the same paths, and its
 * standard error sqrt((Var A - beta Cov(A, G)) / n), never above the plain Asian one. Needs
 * HEXO_ASIAN, HEXO_ASIAN_GEO_CV and HEXO_GEO_ASIAN in payoffs (the covariance comes from the
 * variances of A, G and A - G); formed in the final reduction on the GPU. */
#define HEXO_CV_REGRESSION 16u

/* precision of the path kernel. The FP32 modes need sigma^2 dt >= 1e-18 max(v0, theta) for every
 * step width dt (below it the variance step's moment ratio would overflow single precision;
 * such requests are refused with HEXO_ERR_INVALID and price in HEXO_FP64). */
#define HEXO_FP64 0u             /* plain FP64 kernel: reference form, FP64 everywhere */
#define HEXO_FP32 1u             /* FP32 stepping, FP64 cross-path reduction and estimator */
#define HEXO_FP32_COMPENSATED 2u /* FP32 stepping, compensated-FP32 reduction, FP64 estimator */

/*
 * One pricing request = one call of the reference's price<Scheme> for one underlying.
 * Heston parameters are in HParams order (src/inc/HDistribution.h:9-14).
 * Chains: n_expiries expiries (strictly increasing, > 0), expiry c has n_strikes[c] strikes,
 * strikes concatenated in chain order (the reference's chain-major option order,
 * src/HSimulation.tpp:40). `steps` is the reference's steps per expiry: expiry c is stepped
 * with dt = T_c / steps (src/inc/AsianContract.h:35-38).
 * Path p (0 <= p < n_paths) uses Philox4x32-10 counter block {i, lo32(path_offset+p),
 * hi32(path_offset+p), 0} for step pair i, key {lo32(seed), hi32(seed)}.
 */
typedef struct hexo_request {
    double v0, theta, rho, kappa, sigma; /* HParams {v_0, v_m, rho, kappa, sigma} */
    double S0;                           /* spot */
    double r;                            /* rate: used only with HEXO_DRIFT / HEXO_DISCOUNT */
    uint64_t n_paths;
    uint64_t seed;
    uint64_t path_offset; /* first global path index (sharding, resumable chunks); a multiple of 64 */
    uint32_t steps;
    uint32_t n_expiries;
    const double* expiries;
    const uint32_t* n_strikes;
    const double* strikes;
    uint32_t payoffs;   /* HEXO_ASIAN | HEXO_EUROPEAN | HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN */
    uint32_t flags;     /* HEXO_DRIFT | HEXO_DISCOUNT | HEXO_TRAPEZOID | HEXO_MARTINGALE | HEXO_CV_REGRESSION */
    uint32_t precision; /* HEXO_FP64 | HEXO_FP32 | HEXO_FP32_COMPENSATED */
    uint32_t reserved;  /* must be 0 */
} hexo_request;

/* Launch shape of the FP32 path kernel. Results do not depend on it: every 64-path tile is
 * reduced on its own and tiles are combined in a fixed order. */
typedef struct hexo_tuning {
    uint32_t paths_per_thread; /* 1, 2, 3 or 4 */
    uint32_t block_threads;    /* 64, 128 or 256 */
    uint32_t waves_per_simd;   /* persistent-grid occupancy cap, 1..8 (0 = the occupancy API's maximum) */
    uint32_t reserved;
} hexo_tuning;

typedef struct hexo_batch_stats {
    uint64_t n_requests;
    double path_steps;      /* sum over requests of n_paths * simulated steps */
    double nominal_path_steps; /* sum over requests of n_paths * steps * n_expiries */
    double last_run_ms;     /* wall time of the last hexo_batch_run .. sync (host clock) */
    double path_kernel_ms;  /* HIP-event time of the path kernels of the last run (0 if timing off) */
    double reduce_kernel_ms;
    uint64_t path_kernel_launches;
} hexo_batch_stats;

typedef struct hexo_ctx hexo_ctx;
typedef struct hexo_batch hexo_batch;

/* ---- host-only helpers (no GPU needed) ---- */
const char* hexo_strerror(int status);
const char* hexo_version(void);
/* Checks a request against the reference's preconditions (src/HSimulation.tpp:15-21) and the
 * model's domain; returns HEXO_OK or HEXO_ERR_INVALID. */
int hexo_validate_request(const hexo_request* req);
/* Number of outputs of a request: n_payoffs * sum(n_strikes). Outputs are payoff-major
 * (Asian, European, control-variate Asian, geometric blocks, those requested), options in chain-major
 * order. 0 if the request is invalid. */
uint64_t hexo_request_n_outputs(const hexo_request* req);
/* The reference's time grid (src/HSimulation.tpp:30-45 with src/inc/AsianContract.h:29-38),
 * replayed in double: fin_step[c] = 1-based step at which expiry c is finalised,
 * fin_w[c] = (T_c - t_prev) / dt_c; *total_steps = steps simulated per path. */
int hexo_time_grid(const double* expiries, uint32_t n_expiries, const uint32_t* n_strikes, uint32_t steps,
                   uint64_t* fin_step, double* fin_w, uint64_t* total_steps);
/* Weights of expiry `chain`'s Asian average as a linear functional of the path on the time grid,
 * A = sum_j w[j] X(t[j]) for j = 0..n (t[0] = 0, X(0) = S0): the accumulation of
 * AAsianCallNonAdaptive (src/inc/AsianContract.h:25-34) over the grid of src/HSimulation.tpp:30-45,
 * as the path kernels execute it. Returns n (cap >= n + 1 entries needed) or HEXO_ERR_INVALID. */
int64_t hexo_asian_weights(const hexo_request* req, uint32_t chain, double* t, double* w, uint64_t cap);

/* ---- context ---- */
int hexo_ctx_create(int device, hexo_ctx** out);
int hexo_ctx_destroy(hexo_ctx* ctx);
const char* hexo_ctx_last_error(hexo_ctx* ctx);
int hexo_device_info(hexo_ctx* ctx, char* name, size_t name_len, int* n_cu, int* clock_khz);

/* ---- pricing ----
 * hexo_price replaces HSimulation::price<HQEAnderson<ffloat,Policy>> (src/HSimulation.tpp:10-51):
 * prices[o], std_errs[o] for o < hexo_request_n_outputs(req). Synchronous. std_errs may be NULL. */
int hexo_price(hexo_ctx* ctx, const hexo_request* req, double* prices, double* std_errs);
/* No counterpart in the reference. hexo_price without any device or pinned-host allocation, for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the request's device region lives in a fixed 64 MiB array that each GPU holds
 * for the process's lifetime, its host staging in a fixed 16 MiB host array. Same results as
 * hexo_price. A request needing more is refused with HEXO_ERR_INVALID. Calls are serialised by the
 * library (the arrays are shared); each call overwrites the previous one's contents. */
int hexo_price_noalloc(hexo_ctx* ctx, const hexo_request* req, double* prices, double* std_errs);
/* Many requests in one persistent launch; outputs concatenated in request order. */
int hexo_price_batch(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, double* prices, double* std_errs);

/* A batch kept resident on the device: uploaded once, its launch sequence captured in a HIP
 * graph; hexo_batch_run replays it (asynchronously), hexo_batch_sync waits for it. */
int hexo_batch_create(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, hexo_batch** out);
int hexo_batch_run(hexo_batch* b);
int hexo_batch_run_eager(hexo_batch* b); /* same launches without the graph (HIP-event timed) */
int hexo_batch_sync(hexo_batch* b);
/* Outputs of the last run (waits for a run in flight; layout as hexo_price's). */
int hexo_batch_results(hexo_batch* b, double* prices, double* std_errs);
int hexo_batch_stats_get(hexo_batch* b, hexo_batch_stats* st);
/* Exact branch counts of the batch's paths (runs a counting variant of the kernels once):
 * counts[0] = quadratic-branch steps, counts[1] = exponential-branch steps,
 * counts[2] = AS241 central-region normals, counts[3] = AS241 tail normals. */
int hexo_batch_branch_counts(hexo_batch* b, uint64_t counts[4]);
int hexo_batch_destroy(hexo_batch* b);

/* Exact price, under the Heston model, of the geometric counterpart of every option's Asian
 * average (the control variate of HEXO_ASIAN_GEO_CV): E[max(G - K, 0)] per option, chain-major,
 * times exp(-r T) with HEXO_DISCOUNT; FP64 on the GPU (affine characteristic-function recursion
 * over the grid, Lewis' Fourier inversion). n_paths and the payoffs of req are ignored. No
 * counterpart in the reference, which only suggests the control variate. */
int hexo_geo_asian_price(hexo_ctx* ctx, const hexo_request* req, double* prices);

/* ---- pathwise and component entry points (parity tests) ----
 * hexo_paths: per-path (Asian average A_c, interpolated X_T,c) at every expiry, for paths
 * [first, first+n) of the request, finals[(p*n_expiries + c)*2 + {0,1}], in the request's
 * precision (HQEAnderson::operator++ and the policies' accumulate_final_value,
 * src/HSimulation.tpp:52-86, src/inc/AsianContract.h:29-34, src/inc/VanillaContract.h:28-31).
 * At most 2^26 paths and 2^30 finals per call; expiries never finalised read NaN. */
int hexo_paths(hexo_ctx* ctx, const hexo_request* req, uint64_t first, uint64_t n, double* finals);
/* Raw stream words, computed on the GPU: out[i*n_words + j] = word word0+j of path path0+i
 * (replaces the uniform ring refill of RNG::setup_u, src/RNG.cpp:28-33). */
int hexo_rng_fill(hexo_ctx* ctx, uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths,
                  uint64_t word0, uint64_t n_words, uint32_t* out);
/* Philox4x32-10 on the GPU: out[4i..4i+3] = philox(ctr[4i..], key[2i..]). */
int hexo_philox4x32_10(hexo_ctx* ctx, const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out);
/* Word -> N(0,1) on the GPU: AS241 PPND16 in HEXO_FP64 arithmetic, or the FP32 kernel's inverse
 * normal for HEXO_FP32 (polynomials in log2(s (1 - s)), s = min(u, 1 - u); within 1.1e-6 of PPND16)
 * (replaces RNG::setup_g with ppnd16, src/RNG.cpp:34-43, src/as241.f90:15-119). */
int hexo_normals(hexo_ctx* ctx, const uint32_t* words, uint64_t n, uint32_t precision, double* out);

/* ---- launch shape ---- */
int hexo_get_tuning(hexo_ctx* ctx, hexo_tuning* t);
int hexo_set_tuning(hexo_ctx* ctx, const hexo_tuning* t);
/* Times the candidate launch shapes on the given requests and keeps the fastest;
 * *best receives it. reps: timed repetitions per candidate (>= 1). */
int hexo_autotune(hexo_ctx* ctx, const hexo_request* reqs, uint32_t n, uint32_t reps, hexo_tuning* best);

/* ---- measurement ---- */
/* Plain device copy of `bytes` bytes (a known-size kernel for HBM-traffic calibration). */
int hexo_copy_probe(hexo_ctx* ctx, uint64_t bytes, uint32_t reps, double* ms_per_copy);

/* ---- serving: continuous batching ----
 * Requests are admitted by device-memory budget as they arrive; admitted requests are
 * launched in batches on up to `streams` HIP streams, so new work overlaps work in flight.
 * A request above 1e9 path-steps runs in a batch of its own, one such batch at a time and on at
 * most 4 waves per SIMD; smaller requests are batched together beside it and may pass a large
 * request that waits for its turn. */
typedef struct hexo_server_config {
    uint64_t memory_budget_bytes; /* 0 = a fraction of free device memory */
    uint32_t streams;             /* concurrent in-flight batches, 1..4 */
    uint32_t max_batch;           /* max requests per launch (0 = 4096) */
    double batch_window_s;        /* how long the scheduler may wait to grow a batch (0 = none) */
} hexo_server_config;

typedef struct hexo_ticket_info {
    double arrival_s;  /* host steady-clock seconds at submit */
    double admit_s;    /* when admitted (memory reserved) */
    double launch_s;   /* when its batch was launched */
    double done_s;     /* when results were available */
    uint32_t batch_id;
    uint32_t batch_size;
    int32_t status;
} hexo_ticket_info;

typedef struct hexo_server_stats {
    uint64_t submitted, completed, rejected, cancelled, batches;
    uint64_t max_queue_depth;
    uint64_t bytes_in_use, bytes_peak, budget_bytes;
    double path_steps_done;
} hexo_server_stats;

int hexo_server_start(hexo_ctx* ctx, const hexo_server_config* cfg);
int hexo_submit(hexo_ctx* ctx, const hexo_request* req, uint64_t* ticket);
int hexo_poll(hexo_ctx* ctx, uint64_t ticket, int* done);
/* Waits up to timeout_s seconds (< 0 or > 1e7: forever; NaN: HEXO_ERR_INVALID). Once the ticket
 * has finished it returns the ticket's status, copies the outputs if that is HEXO_OK and releases
 * the ticket (a later call gives HEXO_ERR_NOT_FOUND); on HEXO_ERR_TIMEOUT the ticket stays. */
int hexo_wait(hexo_ctx* ctx, uint64_t ticket, double timeout_s, double* prices, double* std_errs,
              hexo_ticket_info* info);
int hexo_cancel(hexo_ctx* ctx, uint64_t ticket);
int hexo_server_stats_get(hexo_ctx* ctx, hexo_server_stats* st);
int hexo_server_stop(hexo_ctx* ctx);
/* Runs a whole arrival trace through the server from a submitting thread that sleeps until
 * each arrival time (seconds from the start); fills per-request outputs (concatenated),
 * per-ticket info, and returns when every request completed. */
int hexo_server_run_trace(hexo_ctx* ctx, const hexo_request* reqs, const double* arrival_s, uint32_t n,
                          double* prices, double* std_errs, hexo_ticket_info* info);
/* Forces the next `count` batch launches (or admissions, kind = 1) of the server to fail with
 * HEXO_ERR_INJECTED; used by tests of failure handling. kind 0 = launch, 1 = admission. */
int hexo_inject_fault(hexo_ctx* ctx, uint32_t kind, uint32_t count);

/* ---- SWIFT European pricer (FP64) ----
 * Replaces SWIFT / SwiftParameters / HDistribution::chf_chf_grad (src/SWIFT.cpp:18-118,
 * src/inc/SWIFT.h:22-160, src/HDistribution.cpp:8-84): European call prices of one expiry chain
 * from the Heston characteristic function on Shannon wavelets, optionally with the gradient in the
 * five Heston parameters (SWIFT::price_opts_grad, the calibration Jacobian). Many chains run in one
 * call. */
typedef struct hexo_swift_grid {   /* SwiftParameters (src/inc/SWIFT.h:22-31) */
    uint32_t m;                     /* wavelet scale */
    uint32_t exp2_m;                /* 2^m */
    double sqrt_exp2_m;             /* sqrt(2^m) */
    double lower, upper;            /* truncation range of the log-price density */
    int32_t k1, k2;                 /* wavelet index range */
    uint32_t J;                     /* Fourier size, a power of two, <= HEXO_SWIFT_MAX_J */
    uint32_t reserved;              /* must be 0 */
} hexo_swift_grid;

#define HEXO_SWIFT_MAX_J 16384u
#define HEXO_SWIFT_FIXED_GRID 1u    /* use hexo_swift_request.grid as given (the reference's
                                       predefined SwiftParameters); otherwise it is chosen from
                                       the model and the strikes (src/SWIFT.cpp:21-35) */

typedef struct hexo_swift_request {
    double v0, theta, rho, kappa, sigma; /* HParams order (src/inc/HDistribution.h:9-14) */
    double S0;                           /* spot */
    double r;                            /* risk-free rate (HDistribution::risk_free) */
    double T;                            /* expiry in years (HDistribution::tau) */
    uint32_t n_strikes;
    uint32_t flags;                      /* HEXO_SWIFT_FIXED_GRID or 0 */
    const double* strikes;               /* n_strikes, > 0 */
    double trunc;                        /* truncation precision of the dynamic grid (the reference
                                            uses 1e-10, or 1e-7 when built with FASTMATH) */
    hexo_swift_grid grid;                /* with HEXO_SWIFT_FIXED_GRID */
} hexo_swift_request;

/* The grid the pricer would choose for a request (host only; SwiftParameters ctor). */
int hexo_swift_grid_for(const hexo_swift_request* req, hexo_swift_grid* out);
/* Prices n requests (SWIFT::price_opts, src/SWIFT.cpp:86-91). prices: sum of n_strikes, chain
 * order. grads: NULL, or 5 per price in the order v0, theta, rho, kappa, sigma
 * (SWIFT::price_opts_grad, src/SWIFT.cpp:92-97). grids_used: NULL or n. Synchronous. */
int hexo_swift_price(hexo_ctx* ctx, const hexo_swift_request* reqs, uint32_t n, double* prices, double* grads,
                     hexo_swift_grid* grids_used);

/* ---- calibration ----
 * Replaces calibrate() (src/HCalibration.cpp:62-117): Levenberg-Marquardt over SWIFT prices and
 * the SWIFT Jacobian of every chain (get_prices_for_levmar / get_jacobian_for_levmar,
 * src/HCalibration.cpp:37-59, levmar 2.6 dlevmar_der), starting from the volume-weighted average
 * Black-Scholes implied volatility (src/BSM.cpp:51-66): v0 = theta = iv^2, rho = -0.2,
 * kappa = sigma = 1. Every price and Jacobian evaluation is one hexo_swift_price call over all
 * chains. */
typedef struct hexo_calib_chain {   /* options_chain (src/inc/Types.h) */
    double T;                        /* time to expiry in years */
    uint32_t n_options;
    uint32_t reserved;               /* must be 0 */
    const double* strikes;
    const double* prices;            /* market prices (option::price) */
    const double* volumes;           /* weights of the initial implied-volatility average
                                        (option::volume); NULL: all 1 */
} hexo_calib_chain;

#define HEXO_CALIB_INITIAL_GIVEN 1u   /* start from hexo_calib_config.initial */

typedef struct hexo_calib_config {
    uint32_t max_iter;               /* 100 in the reference */
    uint32_t flags;                  /* HEXO_CALIB_INITIAL_GIVEN or 0 */
    double initial[5];               /* v0, theta, rho, kappa, sigma */
    double mu_scale;                 /* initial damping / max diag(J^T J): LM_INIT_MU = 1e-3 */
    double eps_grad;                 /* stop when max |J^T e| <= eps_grad (1e-10) */
    double eps_step;                 /* stop when |dp| <= eps_step |p| (1e-10) */
    double eps_err;                  /* stop when |e|^2 <= eps_err (1e-10) */
    double r;                        /* risk-free rate of the model (0.0005, src/inc/BSM.h:10) */
    double trunc;                    /* SWIFT truncation precision (1e-10) */
} hexo_calib_config;

typedef struct hexo_calib_result {
    double params[5];                /* v0, theta, rho, kappa, sigma */
    double initial[5];               /* the starting point used */
    double err0, err;                /* |e|^2 at the start and at the end */
    uint32_t iterations;
    uint32_t reason;                 /* levmar's stop codes: 1 small gradient, 2 small step,
                                        3 iteration limit, 4 singular system, 5 no further
                                        reduction, 6 small error, 7 invalid values. The
                                        reference accepts 2 and 6. */
    uint32_t n_price_evals, n_jacobian_evals;
} hexo_calib_result;

/* Fills a config with the reference's settings. */
int hexo_calib_defaults(hexo_calib_config* cfg);
/* Calibrates the five Heston parameters to the market prices. cfg NULL: the reference's
 * settings. Returns HEXO_OK when the iteration ran; out->reason says how it stopped. */
int hexo_calibrate(hexo_ctx* ctx, double S0, const hexo_calib_chain* chains, uint32_t n_chains,
                   const hexo_calib_config* cfg, hexo_calib_result* out);

#ifdef __cplusplus
}
#endif
#endif /* HEXO_H */
