/*
 * hexo_oracle.h -- CPU oracle for the Heston QE Monte-Carlo pricing path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference's
 * HSimulation::price<HQEAnderson<ffloat,Policy>> path. It is the checker that
 * the HIP product is compared against, and the "port" CPU baseline timed by
 * bench.py. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; nothing in hestonexotics_amd/ links or calls it.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/HSimulation.tpp:10-51   driver: per-path time loop, multi-expiry finalisation
 *   src/HSimulation.tpp:52-86   HQEAnderson::operator++ (one QE step)
 *   src/HSimulation.tpp:87-94   state reset
 *   src/inc/HSimulation.h:12    PSI_C = 1.5
 *   src/inc/AsianContract.h:25-47, src/inc/VanillaContract.h:28-39  payoff policies
 *   src/RNG.cpp:28-43           uniform mapping, normals through ppnd16
 *   src/as241.f90:15-119        AS241 PPND16 (coefficients and REAL*4 semantics)
 *
 * The RNG is deliberately not shishua (src/inc/RNG.h:9): it is Philox4x32-10
 * (Salmon, Moraes, Dror, Shaw, SC'11), a counter-based generator, restated here.
 *
 * Parity pins (see tests/): Random123 known-answer vectors and rocRAND's
 * philox4x32_10 engine for Philox; scipy.special.ndtri and the AS241 coefficient
 * hash sums for AS241; QE moment invariants (Andersen 2008) for the variance
 * step; the Black-Scholes limit (sigma -> 0), the K = 0 Asian invariant and the
 * SWIFT/closed-form European prices for the full path.
 */
#ifndef HEXO_ORACLE_H
#define HEXO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* request flags (same bit values as the product ABI, include/hexo.h) */
#define ORC_ASIAN 1u
#define ORC_EUROPEAN 2u
#define ORC_DRIFT 1u     /* add r*delta to the log-price step (extension; reference has r = 0) */
#define ORC_DISCOUNT 2u  /* multiply the estimator by exp(-r T) (extension) */
#define ORC_TRAPEZOID 4u /* corrected trapezoid Asian average (extension) */
#define ORC_MARTINGALE 8u /* Andersen's martingale-corrected K0 (extension) */
#define ORC_ASIAN_GEO_CV 4u /* payoff: Asian call with the geometric-average control variate (extension) */
#define ORC_GEO_ASIAN 8u    /* payoff: the geometric-average call itself, plain Monte Carlo (extension) */

/* ---- Philox4x32-10 ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Word w of path p's stream is element w%4 of Philox(ctr = {w/4, lo32(p), hi32(p), stream},
 * key = {lo32(seed), hi32(seed)}). out[i*n_words + j] = word (word0 + j) of path (path0 + i). */
void orc_rng_words(uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths, uint64_t word0,
                   uint64_t n_words, uint32_t* out);

/* ---- uniforms and AS241 ---- */
double orc_uniform_from_word(uint32_t k); /* (k + 1/2) * 2^-32, exact in double */
double orc_ppnd16_f64(double p, int* ifault);     /* AS241 evaluated in double */
double orc_ppnd16_asbuilt(double p, int* ifault); /* AS241 with the reference's REAL*4 arithmetic */
double orc_normal_f64_from_word(uint32_t k);      /* ppnd16_f64(uniform(k)) */
float orc_normal_f32_from_word(uint32_t k);       /* the FP32 kernel's inverse normal of word k */
void orc_as241_coefficients(double out[45]);      /* a0..a7,b1..b7,c0..c7,d1..d7,e0..e7,f1..f7 */

/* ---- one QE step (reference form, FP64) ----
 * params = {v0, theta, rho, kappa, sigma} (HParams order). martingale != 0 applies Andersen's
 * martingale-corrected K0 (ORC_MARTINGALE). Returns 1 for the quadratic branch, 0 for the
 * exponential branch. */
int orc_qe_step(const double params[5], double r_drift, double delta, int martingale, double* V, double* logX,
                uint32_t w_v, uint32_t w_x);
/* E[X'/X | V] of one step by a 2^log2_nv x 2^log2_nx midpoint rule over the two uniforms. */
double orc_qe_step_mean(const double params[5], double r_drift, double delta, int martingale, double V,
                        uint32_t log2_nv, uint32_t log2_nx);

/* ---- the reference time grid for one path (path independent) ----
 * Replays the double-precision loop of src/HSimulation.tpp:35-44 and reports, for every
 * chain c that is finalised: fin_step[c] (1-based step index), fin_w[c] = (T_c - t_prev)/dt.
 * Returns the number of steps simulated, or -1 on bad input. */
int64_t orc_time_grid(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps,
                      uint64_t* fin_step, double* fin_w);

/* ---- Monte-Carlo price (reference semantics, FP64) ----
 * Prices every option of every chain. Outputs are payoff-major: for each payoff present in
 * `payoffs` (Asian, European, the Asian with the geometric control variate, the geometric call),
 * n_opts entries in chain-major order. The control-variate price is mean(A payoff - G payoff) plus the
 * exact geometric price (orc_geo_asian_price), times the discount factor.
 * finals (nullable): [n_paths][n_exp][2] = (Asian average A_c, interpolated X_T) per path.
 * counts (nullable): [2] = (quadratic-branch steps, exponential-branch steps).
 * n_threads: OpenMP threads (0 = library default). Returns 0 or a negative error. */
int orc_price(const double params[5], double S0, double r, uint32_t flags, uint32_t payoffs, uint64_t n_paths,
              uint64_t path_offset, uint64_t seed, uint32_t steps, uint32_t n_exp, const double* T,
              const uint32_t* n_strikes, const double* strikes, double* prices, double* std_errs, double* finals,
              uint64_t* counts, int n_threads);

void orc_qe_moments(const double params[5], double delta, double V, uint64_t n, uint64_t seed, double out[6]);

/* Wall-clock of the last orc_price call (seconds, steady clock, pricing loop only). */
double orc_last_seconds(void);
int orc_max_threads(void);

/* ---- geometric-average control variate (geo_oracle.c) ----
 * orc_asian_weights: weights w[0..n] (and times t[0..n]) of chain `chain`'s Asian functional
 * A = sum_j w_j X_{t_j} (src/HSimulation.tpp:30-45, src/inc/AsianContract.h:23-38); returns n,
 * -2 if the chain is never finalised. orc_geo_cf: E[exp(s sum_j v_j ln(X_{t_j}/S0))] under
 * Heston. orc_geo_price_weights: E[(c e^L - K)^+] with c = W S0, L = sum_j (w_j/W) ln(X_{t_j}/S0),
 * W = sum_j w_j (undiscounted). orc_geo_asian_price: that price for every option of a request
 * (chain-major). */
int64_t orc_asian_weights(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps, uint32_t flags,
                          uint32_t chain, double* t, double* w, uint64_t cap);
#ifndef __cplusplus
double _Complex orc_geo_cf(const double params[5], double r, const double* t, const double* v, int64_t n,
                           double _Complex s);
#endif
void orc_geo_cf_out(const double params[5], double r, const double* t, const double* v, int64_t n, double s_re,
                    double s_im, double out[2]);
int orc_geo_price_weights(const double params[5], double S0, double r, const double* t, const double* w, int64_t n,
                          uint32_t n_strikes, const double* strikes, double* out);
/* the same with the stop rule's tolerance as an argument (orc_geo_price_weights: 1e-15) */
int orc_geo_price_weights_tol(const double params[5], double S0, double r, const double* t, const double* w, int64_t n,
                              uint32_t n_strikes, const double* strikes, double tol, double* out);
int orc_geo_asian_price(const double params[5], double S0, double r, const double* T, uint32_t n_exp,
                        const uint32_t* n_strikes, uint32_t steps, uint32_t flags, const double* strikes,
                        double* out);

/* ---- SWIFT European oracle (swift_oracle.c) ----
 * Restates src/SWIFT.cpp:18-118 and src/HDistribution.cpp:8-38,86-113.
 * params = {v0, theta, rho, kappa, sigma}. If swift_params is non-NULL it holds the
 * predefined parameters {m, exp2_m, sqrt_exp2_m, lower, upper, k_1, k_2, J}; otherwise they
 * are chosen dynamically (src/SWIFT.cpp:21-35) with truncation precision trunc (1e-7 in the
 * reference's release build). Strikes share expiry T. Returns 0 or a negative error. */
int orc_swift_price(const double params[5], double S, double r, double T, uint32_t n_strikes, const double* strikes,
                    const double* swift_params, double trunc, double* prices, double* used_params);
/* The same with the price gradient (SWIFT::price_opts_grad, src/SWIFT.cpp:92-97): grads[5 o + k]
 * is d price_o / d params[k] with the SWIFT parameters held fixed, from the characteristic
 * function evaluated in forward-mode dual arithmetic. grads may be NULL. */
int orc_swift_price_grad(const double params[5], double S, double r, double T, uint32_t n_strikes,
                         const double* strikes, const double* swift_params, double trunc, double* prices,
                         double* grads, double* used_params);
/* Heston characteristic function of the reference (src/HDistribution.cpp:12-38) at real u. */
void orc_heston_chf(const double params[5], double tau, double u, double out[2]);

#ifdef __cplusplus
}
#endif
#endif
