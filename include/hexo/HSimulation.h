/*
 * hexo/HSimulation.h -- drop-in replacement for the reference's src/inc/HSimulation.h and
 * src/HSimulation.tpp, header-only, over the C ABI in hexo.h (libhexo.so).
 *
 * The reference spelling keeps working:
 *
 *   std::vector<ffloat> prices =
 *       HSimulation::price<HSimulation::HQEAnderson<ffloat, AAsianCallNonAdaptive>>(
 *           p, S, all_chains, n_simulations, n_opts, steps);          // src/Main.cpp:88
 *
 * Scheme = HQEAnderson<accumulate_t, OptionPolicy> (src/inc/HSimulation.h:22-47):
 *   OptionPolicy AAsianCallNonAdaptive   -> Asian calls   (src/inc/AsianContract.h:14-48)
 *   OptionPolicy EuropeanCallNonAdaptive -> European calls (src/inc/VanillaContract.h:14-40)
 *   accumulate_t double -> the plain FP64 kernel (HEXO_FP64); float -> FP32 stepping with FP64
 *   reduction (HEXO_FP32).
 * Variance reduction, which the reference suggests beside price (src/inc/HSimulation.h:48-60),
 * as further policies and a scheme (extensions, not in the reference):
 *   OptionPolicy AAsianCallGeometricCV           -> Asian calls with the geometric-average control
 *                                                   variate and the regression coefficient
 *                                                   beta = Cov(A, G) / Var G from the same paths
 *                                                   (HEXO_ASIAN_GEO_CV with HEXO_CV_REGRESSION):
 *                                                   its standard error is never above the plain
 *                                                   Asian one
 *   OptionPolicy AAsianCallGeometricCVRegression -> the same (the explicit name)
 *   OptionPolicy AAsianCallGeometricCVUnitCoefficient -> coefficient 1 instead (opt-in): cheaper by
 *                                                   one payoff block, but worse than plain Monte
 *                                                   Carlo where A and G decorrelate (few steps,
 *                                                   the reference's final price increment
 *                                                   dominating the average)
 *   Scheme HQEAndersonMartingale<accumulate_t, OptionPolicy> -> Andersen's martingale-corrected
 *                                                   step (HEXO_MARTINGALE)
 * The estimator is the reference's: undiscounted, no drift term, its time grid and Asian
 * finalisation; outputs in chain-major option order (src/HSimulation.tpp:40). Differences: the
 * paths come from a counter-based Philox4x32-10 stream (seed set with HSimulation::set_seed, 1 by
 * default) instead of per-thread shishua streams, the path count is exact (the reference rounds
 * down to a multiple of the thread count), and standard errors are available (price_with_errors).
 * Errors raise std::runtime_error (the reference's convention, src/inc/AsianContract.h:30).
 */
#pragma once
#include <cstdint>
#include <cstdlib>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../hexo.h"
#ifndef HEXO_USE_REFERENCE_TYPES
#include "Types.h"
#endif

// Option policies of the reference, as tags (their accumulation logic runs in the kernel):
// the payoff blocks to price, flags, and which block (in payoff bit order) price returns.
struct AAsianCallNonAdaptive {
    static constexpr uint32_t hexo_payoff = HEXO_ASIAN;
    static constexpr uint32_t hexo_flags = 0;
    static constexpr uint32_t hexo_block = 0;
};
struct EuropeanCallNonAdaptive {
    static constexpr uint32_t hexo_payoff = HEXO_EUROPEAN;
    static constexpr uint32_t hexo_flags = 0;
    static constexpr uint32_t hexo_block = 0;
};
struct AAsianCallGeometricCV {  // blocks: Asian, control variate, geometric
    static constexpr uint32_t hexo_payoff = HEXO_ASIAN | HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN;
    static constexpr uint32_t hexo_flags = HEXO_CV_REGRESSION;
    static constexpr uint32_t hexo_block = 1;
};
struct AAsianCallGeometricCVRegression : AAsianCallGeometricCV {};
struct AAsianCallGeometricCVUnitCoefficient {
    static constexpr uint32_t hexo_payoff = HEXO_ASIAN_GEO_CV;
    static constexpr uint32_t hexo_flags = 0;
    static constexpr uint32_t hexo_block = 0;
};

namespace HSimulation {

enum positions { X = 0, V = 1 };

template <typename accumulate_t, class OptionPolicy>
class HQEAnderson {
   public:
    static constexpr uint32_t payoff = OptionPolicy::hexo_payoff;
    static constexpr uint32_t flags = OptionPolicy::hexo_flags;
    static constexpr uint32_t block = OptionPolicy::hexo_block;
    static constexpr uint32_t precision = sizeof(accumulate_t) == sizeof(double) ? HEXO_FP64 : HEXO_FP32;
};

// The QE scheme with Andersen's martingale correction of K0 (HEXO_MARTINGALE).
template <typename accumulate_t, class OptionPolicy>
class HQEAndersonMartingale : public HQEAnderson<accumulate_t, OptionPolicy> {
   public:
    static constexpr uint32_t flags = OptionPolicy::hexo_flags | HEXO_MARTINGALE;
};

namespace detail {

struct Ctx {
    hexo_ctx* ctx = nullptr;
    uint64_t seed = 1;
    std::mutex mu;
    ~Ctx() {
        if (ctx) hexo_ctx_destroy(ctx);
    }
};

inline Ctx& state() {
    static Ctx c;
    return c;
}

inline hexo_ctx* context() {
    Ctx& c = state();
    std::lock_guard<std::mutex> g(c.mu);
    if (!c.ctx) {
        const char* dev = std::getenv("HEXO_DEVICE");
        const int rc = hexo_ctx_create(dev ? std::atoi(dev) : 0, &c.ctx);
        if (rc != HEXO_OK) throw std::runtime_error(std::string("hexo: ") + hexo_strerror(rc));
    }
    return c.ctx;
}

}  // namespace detail

inline void set_seed(uint64_t seed) { detail::state().seed = seed; }

// Prices every option of every chain; also returns standard errors.
template <class Scheme>
std::vector<ffloat> price_with_errors(const HParams& p, const ffloat S, const std::list<options_chain>& all_chains,
                                      unsigned int n_simulations, unsigned int n_opts, unsigned int steps,
                                      std::vector<ffloat>* std_errs) {
    std::vector<double> T;
    std::vector<uint32_t> ns;
    std::vector<double> K;
    for (const options_chain& c : all_chains) {
        T.push_back(c.time_to_expiry);
        ns.push_back((uint32_t)c.options.size());
        for (const option& o : c.options) K.push_back(o.strike);
    }
    if (K.size() != n_opts) throw std::runtime_error("hexo: n_opts does not match the chains");
    hexo_request r{};
    r.v0 = p.v_0;
    r.theta = p.v_m;
    r.rho = p.rho;
    r.kappa = p.kappa;
    r.sigma = p.sigma;
    r.S0 = S;
    r.r = 0.0;
    r.n_paths = n_simulations;
    r.seed = detail::state().seed;
    r.path_offset = 0;
    r.steps = steps;
    r.n_expiries = (uint32_t)T.size();
    r.expiries = T.data();
    r.n_strikes = ns.data();
    r.strikes = K.data();
    r.payoffs = Scheme::payoff;
    r.flags = Scheme::flags;
    r.precision = Scheme::precision;
    const uint64_t n_out = hexo_request_n_outputs(&r);
    std::vector<double> prices(n_out ? n_out : n_opts), ses(prices.size());
    hexo_ctx* ctx = detail::context();
    const int rc = hexo_price(ctx, &r, prices.data(), ses.data());
    if (rc != HEXO_OK)
        throw std::runtime_error(std::string("hexo: ") + hexo_strerror(rc) + ": " + hexo_ctx_last_error(ctx));
    const size_t b0 = (size_t)Scheme::block * n_opts;
    if (std_errs) std_errs->assign(ses.begin() + b0, ses.begin() + b0 + n_opts);
    return std::vector<ffloat>(prices.begin() + b0, prices.begin() + b0 + n_opts);
}

// src/inc/HSimulation.h:61-63
template <class Scheme>
std::vector<ffloat> price(const HParams& p, const ffloat S, const std::list<options_chain>& all_chains,
                          unsigned int n_simulations, unsigned int n_opts, unsigned int steps) {
    return price_with_errors<Scheme>(p, S, all_chains, n_simulations, n_opts, steps, nullptr);
}

}  // namespace HSimulation
