/*
 * hexo/HCalibration.h -- drop-in replacement for the reference's calibrate()
 * (src/inc/HCalibration.h, src/HCalibration.cpp:62-117), header-only, over hexo_calibrate.
 *
 *   HParams p = calibrate(S, all_chains);        // src/Main.cpp:71, :83
 *
 * Same starting point (volume-weighted average Black-Scholes implied volatility), same
 * Levenberg-Marquardt settings (mu 1e-3, thresholds 1e-10, 100 iterations) and the same
 * acceptance (levmar stop reason 2 or 6, else std::runtime_error "levmar failed!"). Prices and
 * Jacobians come from the GPU SWIFT pricer. The model's risk-free rate is the reference's
 * 0.0005 (HDistribution in calibrate(), src/HCalibration.cpp:75).
 */
#pragma once
#include <list>
#include <stdexcept>
#include <string>
#include <vector>

#include "HSimulation.h"

inline HParams calibrate(const ffloat S, const std::list<options_chain>& market_data) {
    std::vector<hexo_calib_chain> chains;
    std::vector<std::vector<double>> cols;  // strikes, prices, volumes per chain
    cols.reserve(3 * market_data.size());
    for (const options_chain& c : market_data) {
        if (c.options.empty()) continue;
        cols.emplace_back();
        cols.emplace_back();
        cols.emplace_back();
        std::vector<double>&K = cols[cols.size() - 3], &P = cols[cols.size() - 2], &W = cols[cols.size() - 1];
        for (const option& o : c.options) {
            K.push_back(o.strike);
            P.push_back(o.price);
            W.push_back((double)o.volume);
        }
        hexo_calib_chain h{};
        h.T = c.time_to_expiry;
        h.n_options = (uint32_t)K.size();
        h.strikes = K.data();
        h.prices = P.data();
        h.volumes = W.data();
        chains.push_back(h);
    }
    if (chains.empty()) throw std::runtime_error("calibrate: no options");
    hexo_ctx* ctx = HSimulation::detail::context();
    hexo_calib_result res{};
    const int rc = hexo_calibrate(ctx, S, chains.data(), (uint32_t)chains.size(), nullptr, &res);
    if (rc != HEXO_OK) throw std::runtime_error(std::string("hexo: ") + hexo_ctx_last_error(ctx));
    if (res.reason != 6 && res.reason != 2) throw std::runtime_error("levmar failed! ");
    return HParams{res.params[0], res.params[1], res.params[2], res.params[3], res.params[4]};
}
