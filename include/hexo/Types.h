/*
 * hexo/Types.h -- the data types of the reference's pricing interface, for builds that do not
 * have the reference's own headers. Field names, order and meaning follow
 * src/inc/Types.h:14-58 (ffloat, option, options_chain) and src/inc/HDistribution.h:9-24 (HParams).
 * A build inside the reference tree defines HEXO_USE_REFERENCE_TYPES and uses those headers instead.
 */
#pragma once
#include <cstdint>
#include <limits>
#include <utility>
#include <vector>

typedef double ffloat;
typedef uint64_t fuint;

typedef struct {
    ffloat price;    // ask price
    ffloat bid;
    ffloat strike;
    int64_t volume;  // trading volume or open interest, as selected by the caller
} option;

typedef struct OptionsChain {
    std::vector<option> options;
    unsigned int days_to_expiry;
    ffloat time_to_expiry;  // years
    ffloat max_strike;
    ffloat min_strike;
    OptionsChain(unsigned int days_until, ffloat time_until)
        : days_to_expiry(days_until),
          time_to_expiry(time_until),
          max_strike(std::numeric_limits<ffloat>::lowest()),
          min_strike(std::numeric_limits<ffloat>::max()) {}
} options_chain;

typedef struct HP {
    ffloat v_0;    // initial variance
    ffloat v_m;    // long-term variance (theta)
    ffloat rho;    // spot/variance correlation
    ffloat kappa;  // mean-reversion rate
    ffloat sigma;  // volatility of variance
    friend bool operator==(const HP& l, const HP& r) {
        return l.v_0 == r.v_0 && l.v_m == r.v_m && l.rho == r.rho && l.kappa == r.kappa && l.sigma == r.sigma;
    }
} HParams;
