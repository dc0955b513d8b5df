// calib.cpp -- Heston calibration over the GPU SWIFT pricer: the reference's calibrate()
// (src/HCalibration.cpp:62-117) with its initial point (src/BSM.cpp:22-66) and
// Levenberg-Marquardt as in levmar 2.6's dlevmar_der (the reference's dependency, fetched by
// Makefile.am:102-108; Madsen, Nielsen and Tingleff 2004, Alg. 3.16, with levmar's stopping
// rules). Prices and Jacobians of all chains at a parameter point are one hexo_swift_price call.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/hexo.h"
#include "engine.hpp"

namespace {

constexpr double kYearlyRiskFree = 0.0005;  // src/inc/BSM.h:10

double norm_cdf(double v) { return 0.5 * std::erfc(-v * M_SQRT1_2); }

// Black-Scholes call at the reference's fixed rate (src/BSM.cpp:12-21).
double bs_call(double S, double K, double sigma, double T) {
    const double sd = sigma * std::sqrt(T);
    const double d1 = (std::log(S / K) + (kYearlyRiskFree + 0.5 * sigma * sigma) * T) / sd;
    const double d2 = (std::log(S / K) + (kYearlyRiskFree - 0.5 * sigma * sigma) * T) / sd;
    return S * norm_cdf(d1) - K * std::exp(-kYearlyRiskFree * T) * norm_cdf(d2);
}

// Implied volatility by bracketing from 1/64 and bisection to 1e-6 (src/BSM.cpp:22-50);
// -1 below the arbitrage bound, -2 at or above the spot.
double imp_vol(double S, double K, double price, double T) {
    if (price <= S - std::exp(-kYearlyRiskFree * T) * K) return -1.0;
    if (price >= S) return -2.0;
    double i = 1.0 / 64, lo, hi;
    if (bs_call(S, K, i, T) < price) {
        lo = i;
        do i *= 2;
        while (bs_call(S, K, i, T) < price);
        hi = i;
    } else {
        hi = i;
        do i *= 0.5;
        while (bs_call(S, K, i, T) >= price);
        lo = i;
    }
    double mid;
    do {
        mid = 0.5 * (lo + hi);
        if (bs_call(S, K, mid, T) < price)
            lo = mid;
        else
            hi = mid;
    } while (std::fabs(hi - lo) >= 1e-6);
    return mid;
}

bool in_domain(const double p[5]) { return p[0] > 0 && p[1] > 0 && p[3] > 0 && p[4] > 0 && p[2] > -1 && p[2] < 1; }

// Solves the 5x5 system A x = b by Gaussian elimination with partial pivoting; false if singular.
bool solve5(double A[5][5], double b[5], double x[5]) {
    for (int c = 0; c < 5; ++c) {
        int piv = c;
        for (int r = c + 1; r < 5; ++r)
            if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        if (!(std::fabs(A[piv][c]) > 0.0) || !std::isfinite(A[piv][c])) return false;
        if (piv != c) {
            for (int k = 0; k < 5; ++k) std::swap(A[c][k], A[piv][k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < 5; ++r) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 5; ++k) A[r][k] -= f * A[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 4; c >= 0; --c) {
        double s = b[c];
        for (int k = c + 1; k < 5; ++k) s -= A[c][k] * x[k];
        x[c] = s / A[c][c];
    }
    for (int c = 0; c < 5; ++c)
        if (!std::isfinite(x[c])) return false;
    return true;
}

// The model side: SWIFT prices (and Jacobian) of every chain at p, one GPU call.
struct Model {
    hexo_ctx* ctx;
    double S0, r, trunc;
    std::vector<hexo_swift_request> reqs;
    size_t n_obs = 0;
    // status of the last evaluation: HEXO_OK, or the failing call's code
    int eval(const double p[5], double* prices, double* jac) {
        for (auto& q : reqs) {
            q.v0 = p[0];
            q.theta = p[1];
            q.rho = p[2];
            q.kappa = p[3];
            q.sigma = p[4];
        }
        return hexo_swift_price(ctx, reqs.data(), (uint32_t)reqs.size(), prices, jac, nullptr);
    }
};

}  // namespace

extern "C" int hexo_calib_defaults(hexo_calib_config* c) {
    if (!c) return HEXO_ERR_INVALID;
    std::memset(c, 0, sizeof(*c));
    c->max_iter = 100;
    c->mu_scale = 1e-3;
    c->eps_grad = 1e-10;
    c->eps_step = 1e-10;
    c->eps_err = 1e-10;
    c->r = kYearlyRiskFree;
    c->trunc = 1e-10;
    return HEXO_OK;
}

extern "C" int hexo_calibrate(hexo_ctx* ctx, double S0, const hexo_calib_chain* chains, uint32_t n_chains,
                              const hexo_calib_config* cfg_in, hexo_calib_result* out) {
    if (!ctx) return HEXO_ERR_INVALID;
    if (!chains || !out || n_chains == 0 || !(S0 > 0.0) || !std::isfinite(S0))
        return ctx->set_error(HEXO_ERR_INVALID, "calibration needs S0 > 0, chains and a result");
    hexo_calib_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        hexo_calib_defaults(&cfg);
    if (cfg.max_iter == 0 || (cfg.flags & ~HEXO_CALIB_INITIAL_GIVEN) || !(cfg.mu_scale > 0.0) ||
        !(cfg.eps_grad >= 0.0) || !(cfg.eps_step >= 0.0) || !(cfg.eps_err >= 0.0) || !std::isfinite(cfg.r))
        return ctx->set_error(HEXO_ERR_INVALID, "bad calibration config");
    Model M{ctx, S0, cfg.r, cfg.trunc, {}, 0};
    std::vector<double> x;  // market prices
    for (uint32_t c = 0; c < n_chains; ++c) {
        const hexo_calib_chain& ch = chains[c];
        if (ch.n_options == 0 || !ch.strikes || !ch.prices || ch.reserved != 0 || !(ch.T > 0.0))
            return ctx->set_error(HEXO_ERR_INVALID, "chain " + std::to_string(c) + " is malformed");
        hexo_swift_request q{};
        q.S0 = S0;
        q.r = cfg.r;
        q.T = ch.T;
        q.n_strikes = ch.n_options;
        q.strikes = ch.strikes;
        q.trunc = cfg.trunc;
        M.reqs.push_back(q);
        for (uint32_t i = 0; i < ch.n_options; ++i) {
            if (!std::isfinite(ch.prices[i])) return ctx->set_error(HEXO_ERR_INVALID, "market prices must be finite");
            x.push_back(ch.prices[i]);
        }
    }
    const size_t n = x.size();
    // initial point (src/HCalibration.cpp:67-78)
    double p[5];
    if (cfg.flags & HEXO_CALIB_INITIAL_GIVEN) {
        std::memcpy(p, cfg.initial, sizeof(p));
    } else {
        double num = 0.0, den = 0.0;
        for (uint32_t c = 0; c < n_chains; ++c)
            for (uint32_t i = 0; i < chains[c].n_options; ++i) {
                const double iv = imp_vol(S0, chains[c].strikes[i], chains[c].prices[i], chains[c].T);
                const double w = chains[c].volumes ? chains[c].volumes[i] : 1.0;
                if (iv >= 0.0) {
                    num += w * iv;
                    den += w;
                }
            }
        const double iv = den != 0.0 ? num / den : 0.3;
        p[0] = p[1] = iv * iv;
        p[2] = -0.2;
        p[3] = 1.0;
        p[4] = 1.0;
    }
    if (!in_domain(p)) return ctx->set_error(HEXO_ERR_INVALID, "initial point outside the Heston domain");
    std::memset(out, 0, sizeof(*out));
    std::memcpy(out->initial, p, sizeof(p));

    std::vector<double> hx(n), e(n), J(5 * n), hn(n), en(n);
    int rc = M.eval(p, hx.data(), nullptr);
    if (rc) return rc;
    uint32_t nfev = 1, njev = 0;
    for (size_t i = 0; i < n; ++i) e[i] = x[i] - hx[i];
    double eL2 = 0.0;
    for (double v : e) eL2 += v * v;
    out->err0 = eL2;
    double mu = 0.0;
    uint64_t nu = 2;
    uint32_t stop = 0, k = 0;
    const double tau = cfg.mu_scale, eps1 = cfg.eps_grad, eps2 = cfg.eps_step, eps3 = cfg.eps_err;
    while (k < cfg.max_iter) {
        if (eL2 <= eps3) {
            stop = 6;
            break;
        }
        if ((rc = M.eval(p, hn.data(), J.data()))) return rc;
        ++njev;
        double JtJ[5][5] = {}, Jte[5] = {};
        for (size_t i = 0; i < n; ++i) {
            const double* g = &J[5 * i];
            for (int a = 0; a < 5; ++a) {
                Jte[a] += g[a] * e[i];
                for (int b = 0; b < 5; ++b) JtJ[a][b] += g[a] * g[b];
            }
        }
        double jte_inf = 0.0, pL2 = 0.0;
        for (int a = 0; a < 5; ++a) {
            jte_inf = std::max(jte_inf, std::fabs(Jte[a]));
            pL2 += p[a] * p[a];
        }
        if (jte_inf <= eps1) {
            stop = 1;
            break;
        }
        if (k == 0) {
            double dmax = DBL_MIN;
            for (int a = 0; a < 5; ++a) dmax = std::max(dmax, JtJ[a][a]);
            mu = tau * dmax;
        }
        while (true) {
            double A[5][5], b[5], dp[5];
            std::memcpy(A, JtJ, sizeof(A));
            std::memcpy(b, Jte, sizeof(b));
            for (int a = 0; a < 5; ++a) A[a][a] += mu;
            if (solve5(A, b, dp)) {
                double dpL2 = 0.0, pn[5];
                for (int a = 0; a < 5; ++a) {
                    dpL2 += dp[a] * dp[a];
                    pn[a] = p[a] + dp[a];
                }
                if (dpL2 <= eps2 * eps2 * pL2) {
                    stop = 2;
                    break;
                }
                if (dpL2 >= (pL2 + eps2) / (DBL_EPSILON * DBL_EPSILON)) {
                    stop = 4;
                    break;
                }
                // trial points outside the Heston domain, or where the pricer rejects the model
                // (no SWIFT grid within its limits), are rejected steps
                bool priced = false;
                if (in_domain(pn)) {
                    rc = M.eval(pn, hn.data(), nullptr);
                    if (rc != HEXO_OK && rc != HEXO_ERR_INVALID) return rc;
                    priced = rc == HEXO_OK;
                }
                if (priced) {
                    ++nfev;
                    double enL2 = 0.0;
                    for (size_t i = 0; i < n; ++i) {
                        en[i] = x[i] - hn[i];
                        enL2 += en[i] * en[i];
                    }
                    if (!std::isfinite(enL2)) {
                        stop = 7;
                        break;
                    }
                    double dL = 0.0;
                    for (int a = 0; a < 5; ++a) dL += dp[a] * (mu * dp[a] + Jte[a]);
                    const double dF = eL2 - enL2;
                    if (dL > 0.0 && dF > 0.0) {
                        double t = 2.0 * dF / dL - 1.0;
                        t = 1.0 - t * t * t;
                        mu *= std::max(t, 1.0 / 3.0);
                        nu = 2;
                        std::memcpy(p, pn, sizeof(p));
                        e.swap(en);
                        eL2 = enL2;
                        break;
                    }
                }
            }
            mu *= (double)nu;
            const uint64_t nu2 = nu << 1;
            if (nu2 <= nu || nu2 > (1ull << 62)) {
                stop = 5;
                break;
            }
            nu = nu2;
        }
        ++k;
        if (stop) break;
    }
    if (!stop) stop = 3;
    std::memcpy(out->params, p, sizeof(p));
    out->err = eL2;
    out->iterations = k;
    out->reason = stop;
    out->n_price_evals = nfev;
    out->n_jacobian_evals = njev;
    return HEXO_OK;
}
