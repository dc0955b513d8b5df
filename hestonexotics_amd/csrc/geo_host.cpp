// geo_host.cpp -- host side of the geometric-average control variate (payoff HEXO_ASIAN_GEO_CV).
//
// The reference prices the arithmetic Asian call by plain Monte Carlo and suggests a
// geometric-Asian control variate (src/inc/HSimulation.h:48-60). The Asian average of expiry c is
// a linear functional of the path on the time grid, A_c = sum_j w_j X_{t_j} (the trapezoid
// accumulation of src/inc/AsianContract.h:25-28 with the step-width changes of
// src/HSimulation.tpp:36-44 and the final term of src/inc/AsianContract.h:29-34). Its geometric
// counterpart G_c = W S0 exp(sum_j (w_j/W) ln(X_{t_j}/S0)), W = sum_j w_j, moves with A_c path by
// path and has an exact price under the Heston model (the affine recursion of geo.hip). The
// estimator is mean(max(A - K, 0) - max(G - K, 0)) + E[max(G - K, 0)].
//
// Here: the weights w_j from the run schedule the path kernels execute, the per-expiry work items
// of the exact-price kernel, and the host-only hexo_asian_weights entry point.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hexo.h"
#include "engine.hpp"

namespace hexo {

void gauss_legendre64(double* x, double* w) {
    // Newton's method on P_64 from the standard cosine initial guesses
    const int n = 64;
    for (int i = 0; i < n / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// One pass over the run schedule: the accumulation weights and grid times as the path kernels
// build them, and at every finalisation fn(F, p, t, w) with the finalised functional's weights
// w[0..p] (divided by T).
template <typename Fn>
static void walk_functionals(const std::vector<HostSeg>& segs, uint32_t flags, Fn fn) {
    std::vector<double> acc{0.0}, tt{0.0}, w;
    uint64_t p = 0;
    for (const HostSeg& S : segs) {
        for (uint32_t i = 1; i <= S.n_steps; ++i) {
            ++p;
            tt.push_back(tt.back() + S.delta);  // the reference's cur_time += delta
            acc.push_back(0.0);
            if (i < S.n_steps) {  // accumulate_value after every step of the run but its last
                acc[p - 1] += 0.5 * S.delta;
                acc[p] += 0.5 * S.delta;
            }
        }
        for (const HostFin& F : S.fins) {
            w = acc;
            if (flags & HEXO_TRAPEZOID) {
                const double h = 0.5 * (F.T - F.tp);
                w[p - 1] += h * (2.0 - F.w);
                w[p] += h * F.w;
            } else {  // (X - Xp)(T - t_prev)/dt
                w[p] += F.w;
                w[p - 1] -= F.w;
            }
            for (double& x : w) x /= F.T;
            if (!fn(F, p, tt, w)) return;
        }
        // the accumulation after the run's last step uses the step width that follows it
        acc[p - 1] += 0.5 * S.dnext;
        acc[p] += 0.5 * S.dnext;
    }
}

// Grid steps simulated up to the finalisation of expiry `chain` (-1 when it is never finalised),
// from the schedule alone: lets callers bound a walk before it allocates per-step arrays.
static int64_t points_at_fin(const std::vector<HostSeg>& segs, uint32_t chain) {
    uint64_t p = 0;
    for (const HostSeg& S : segs) {
        p += S.n_steps;
        for (const HostFin& F : S.fins)
            if (F.chain == chain) return (int64_t)p;
    }
    return -1;
}

int64_t functional_weights(const std::vector<HostSeg>& segs, uint32_t flags, uint32_t chain, std::vector<double>& t,
                           std::vector<double>& w) {
    int64_t n = -1;
    walk_functionals(segs, flags, [&](const HostFin& F, uint64_t p, const std::vector<double>& tt,
                                      const std::vector<double>& ww) {
        if (F.chain != chain) return true;
        t = tt;
        w = ww;
        n = (int64_t)p;
        return false;
    });
    return n;
}

bool plan_geo(const hexo_request* r, const std::vector<HostSeg>& segs, GeoPlan& g, std::string& why) {
    g = GeoPlan();
    // W for both geometric payoffs; exact prices for the control variate only
    const bool cv = (r->payoffs & HEXO_ASIAN_GEO_CV) != 0;
    if (!(r->payoffs & (HEXO_ASIAN_GEO_CV | HEXO_GEO_ASIAN))) {
        for (const HostSeg& S : segs) g.fin_W.insert(g.fin_W.end(), S.fins.size(), 1.0);
        return true;
    }
    std::vector<uint32_t> opt_begin(r->n_expiries + 1, 0);
    for (uint32_t c = 0; c < r->n_expiries; ++c) opt_begin[c + 1] = opt_begin[c] + r->n_strikes[c];
    // the walk below holds every grid step up to the last expiry it reaches: refuse oversized
    // expiries from the schedule first
    for (uint32_t c = 0; c < r->n_expiries; ++c) {
        const int64_t n = r->n_strikes[c] ? points_at_fin(segs, c) : 0;
        if (n > (int64_t)kGeoMaxPoints) {
            why = "geometric control variate: expiry " + std::to_string(c) + " has " + std::to_string(n) +
                  " grid steps (at most " + std::to_string(kGeoMaxPoints) + " per expiry)";
            return false;
        }
    }
    const double rd = (r->flags & HEXO_DRIFT) ? r->r : 0.0;
    bool ok = true;
    walk_functionals(segs, r->flags, [&](const HostFin& F, uint64_t p, const std::vector<double>& t,
                                        const std::vector<double>& w) {
        if (r->n_strikes[F.chain] == 0) {
            g.fin_W.push_back(1.0);
            return true;
        }
        const int64_t n = (int64_t)p;
        double W = 0.0;
        for (int64_t j = 0; j <= n; ++j) W += w[j];
        if (n < 1 || !(W >= 0.25)) {
            why = "geometric control variate: the Asian average of expiry " + std::to_string(F.chain) +
                  " has weights summing to " + std::to_string(W) + " (< 0.25; use at least 2 steps)";
            return ok = false;
        }
        if (n > (int64_t)kGeoMaxPoints || g.pts.size() / 2 + (size_t)n + 1 > kGeoMaxRequestPoints) {
            why = "geometric control variate: expiry " + std::to_string(F.chain) + " has " + std::to_string(n) +
                  " grid steps (at most " + std::to_string(kGeoMaxPoints) + " per expiry and " +
                  std::to_string(kGeoMaxRequestPoints) + " over a request's expiries)";
            return ok = false;
        }
        g.fin_W.push_back(W);
        const uint32_t pts_begin = (uint32_t)(g.pts.size() / 2);
        // scale of L = sum_j v_j ln(X_{t_j}/S0): mean and variance along E[V_t] (panel width only)
        double mu = 0.0, var = 0.0, suffix = 0.0;
        for (int64_t j = n; j >= 1; --j) {
            suffix += w[j] / W;
            const double tm = 0.5 * (t[j] + t[j - 1]);
            const double ev = r->theta + (r->v0 - r->theta) * std::exp(-r->kappa * tm);
            mu += (t[j] - t[j - 1]) * (rd - 0.5 * ev) * suffix;
            var += (t[j] - t[j - 1]) * ev * suffix * suffix;
        }
        if (!cv) return true;
        for (int64_t j = 0; j <= n; ++j) {
            g.pts.push_back(t[j]);
            g.pts.push_back(w[j] / W);
        }
        const double sd = std::sqrt(std::max(var, 1e-12));
        const double c = W * r->S0;
        const uint32_t ns = r->n_strikes[F.chain];
        for (uint32_t kb = 0; kb < ns; kb += 64) {
            HostGeoItem it;
            it.pts_begin = pts_begin;
            it.n_pts = (uint32_t)n;
            it.n_strikes = std::min<uint32_t>(64, ns - kb);
            it.opt = opt_begin[F.chain] + kb;
            it.c = c;
            double kmax = 0.0;
            for (uint32_t i = 0; i < it.n_strikes; ++i) {
                const double K = r->strikes[it.opt + i];
                if (K > 0.0) kmax = std::max(kmax, std::fabs(std::log(K / c) - mu));
            }
            // a 64-node panel spans at most 1/sd in u and four periods of e^{iu(k - mu)}
            it.h = std::min(1.0 / sd, 8.0 * M_PI / (kmax + 4.0 * sd + 1e-3));
            // bounded work per wavefront: panels x grid steps <= 2^26 recursion steps
            it.max_panels = (uint32_t)std::min<uint64_t>(kGeoMaxPanels, (1ull << 26) / (uint64_t)n);
            // panels evaluated in parallel: up to ten times the u where a Gaussian L's
            // characteristic function falls below 1e-16 (Heston's decays slower; the sum pass
            // continues by itself where needed, and stops at the first decayed panel)
            const double u_max = 10.0 * std::sqrt(2.0 * 36.9 / std::max(var, 1e-12));
            it.p0 = (uint32_t)std::min<double>(std::min<uint32_t>(kGeoMaxP0, it.max_panels),
                                               std::max(2.0, std::ceil(u_max / it.h)));
            g.items.push_back(it);
        }
        return true;
    });
    return ok;
}

}  // namespace hexo

using namespace hexo;

extern "C" int64_t hexo_asian_weights(const hexo_request* r, uint32_t chain, double* t, double* w, uint64_t cap) {
    if (!r || !t || !w) return HEXO_ERR_INVALID;
    if (!validate(r).empty() || chain >= r->n_expiries) return HEXO_ERR_INVALID;
    std::vector<HostSeg> segs;
    uint64_t total = 0;
    if (!build_schedule(r->expiries, r->n_expiries, r->n_strikes, r->steps, segs, total)) return HEXO_ERR_INVALID;
    // refuse before the walk allocates per-step arrays the caller has no room for
    const int64_t p = points_at_fin(segs, chain);
    if (p < 0 || (uint64_t)p + 1 > cap) return HEXO_ERR_INVALID;
    std::vector<double> tt, ww;
    const int64_t n = functional_weights(segs, r->flags, chain, tt, ww);
    if (n < 0) return HEXO_ERR_INVALID;  // never finalised (a trailing expiry without strikes)
    if ((uint64_t)n + 1 > cap) return HEXO_ERR_INVALID;
    std::memcpy(t, tt.data(), sizeof(double) * (size_t)(n + 1));
    std::memcpy(w, ww.data(), sizeof(double) * (size_t)(n + 1));
    return n;
}
