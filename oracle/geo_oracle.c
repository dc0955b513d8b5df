/* geo_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker for the geometric-average control
 * variate, never shipped or measured).
 *
 * The reference prices the arithmetic Asian call AAsianCallNonAdaptive by plain Monte Carlo and
 * suggests a geometric-Asian control variate (src/inc/HSimulation.h:48-60) without building it.
 * This file restates the two pieces such a control variate needs:
 *   1. the weights of the reference's Asian functional A_c = sum_j w_j X_{t_j} over the grid
 *      points, from the time loop and accumulation of src/HSimulation.tpp:30-45 and
 *      src/inc/AsianContract.h:23-38 (the same loop as simulate_path in hexo_oracle.c, run on
 *      weights instead of values);
 *   2. the exact price of the geometric counterpart G_c = W S0 exp(sum_j (w_j / W) ln(X_{t_j}/S0)),
 *      W = sum_j w_j, under the Heston model: the characteristic function of the weighted sum of
 *      log-prices by the affine backward recursion (at each grid point the log-price coefficient
 *      grows by the point's weight; between points the Heston Riccati equations are solved in
 *      closed form with the running variance coefficient as initial value), and the call price
 *      by Lewis' inversion E[(e^L - e^k)^+] = E[e^L] - e^{k/2}/pi int_0^inf
 *      Re(e^{iuk} E[e^{(1/2 - iu)L}]) / (u^2 + 1/4) du.
 * Pinned by the GBM limit (sigma -> 0: L is Gaussian, a Black-type closed form), by the one-date
 * functional (the Heston European call of the SWIFT oracle, itself pinned by the reference's
 * golden prices) and by Monte Carlo with the oracle's own QE paths.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hexo_oracle.h"

int64_t orc_asian_weights(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps, uint32_t flags,
                          uint32_t chain, double* t_out, double* w_out, uint64_t cap) {
    if (n_exp == 0 || steps == 0 || chain >= n_exp || cap < 2) return -1;
    uint64_t n_opts = 0;
    for (uint32_t c = 0; c < n_exp; ++c) n_opts += n_strikes[c];
    if (n_opts == 0) return -1;
    double* wacc = (double*)calloc(cap, sizeof(double));
    double* tp_ = (double*)calloc(cap, sizeof(double));
    if (!wacc || !tp_) {
        free(wacc);
        free(tp_);
        return -3;
    }
    double t = 0.0, tp = 0.0, dt = T[0] / steps;
    uint32_t c = 0;
    uint64_t priced = 0, k = 0;
    int64_t result = -2; /* chain never finalised */
#define ORC_WSTEP()                 \
    do {                            \
        tp = t;                     \
        t += dt;                    \
        ++k;                        \
        if (k >= cap) goto done;    \
        tp_[k] = t;                 \
    } while (0)
    ORC_WSTEP();
    while (priced < n_opts) {
        while (t >= T[c] && priced < n_opts) {
            if (c == chain) {
                const double wf = (T[c] - tp) / dt;
                for (uint64_t j = 0; j <= k; ++j) w_out[j] = wacc[j];
                if (flags & ORC_TRAPEZOID) {
                    const double h = 0.5 * (T[c] - tp);
                    w_out[k - 1] += h * (2.0 - wf);
                    w_out[k] += h * wf;
                } else {
                    w_out[k] += wf;
                    w_out[k - 1] -= wf;
                }
                for (uint64_t j = 0; j <= k; ++j) {
                    w_out[j] /= T[c];
                    t_out[j] = tp_[j];
                }
                result = (int64_t)k;
                goto done;
            }
            priced += n_strikes[c];
            if (priced < n_opts) {
                ++c;
                dt = T[c] / steps;
            }
        }
        if (priced >= n_opts) break;
        wacc[k] += 0.5 * dt; /* acc += dt (X_k + X_{k-1}) / 2 */
        wacc[k - 1] += 0.5 * dt;
        ORC_WSTEP();
    }
#undef ORC_WSTEP
done:
    free(wacc);
    free(tp_);
    return result;
}

/* ---- complex helpers ---- */
static double complex clog1p_(double complex z) {
    if (cabs(z) < 1e-3) return z * (1.0 - z * (0.5 - z * (1.0 / 3.0 - z * 0.25)));
    return clog(1.0 + z);
}
static double complex cexpm1_(double complex w) {
    if (cabs(w) < 1e-3) return w * (1.0 + w * (0.5 + w * (1.0 / 6.0 + w / 24.0)));
    return cexp(w) - 1.0;
}

/* E[exp(s L)], L = sum_{j=1..n} v_j ln(X_{t_j}/S0), t_0 = 0 < t_1 < ... < t_n.
 * Backward over the points: a = s sum_{i >= j} v_i is the log-price coefficient at t_j; over
 * (t_{j-1}, t_j] the pair (A, B) of exp(A + a y + B V) solves
 *   B' = sigma^2 B^2 / 2 + (rho sigma a - kappa) B + (a^2 - a) / 2,   A' = kappa theta B + r a,
 * with B(0) = the later points' B. With beta = kappa - rho sigma a, d = sqrt(beta^2 - sigma^2 (a^2 - a)),
 * the stable root B- = (a^2 - a)/(beta + d) and e = exp(-d tau):
 *   B(tau) = B- + (B - B-) e 2d / D,  D = (beta + d - sigma^2 B) - sigma^2 (B- - B) e,
 *   A(tau) = A + r a tau + kappa theta (B- tau - 2/sigma^2 log(1 + sigma^2 (B- - B)(1 - e)/(2d))),
 * forms without 1/sigma^2 cancellation, so that sigma -> 0 stays accurate. */
double complex orc_geo_cf(const double p[5], double r, const double* t, const double* v, int64_t n,
                          double complex s) {
    const double v0 = p[0], th = p[1], rho = p[2], kap = p[3], sig = p[4], s2 = sig * sig;
    double complex a = 0.0, A = 0.0, B = 0.0;
    for (int64_t j = n; j >= 1; --j) {
        const double tau = t[j] - t[j - 1];
        a += s * v[j];
        const double complex bt = kap - rho * sig * a;
        const double complex q = a * a - a;
        const double complex d = csqrt(bt * bt - s2 * q);
        const double complex Bm = cabs(bt + d) >= cabs(bt - d) ? q / (bt + d) : (bt - d) / s2;
        const double complex om = -cexpm1_(-d * tau);
        const double complex e = 1.0 - om;
        const double complex D = (bt + d - s2 * B) - s2 * (Bm - B) * e;
        const double complex Bn = Bm + (B - Bm) * e * 2.0 * d / D;
        const double complex z = s2 * (Bm - B) * om / (2.0 * d);
        /* log(1 + z) * 2 / sigma^2 -> (B- - B)(1 - e)/d as sigma -> 0 */
        const double complex lz = s2 > 0.0 ? clog1p_(z) * (2.0 / s2) : (Bm - B) * om / d;
        A += r * a * tau + kap * th * (Bm * tau - lz);
        B = Bn;
    }
    return cexp(A + B * v0);
}

void orc_geo_cf_out(const double p[5], double r, const double* t, const double* v, int64_t n, double s_re,
                    double s_im, double out[2]) {
    const double complex z = orc_geo_cf(p, r, t, v, n, s_re + I * s_im);
    out[0] = creal(z);
    out[1] = cimag(z);
}

/* Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n from Tricomi's initial guesses). */
static void gauss_legendre(int n, double* x, double* w) {
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        x[i] = -z;
        x[n - 1 - i] = z;
        w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

#define ORC_GL 64
#define ORC_MAX_PANELS 20000

/* Stop rule: beyond u0 = pnl h the integrand |phi| / (u^2 + 1/4) leaves at most sup|phi| / u0 to
 * the integral. The sum stops after the first two consecutive panels (after the first) whose
 * largest |phi| is below tol u0; tol = 1e-15 in orc_geo_price_weights. Two panels, not one: |phi|
 * need not fall monotonically, and one panel can sit in a dip. */
int orc_geo_price_weights_tol(const double p[5], double S0, double r, const double* t, const double* w, int64_t n,
                              uint32_t n_strikes, const double* strikes, double tol, double* out) {
    if (n < 1) return -1;
    double W = 0.0;
    for (int64_t j = 0; j <= n; ++j) W += w[j];
    if (!(W > 0.0)) return -2;
    double* v = (double*)malloc(sizeof(double) * (size_t)(n + 1));
    double* kk = (double*)malloc(sizeof(double) * n_strikes);
    double* integ = (double*)calloc(n_strikes, sizeof(double));
    if (!v || !kk || !integ) {
        free(v);
        free(kk);
        free(integ);
        return -3;
    }
    for (int64_t j = 0; j <= n; ++j) v[j] = w[j] / W;
    const double c = W * S0;
    /* scale of L: mean and variance from E[V_t] along the grid (panel width only) */
    double mu = 0.0, var = 0.0, suffix = 0.0;
    for (int64_t j = n; j >= 1; --j) {
        suffix += v[j];
        const double tm = 0.5 * (t[j] + t[j - 1]);
        const double ev = p[1] + (p[0] - p[1]) * exp(-p[3] * tm);
        mu += (t[j] - t[j - 1]) * (r - 0.5 * ev) * suffix;
        var += (t[j] - t[j - 1]) * ev * suffix * suffix;
    }
    const double sd = sqrt(var > 1e-12 ? var : 1e-12);
    double kmax = 0.0;
    for (uint32_t i = 0; i < n_strikes; ++i) {
        kk[i] = strikes[i] > 0.0 ? log(strikes[i] / c) : -INFINITY;
        if (isfinite(kk[i]) && fabs(kk[i] - mu) > kmax) kmax = fabs(kk[i] - mu);
    }
    /* panels: 64 nodes over at most ~4 periods of e^{iu(k - mu)} and 1/sd of u */
    double h = 1.0 / sd;
    const double hosc = 8.0 * M_PI / (kmax + 4.0 * sd + 1e-3);
    if (hosc < h) h = hosc;
    double gx[ORC_GL], gw[ORC_GL];
    gauss_legendre(ORC_GL, gx, gw);
    const double complex e1 = orc_geo_cf(p, r, t, v, n, 1.0);
    int converged = 0, below = 0;
    /* panel 0 holds the integrand's 1/(u^2 + 1/4) peak: min(64, ceil(h/16)) sub-panels of 64 nodes
       (64 nodes over [0, 16] integrate it to 1e-13, over [0, 120] only to 1e-4) */
    const int n_first = (int)fmin(64.0, ceil(h / 16.0));
    for (int pnl = 0; pnl < ORC_MAX_PANELS && !converged; ++pnl) {
        double amax = 0.0;
        const int n_sub = pnl == 0 ? n_first : 1;
        const double hs = h / n_sub;
        for (int j = 0; j < n_sub; ++j)
            for (int q = 0; q < ORC_GL; ++q) {
                const double u = (pnl + j + 0.5 * (gx[q] + 1.0)) * hs, wq = 0.5 * hs * gw[q];
                const double complex ph = orc_geo_cf(p, r, t, v, n, 0.5 - I * u);
                const double a = cabs(ph);
                if (a > amax) amax = a;
                for (uint32_t i = 0; i < n_strikes; ++i)
                    if (isfinite(kk[i])) integ[i] += wq * creal(cexp(I * u * kk[i]) * ph) / (u * u + 0.25);
            }
        const int now_below = pnl > 0 && amax < tol * pnl * h;
        converged = now_below && below;
        below = now_below;
    }
    for (uint32_t i = 0; i < n_strikes; ++i)
        out[i] = !converged ? NAN
                 : isfinite(kk[i]) ? c * (creal(e1) - exp(0.5 * kk[i]) / M_PI * integ[i])
                                   : c * creal(e1) - strikes[i];
    free(v);
    free(kk);
    free(integ);
    return converged ? 0 : -4;
}

int orc_geo_price_weights(const double p[5], double S0, double r, const double* t, const double* w, int64_t n,
                          uint32_t n_strikes, const double* strikes, double* out) {
    return orc_geo_price_weights_tol(p, S0, r, t, w, n, n_strikes, strikes, 1e-15, out);
}

int orc_geo_asian_price(const double params[5], double S0, double r, const double* T, uint32_t n_exp,
                        const uint32_t* n_strikes, uint32_t steps, uint32_t flags, const double* strikes,
                        double* out) {
    const uint64_t cap = 4ull * steps * n_exp + 64; /* the time loop's own runaway bound */
    double* t = (double*)malloc(sizeof(double) * cap);
    double* w = (double*)malloc(sizeof(double) * cap);
    if (!t || !w) {
        free(t);
        free(w);
        return -3;
    }
    uint64_t off = 0;
    int rc = 0;
    for (uint32_t c = 0; c < n_exp && rc == 0; ++c) {
        if (n_strikes[c] == 0) continue;
        const int64_t n = orc_asian_weights(T, n_exp, n_strikes, steps, flags, c, t, w, cap);
        if (n < 0) {
            for (uint32_t i = 0; i < n_strikes[c]; ++i) out[off + i] = NAN;
        } else {
            const int e = orc_geo_price_weights(params, S0, r, t, w, n, n_strikes[c], strikes + off, out + off);
            if (e != 0 && e != -4) rc = e;
        }
        off += n_strikes[c];
    }
    free(t);
    free(w);
    return rc;
}
