/*
 * swift_oracle.c -- SWIFT European-call oracle (TEST INFRASTRUCTURE ONLY; see hexo_oracle.h).
 *
 * Restates the reference's Shannon-wavelet (SWIFT) pricer, src/SWIFT.cpp:18-118, and its
 * Heston characteristic function, src/HDistribution.cpp:8-38 and :86-113. The reference uses
 * FFTW (two transforms) and Eigen (one complex matrix product); neither is available here, so
 * both transforms are evaluated as direct DFTs over the few hundred non-zero inputs and the
 * product as a plain sum. Pinned by the 40 golden prices of src/UnitTest.cpp:221-261.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>

#include "hexo_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Heston characteristic function in the reference's parameterisation (Helpers, chf). */
static double complex chf(const double p[5], double tau, double complex u) {
    const double v0 = p[0], vm = p[1], rho = p[2], kappa = p[3], sigma = p[4];
    const double complex xi = kappa + sigma * rho * I * u;
    const double complex fac = u * u - I * u;
    const double complex d = csqrt(xi * xi + sigma * sigma * fac);
    const double complex sh = csinh(d * tau * 0.5), ch = ccosh(d * tau * 0.5);
    const double complex a1 = fac * sh;
    const double complex a2 = d / v0 * ch + xi / v0 * sh;
    const double complex a = a1 / a2;
    const double complex dd = clog((2.0 * d) / (d + xi + (d - xi) * cexp(-d * tau))) + (kappa - d) * tau * 0.5;
    return cexp((kappa * vm * rho * tau * u * I) / sigma - a + (2.0 * dd * kappa * vm) / (sigma * sigma));
}

/* ---- price gradients: the same characteristic function in forward-mode dual arithmetic ----
 * cd carries a value and its partial derivatives with respect to (v0, theta, rho, kappa, sigma);
 * every operation applies the chain rule, so chf_dual returns chf and its exact gradient (up to
 * rounding) without hand-derived formulas. The reference's analytic gradient
 * (src/HDistribution.cpp:47-84) is checked against this through its golden values. */
typedef struct {
    double complex v;
    double complex d[5];
} cd;

static cd cd_const(double complex v) {
    cd r;
    r.v = v;
    for (int k = 0; k < 5; ++k) r.d[k] = 0;
    return r;
}
static cd cd_var(double v, int k) {
    cd r = cd_const(v);
    r.d[k] = 1;
    return r;
}
static cd cd_add(cd a, cd b) {
    for (int k = 0; k < 5; ++k) a.d[k] += b.d[k];
    a.v += b.v;
    return a;
}
static cd cd_sub(cd a, cd b) {
    for (int k = 0; k < 5; ++k) a.d[k] -= b.d[k];
    a.v -= b.v;
    return a;
}
static cd cd_mul(cd a, cd b) {
    cd r;
    r.v = a.v * b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static cd cd_scale(cd a, double complex c) {
    a.v *= c;
    for (int k = 0; k < 5; ++k) a.d[k] *= c;
    return a;
}
static cd cd_div(cd a, cd b) {
    cd r;
    r.v = a.v / b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
    return r;
}
/* unary f with derivative f'(a.v) */
static cd cd_apply(cd a, double complex fv, double complex dfv) {
    cd r;
    r.v = fv;
    for (int k = 0; k < 5; ++k) r.d[k] = dfv * a.d[k];
    return r;
}
static cd cd_sqrt(cd a) {
    const double complex s = csqrt(a.v);
    return cd_apply(a, s, 0.5 / s);
}
static cd cd_exp(cd a) {
    const double complex e = cexp(a.v);
    return cd_apply(a, e, e);
}
static cd cd_log(cd a) { return cd_apply(a, clog(a.v), 1.0 / a.v); }
static cd cd_sinh(cd a) { return cd_apply(a, csinh(a.v), ccosh(a.v)); }
static cd cd_cosh(cd a) { return cd_apply(a, ccosh(a.v), csinh(a.v)); }

/* chf(u) and its gradient; the expression is chf() above, step for step */
static cd chf_dual(const double p[5], double tau, double u) {
    const cd v0 = cd_var(p[0], 0), vm = cd_var(p[1], 1), rho = cd_var(p[2], 2), kappa = cd_var(p[3], 3),
             sigma = cd_var(p[4], 4);
    const cd iu = cd_const(I * u);
    const cd xi = cd_add(kappa, cd_mul(cd_mul(sigma, rho), iu));
    const cd fac = cd_const(u * u - I * u);
    const cd d = cd_sqrt(cd_add(cd_mul(xi, xi), cd_mul(cd_mul(sigma, sigma), fac)));
    const cd half = cd_scale(d, tau * 0.5);
    const cd sh = cd_sinh(half), ch = cd_cosh(half);
    const cd a1 = cd_mul(fac, sh);
    const cd a2 = cd_add(cd_mul(cd_div(d, v0), ch), cd_mul(cd_div(xi, v0), sh));
    const cd a = cd_div(a1, a2);
    const cd num = cd_scale(d, 2.0);
    const cd den = cd_add(cd_add(d, xi), cd_mul(cd_sub(d, xi), cd_exp(cd_scale(d, -tau))));
    const cd dd = cd_add(cd_log(cd_div(num, den)), cd_scale(cd_sub(kappa, d), tau * 0.5));
    const cd t1 = cd_div(cd_scale(cd_mul(cd_mul(kappa, vm), rho), tau * u * I), sigma);
    const cd t3 = cd_div(cd_scale(cd_mul(cd_mul(dd, kappa), vm), 2.0), cd_mul(sigma, sigma));
    return cd_exp(cd_add(cd_sub(t1, a), t3));
}

void orc_heston_chf(const double params[5], double tau, double u, double out[2]) {
    double complex z = chf(params, tau, u);
    out[0] = creal(z);
    out[1] = cimag(z);
}

/* Cumulants used for the truncation range (HDistribution::first/second/fourth_order_moment). */
static double moment1(const double p[5], double tau) { return -0.5 * p[1] * tau; }
static double moment2(const double p[5], double tau) {
    const double s2 = p[1], r = p[2], kappa = p[3], k = p[4];
    const double kappa2 = kappa * kappa, kappa3 = kappa2 * kappa, k2 = k * k, t = tau;
    return s2 / (8 * kappa3) *
           (-k2 * exp(-2 * kappa * t) + 4 * k * exp(-kappa * t) * (k - 2 * kappa * r) +
            2 * kappa * t * (4 * kappa2 + k2 - 4 * kappa * k * r) + k * (8 * kappa * r - 3 * k));
}
static double moment4(const double p[5], double tau) {
    const double s2 = p[1], r = p[2], a = p[3], k = p[4];
    const double a2 = a * a, a3 = a2 * a, a4 = a3 * a;
    const double k2 = k * k, k3 = k2 * k, k4 = k3 * k;
    const double t = tau, t2 = t * t, r2 = r * r;
    return (3 * k2 * s2) / (64 * pow(a, 7)) *
           (-3 * k4 * exp(-4 * a * t) -
            8 * k2 * exp(-3 * a * t) * (2 * a * k * t * (k - 2 * a * r) + 4 * a2 + k2 - 6 * a * k * r) -
            4 * exp(-2 * a * t) *
                (4 * a2 * k2 * t2 * pow(k - 2 * a * r, 2) +
                 2 * a * k * t * (k3 - 16 * a3 * r - 12 * a * k2 * r + 4 * a2 * k * (3 + 4 * r2)) + 8 * a4 - 3 * k4 -
                 32 * a3 * k * r + 8 * a * k3 * r + 16 * a2 * k2 * r2) -
            8 * exp(-a * t) *
                (-2 * a2 * k * t2 * pow(k - 2 * a * r, 3) -
                 8 * a * t * (k4 - 7 * a * k3 * r + 4 * a4 * r2 - 8 * a3 * k * r * (1 + r2) + a2 * k2 * (3 + 14 * r2)) -
                 9 * k4 + 70 * a * k3 * r + 32 * a3 * k * r * (4 + 3 * r2) - 16 * a4 * (1 + 4 * r2) -
                 4 * a2 * k2 * (9 + 40 * r2)) +
            4 * a * t *
                (5 * k4 - 40 * a * k3 * r - 32 * a3 * k * r * (3 + 2 * r2) + 16 * a4 * (1 + 4 * r2) +
                 24 * a2 * k2 * (1 + 4 * r2)) -
            73 * k4 + 544 * a * k3 * r + 128 * a3 * k * r * (7 + 6 * r2) - 32 * a4 * (3 + 16 * r2) -
            64 * a2 * k2 * (4 + 19 * r2));
}

/* Truncation error bound for scale m (HDistribution::int_error). */
static double int_error(const double p[5], double tau, unsigned m) {
    const double e2m = exp2((double)m);
    return cabs(chf(p, tau, e2m * M_PI) + chf(p, tau, -e2m * M_PI)) / (4 * e2m * M_PI * M_PI * tau);
}

typedef struct {
    unsigned m, exp2_m;
    double sqrt_exp2_m, lower, upper;
    int k1, k2;
    unsigned J;
} swift_params;

static double u_of(const swift_params* sp, unsigned i) {
    return M_PI * (2 * (double)i + 1) / (2. * (double)sp->J) * (double)sp->exp2_m;
}

/* Dynamic parameters (SwiftParameters ctor, src/SWIFT.cpp:21-35). */
static void dynamic_params(const double p[5], double S, double r, double tau, double kmin, double kmax, double trunc,
                           swift_params* sp) {
    unsigned m = 0;
    while (int_error(p, tau, m++) > trunc && m < 30) {
    }
    sp->m = m;
    const double mx = r * tau + log(S / kmin);
    const double mn = r * tau + log(S / kmax);
    const double c = fabs(moment1(p, tau)) + 10. * sqrt(fabs(moment2(p, tau)) + sqrt(fabs(moment4(p, tau))));
    sp->exp2_m = (unsigned)exp2((double)m);
    sp->sqrt_exp2_m = sqrt((double)sp->exp2_m);
    sp->lower = mn - c;
    sp->upper = mx + c;
    sp->k1 = (int)ceil(sp->exp2_m * sp->lower);
    sp->k2 = (int)floor(sp->exp2_m * sp->upper);
    const double iota = ceil(log2(M_PI * abs(sp->k1 - sp->k2))) - 1;
    sp->J = (unsigned)exp2(iota - 1);
}

static inline int pmod(long a, long n) { return (int)(((a % n) + n) % n); }

int orc_swift_price_grad(const double params[5], double S, double r, double T, uint32_t n_strikes,
                         const double* strikes, const double* swift_in, double trunc, double* prices, double* grads,
                         double* used) {
    if (n_strikes == 0 || !(T > 0) || !(S > 0)) return -1;
    swift_params sp;
    if (swift_in) {
        sp.m = (unsigned)swift_in[0];
        sp.exp2_m = (unsigned)swift_in[1];
        sp.sqrt_exp2_m = swift_in[2];
        sp.lower = swift_in[3];
        sp.upper = swift_in[4];
        sp.k1 = (int)swift_in[5];
        sp.k2 = (int)swift_in[6];
        sp.J = (unsigned)swift_in[7];
    } else {
        double kmin = strikes[0], kmax = strikes[0];
        for (uint32_t i = 1; i < n_strikes; ++i) {
            if (strikes[i] < kmin) kmin = strikes[i];
            if (strikes[i] > kmax) kmax = strikes[i];
        }
        dynamic_params(params, S, r, T, kmin, kmax, trunc, &sp);
    }
    if (used) {
        used[0] = sp.m;
        used[1] = sp.exp2_m;
        used[2] = sp.sqrt_exp2_m;
        used[3] = sp.lower;
        used[4] = sp.upper;
        used[5] = sp.k1;
        used[6] = sp.k2;
        used[7] = sp.J;
    }
    const unsigned J = sp.J;
    if (J == 0 || J > (1u << 20)) return -2;
    const long n2 = 2L * J, n4 = 4L * J;
    double complex* pay_in = (double complex*)calloc((size_t)J, sizeof(double complex));
    double complex* tw2 = (double complex*)malloc(sizeof(double complex) * (size_t)n2);
    double complex* tw4 = (double complex*)malloc(sizeof(double complex) * (size_t)n4);
    double* dens_in = (double*)calloc((size_t)n4, sizeof(double));
    double complex* coeff = (double complex*)malloc(sizeof(double complex) * J);
    if (!pay_in || !tw2 || !tw4 || !dens_in || !coeff) return -3;
    for (long k = 0; k < n2; ++k) tw2[k] = cexp(2.0 * M_PI * I * (double)k / (double)n2);
    for (long k = 0; k < n4; ++k) tw4[k] = cexp(-2.0 * M_PI * I * (double)k / (double)n4);

    /* payoff coefficients of the call's integrand on [max(lower,0), upper] (src/SWIFT.cpp:41-58) */
    const double ylo = sp.lower > 0. ? sp.lower : 0.;
    const double eu = exp(sp.upper), el = exp(ylo);
    for (unsigned j = 0; j < J; ++j) {
        const double u = u_of(&sp, j);
        const double complex hu = -I * cexp(-I * u * sp.upper) * (1 / u - eu / (I + u));
        const double complex hl = -I * cexp(-I * u * ylo) * (1 / u - el / (I + u));
        pay_in[j] = hu - hl;
    }
    /* backward DFT of length 2J (zero-padded), sampled at i mod 2J, then rotated and scaled
     * into the real input of the length-4J forward transform (src/SWIFT.cpp:63-71) */
    for (int i = sp.k1; i <= sp.k2; ++i) {
        const int i2 = pmod(i, n2), i4 = pmod(i, n4);
        double complex acc = 0;
        for (unsigned j = 0; j < J; ++j) acc += pay_in[j] * tw2[((long)j * i2) % n2];
        const double complex rot = cexp(I * (double)i * M_PI / (double)(2 * J));
        dens_in[i4] = creal(rot * acc) * sp.sqrt_exp2_m / (double)J;
    }
    /* forward real DFT of length 4J at the odd bins, conjugated (src/SWIFT.cpp:71-72) */
    for (unsigned j = 0; j < J; ++j) {
        const long k = 2L * j + 1;
        double complex acc = 0;
        for (int i = sp.k1; i <= sp.k2; ++i) {
            const int i4 = pmod(i, n4);
            acc += dens_in[i4] * tw4[((long)i4 * k) % n4];
        }
        coeff[j] = conj(acc) * sp.sqrt_exp2_m / (double)J;
    }
    /* price = Re sum_j chf(u_j) c_j e^{-rT} K e^{-i u_j x}, x = rT + log(S/K) (src/SWIFT.cpp:102-117) */
    /* with grads: rows 1..5 are the partials of chf, as in the reference's 6 x J pricing
     * matrix (src/SWIFT.cpp:106-110); the strike factors are the same for every row */
    const double disc = exp(-r * T);
    const int rows = grads ? 6 : 1;
    double complex* pm = (double complex*)malloc(sizeof(double complex) * J * (size_t)rows);
    if (!pm) return -3;
    for (unsigned j = 0; j < J; ++j) {
        if (grads) {
            const cd c = chf_dual(params, T, u_of(&sp, j));
            pm[j] = c.v * coeff[j];
            for (int k = 0; k < 5; ++k) pm[(size_t)(k + 1) * J + j] = c.d[k] * coeff[j];
        } else {
            pm[j] = chf(params, T, u_of(&sp, j)) * coeff[j];
        }
    }
    for (uint32_t o = 0; o < n_strikes; ++o) {
        const double K = strikes[o];
        const double x = r * T + log(S / K);
        double complex acc[6] = {0, 0, 0, 0, 0, 0};
        for (unsigned j = 0; j < J; ++j) {
            const double complex e = disc * K * cexp(-I * u_of(&sp, j) * x);
            for (int row = 0; row < rows; ++row) acc[row] += pm[(size_t)row * J + j] * e;
        }
        prices[o] = creal(acc[0]);
        if (grads)
            for (int k = 0; k < 5; ++k) grads[(size_t)o * 5 + k] = creal(acc[k + 1]);
    }
    free(pm);
    free(pay_in);
    free(tw2);
    free(tw4);
    free(dens_in);
    free(coeff);
    return 0;
}

int orc_swift_price(const double params[5], double S, double r, double T, uint32_t n_strikes, const double* strikes,
                    const double* swift_in, double trunc, double* prices, double* used) {
    return orc_swift_price_grad(params, S, r, T, n_strikes, strikes, swift_in, trunc, prices, NULL, used);
}
