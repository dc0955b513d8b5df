/*
 * hexo_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see hexo_oracle.h).
 *
 * A plain-C restatement of the reference's Heston QE Monte-Carlo path, used as the checker
 * for the HIP product and as the "port" CPU baseline. It keeps the reference's control flow:
 * the per-path loop replays the double-precision time accumulation of
 * src/HSimulation.tpp:35-44 literally, rather than using a precomputed schedule, so the
 * product's host schedule builder is checked against it.
 */
#include "hexo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11).
 * Round: (hi0,lo0) = M0*c0, (hi1,lo1) = M1*c2, c' = (hi1^c1^k0, lo1, hi0^c3^k1, lo0);
 * key bumped by the Weyl constants between rounds.
 * ------------------------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static inline void stream_block(uint64_t seed, uint32_t stream, uint64_t path, uint64_t block, uint32_t out[4]) {
    uint32_t ctr[4] = {(uint32_t)block, (uint32_t)path, (uint32_t)(path >> 32), stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, out);
}

void orc_rng_words(uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths, uint64_t word0,
                   uint64_t n_words, uint32_t* out) {
    for (uint64_t i = 0; i < n_paths; ++i) {
        uint32_t blk[4];
        uint64_t have = ~(uint64_t)0;
        for (uint64_t j = 0; j < n_words; ++j) {
            uint64_t w = word0 + j;
            if ((w >> 2) != have) {
                have = w >> 2;
                stream_block(seed, stream, path0 + i, have, blk);
            }
            out[i * n_words + j] = blk[w & 3];
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Uniforms. u = (k + 1/2) 2^-32 lies strictly inside (0,1) and is symmetric: 1 - u(k) = u(~k).
 * ------------------------------------------------------------------------------------------ */
double orc_uniform_from_word(uint32_t k) { return ((double)k + 0.5) * 0x1p-32; }

/* ------------------------------------------------------------------------------------------
 * AS241 PPND16 (Wichura 1988, Appl. Statist. 37(3)); coefficients as published and as
 * listed at src/as241.f90:31-44 (A,B), :49-63 (C,D), :68-82 (E,F).
 * ------------------------------------------------------------------------------------------ */
static const double AS_A[8] = {3.3871328727963666080e0, 1.3314166789178437745e+2, 1.9715909503065514427e+3,
                               1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4,
                               3.3430575583588128105e+4, 2.5090809287301226727e+3};
static const double AS_B[8] = {1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3,
                               2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4,
                               5.2264952788528545610e+3};
static const double AS_C[8] = {1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0,
                               3.64784832476320460504e0, 1.27045825245236838258e0, 2.41780725177450611770e-1,
                               2.27238449892691845833e-2, 7.74545014278341407640e-4};
static const double AS_D[8] = {1.0, 2.05319162663775882187e0, 1.67638483018380384940e0, 6.89767334985100004550e-1,
                               1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4,
                               1.05075007164441684324e-9};
static const double AS_E[8] = {6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0,
                               2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3,
                               2.71155556874348757815e-5, 2.01033439929228813265e-7};
static const double AS_F[8] = {1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2,
                               7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7,
                               2.04426310338993978564e-15};

void orc_as241_coefficients(double out[45]) {
    int n = 0;
    for (int i = 0; i < 8; ++i) out[n++] = AS_A[i];
    for (int i = 1; i < 8; ++i) out[n++] = AS_B[i];
    for (int i = 0; i < 8; ++i) out[n++] = AS_C[i];
    for (int i = 1; i < 8; ++i) out[n++] = AS_D[i];
    for (int i = 0; i < 8; ++i) out[n++] = AS_E[i];
    for (int i = 1; i < 8; ++i) out[n++] = AS_F[i];
}

/* Horner from the highest coefficient down: ((c7 r + c6) r + ...) r + c0. */
static inline double horner8(const double* c, double r) {
    double acc = c[7];
    for (int i = 6; i >= 0; --i) acc = acc * r + c[i];
    return acc;
}
static inline float horner8f(const float* c, float r) {
    float acc = c[7];
    for (int i = 6; i >= 0; --i) acc = acc * r + c[i];
    return acc;
}

/* AS241 evaluated entirely in double (the build's FP64 mode; src/as241.f90:85-118 logic). */
double orc_ppnd16_f64(double p, int* ifault) {
    *ifault = 0;
    double q = p - 0.5;
    if (fabs(q) <= 0.425) {
        double r = 0.180625 - q * q;
        return q * horner8(AS_A, r) / horner8(AS_B, r);
    }
    double r = (q < 0.0) ? p : 1.0 - p;
    if (r <= 0.0) {
        *ifault = 1;
        return 0.0;
    }
    r = sqrt(-log(r));
    double z;
    if (r <= 5.0) {
        r -= 1.6;
        z = horner8(AS_C, r) / horner8(AS_D, r);
    } else {
        r -= 5.0;
        z = horner8(AS_E, r) / horner8(AS_F, r);
    }
    return (q < 0.0) ? -z : z;
}

/* AS241 with the reference's arithmetic: every coefficient and the locals Q, R are default
 * REAL (single precision) in src/as241.f90:20-27, so the evaluation is FP32 behind a double
 * interface. P - HALF and ONE - P are formed in double and rounded on assignment. */
double orc_ppnd16_asbuilt(double p, int* ifault) {
    float fa[8], fb[8], fc[8], fd[8], fe[8], ff[8];
    for (int i = 0; i < 8; ++i) {
        fa[i] = (float)AS_A[i];
        fb[i] = (float)AS_B[i];
        fc[i] = (float)AS_C[i];
        fd[i] = (float)AS_D[i];
        fe[i] = (float)AS_E[i];
        ff[i] = (float)AS_F[i];
    }
    *ifault = 0;
    volatile float q = (float)(p - 0.5);
    if (fabsf(q) <= (float)0.425) {
        volatile float qq = q * q;
        volatile float r = (float)0.180625 - qq;
        volatile float num = q * horner8f(fa, r);
        volatile float z = num / horner8f(fb, r);
        return (double)z;
    }
    volatile float r = (q < 0.0f) ? (float)p : (float)(1.0 - p);
    if (r <= 0.0f) {
        *ifault = 1;
        return 0.0;
    }
    r = sqrtf(-logf(r));
    volatile float z;
    if (r <= 5.0f) {
        r = r - (float)1.6;
        z = horner8f(fc, r) / horner8f(fd, r);
    } else {
        r = r - 5.0f;
        z = horner8f(fe, r) / horner8f(ff, r);
    }
    if (q < 0.0f) z = -z;
    return (double)z;
}

double orc_normal_f64_from_word(uint32_t k) {
    int f;
    return orc_ppnd16_f64(orc_uniform_from_word(k), &f);
}

/* The FP32 kernel's normal (hestonexotics_amd/csrc/device.hpp, normal_word_f32), restated with
 * libm's log2f and sqrtf where the kernel uses the hardware's: z = Phi^-1(u), u = (k + 1/2) 2^-32.
 * The word is folded onto the lower half, j = min(k, ~k), so s = (j + 1/2) 2^-32 = min(u, 1 - u)
 * keeps full relative precision in both tails; with t = s (1 - s) and l = log2 t, |z| = (1 - 2s) h(l)
 * where h is a degree-9 polynomial in l for l >= -11 and a degree-8 polynomial in
 * sqrt(-l) - 4.5305938720703125 below (Giles' erfinv variable, M. Giles, "Approximating the erfinv
 * function", GPU Computing Gems Jade Edition, 2011, in base 2). The coefficients are the product's
 * own least-squares fits; tests/test_oracle_normal_f32.py pins this function against the FP64
 * PPND16 and scipy's ndtri. */
static const float NF32_C[10] = {0.8369221091270447f,     -0.18630073964595795f,   0.011675846762955189f,
                                 1.11724120870349e-05f,   -0.0002003998088184744f, -9.236554433300626e-06f,
                                 1.7887376770886476e-06f, 2.4262888587145426e-07f, 1.1781071229677309e-08f,
                                 2.1350940460074241e-10f};
static const float NF32_D[9] = {4.8360700607299805f,     1.2490957975387573f,      -0.011836512014269829f,
                                0.0019035495351999998f,  -0.00013401612523011863f, -0.00016207653970923275f,
                                0.00019050169794354588f, -0.00012945415801368654f, 3.693158214446157e-05f};

static inline float horner_fmaf(const float* c, int n, float x) {
    float acc = c[n - 1];
    for (int i = n - 2; i >= 0; --i) acc = fmaf(acc, x, c[i]);
    return acc;
}

float orc_normal_f32_from_word(uint32_t k) {
    const uint32_t j = (k & 0x80000000u) ? ~k : k;
    const float s = fmaf((float)j, 0x1p-32f, 0x1p-33f);
    const float t = fmaf(-s, s, s);
    const float l = log2f(t);
    const float h = (l < -11.0f) ? horner_fmaf(NF32_D, 9, sqrtf(-l) - 4.5305938720703125f)
                                 : horner_fmaf(NF32_C, 10, l);
    const float pr = h * fmaf(-2.0f, s, 1.0f);
    return (k & 0x80000000u) ? pr : -pr;
}

/* ------------------------------------------------------------------------------------------
 * One QE step, FP64, in the reference's form (src/HSimulation.tpp:52-86): Andersen's
 * quadratic-exponential variance step with switch at PSI_C = 1.5 (src/inc/HSimulation.h:12)
 * and the central (gamma1 = gamma2 = 1/2) log-price step. The V-draw uses word w_v (as a
 * normal in the quadratic branch, as a uniform in the exponential branch); Z_X uses w_x.
 * r_drift is added as r*delta (0 reproduces the reference). With martingale set, K0 is
 * Andersen's martingale-corrected K0* (L. Andersen, "Simple and efficient simulation of the
 * Heston stochastic volatility model", J. Comput. Finance 11(3), 2008; an extension the
 * reference suggests at src/inc/HSimulation.h:48-60 but does not implement).
 * ------------------------------------------------------------------------------------------ */
int orc_qe_step(const double params[5], double r_drift, double delta, int martingale, double* Vio, double* logXio,
                uint32_t w_v, uint32_t w_x) {
    const double theta = params[1], rho = params[2], kappa = params[3], eps = params[4];
    const double V = *Vio;
    const double e = exp(-kappa * delta);
    const double m = theta + (V - theta) * e;
    const double s2 = fabs(V * eps * eps * e / kappa * (1.0 - e) + theta * eps * eps / (2.0 * kappa) * (1.0 - e) * (1.0 - e));
    const double psi = s2 / (m * m);
    double Vn, a = 0.0, b2 = 0.0, p = 0.0, beta = 0.0;
    int quad;
    if (psi < 1.5) {
        const double two_over_psi = 2.0 / psi;
        b2 = two_over_psi - 1.0 + sqrt(two_over_psi * (two_over_psi - 1.0));
        const double b = sqrt(b2);
        a = m / (1.0 + b2);
        const double zv = orc_normal_f64_from_word(w_v);
        Vn = a * (b + zv) * (b + zv);
        quad = 1;
    } else {
        p = (psi - 1.0) / (psi + 1.0);
        beta = 2.0 / (m * (psi + 1.0));
        const double u = orc_uniform_from_word(w_v);
        Vn = (p < u) ? log((1.0 - p) / (1.0 - u)) / beta : 0.0;
        quad = 0;
    }
    double k0 = -rho * kappa * theta / eps * delta + r_drift * delta;
    const double k1 = 0.5 * delta * (kappa * rho / eps - 0.5) - rho / eps;
    const double k2 = 0.5 * delta * (kappa * rho / eps - 0.5) + rho / eps;
    const double k3 = 0.5 * delta * (1.0 - rho * rho);
    const double k4 = k3;
    if (martingale) {
        /* Andersen's martingale correction: K0* = -ln M - (K1 + K3/2) V with M = E[exp(A V') | V],
         * A = K2 + K4/2. Quadratic branch: M = exp(A b^2 a / (1 - 2 A a)) / sqrt(1 - 2 A a) for
         * 2 A a < 1; exponential branch: M = p + beta (1 - p) / (beta - A) for A < beta. */
        const double A = k2 + 0.5 * k4;
        double lnM = 0.0;
        int ok = 0;
        if (quad && 2.0 * A * a < 1.0) {
            lnM = A * b2 * a / (1.0 - 2.0 * A * a) - 0.5 * log1p(-2.0 * A * a);
            ok = 1;
        } else if (!quad && A < beta) {
            lnM = log(p + beta * (1.0 - p) / (beta - A));
            ok = 1;
        }
        if (ok) k0 = -lnM - (k1 + 0.5 * k3) * V + r_drift * delta;
    }
    const double zx = orc_normal_f64_from_word(w_x);
    *logXio = *logXio + k0 + k1 * V + k2 * Vn + sqrt(k3 * V + k4 * Vn) * zx;
    *Vio = Vn;
    return quad;
}

/* Conditional mean E[X'/X | V] of one step (r_drift included) by the product midpoint rule over
 * the step's two uniforms: 2^log2_nv x 2^log2_nx words, each the midpoint of its interval. A
 * test helper for the martingale correction, under which the mean is exp(r_drift delta). */
double orc_qe_step_mean(const double params[5], double r_drift, double delta, int martingale, double V,
                        uint32_t log2_nv, uint32_t log2_nx) {
    if (log2_nv < 1 || log2_nv > 24 || log2_nx < 1 || log2_nx > 24) return NAN;
    const uint32_t nv = 1u << log2_nv, nx = 1u << log2_nx, sv = 32 - log2_nv, sx = 32 - log2_nx;
    double total = 0.0;
#pragma omp parallel for reduction(+ : total) schedule(static)
    for (int64_t i = 0; i < (int64_t)nv; ++i) {
        const uint32_t wv = ((uint32_t)i << sv) | (1u << (sv - 1));
        double row = 0.0;
        for (uint32_t j = 0; j < nx; ++j) {
            const uint32_t wx = (j << sx) | (1u << (sx - 1));
            double v = V, lx = 0.0;
            orc_qe_step(params, r_drift, delta, martingale, &v, &lx, wv, wx);
            row += exp(lx);
        }
        total += row / nx;
    }
    return total / nv;
}

/* ------------------------------------------------------------------------------------------
 * Reference time grid (src/HSimulation.tpp:30-45, src/inc/AsianContract.h:35-38).
 * ------------------------------------------------------------------------------------------ */
int64_t orc_time_grid(const double* T, uint32_t n_exp, const uint32_t* n_strikes, uint32_t steps,
                      uint64_t* fin_step, double* fin_w) {
    if (n_exp == 0 || steps == 0) return -1;
    uint64_t n_opts = 0;
    for (uint32_t c = 0; c < n_exp; ++c) n_opts += n_strikes[c];
    if (n_opts == 0) return -1;
    double t = 0.0, tp = 0.0, dt = T[0] / steps;
    uint32_t c = 0;
    uint64_t priced = 0;
    int64_t k = 0;
    tp = t;
    t += dt;
    ++k;
    while (priced < n_opts) {
        while (t >= T[c] && priced < n_opts) {
            fin_step[c] = (uint64_t)k;
            fin_w[c] = (T[c] - tp) / dt;
            priced += n_strikes[c];
            if (priced < n_opts) {
                ++c;
                dt = T[c] / steps;
            }
        }
        if (priced >= n_opts) break;
        tp = t;
        t += dt;
        ++k;
        if (k > (int64_t)4 * steps * (int64_t)n_exp + 16) return -1; /* runaway guard */
    }
    for (uint32_t cc = c + 1; cc < n_exp; ++cc) {
        fin_step[cc] = 0;
        fin_w[cc] = 0.0;
    }
    return k;
}

/* ------------------------------------------------------------------------------------------
 * Monte-Carlo price.
 * ------------------------------------------------------------------------------------------ */
static double g_last_seconds = 0.0;
double orc_last_seconds(void) { return g_last_seconds; }
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Neumaier-compensated accumulator (so the checker's sums are accurate to ~1 ulp). */
typedef struct {
    double s, c;
} acc2;
static inline void acc_add(acc2* a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x))
        a->c += (a->s - t) + x;
    else
        a->c += (x - t) + a->s;
    a->s = t;
}
static inline double acc_val(const acc2* a) { return a->s + a->c; }

#define ORC_BLOCK 4096 /* paths per partial block; partials are combined in block order */

/* Simulate one path; adds payoffs into sums[o] and sumsq[o] (payoff-major outputs). */
static void simulate_path(const double params[5], double S0, double r, uint32_t flags, uint32_t payoffs,
                          uint64_t seed, uint64_t path, uint32_t steps, uint32_t n_exp, const double* T,
                          const uint32_t* n_strikes, const uint32_t* opt_off, const double* strikes, uint64_t n_opts,
                          acc2* sums, acc2* sumsq, double* finals, uint64_t* counts) {
    const double r_drift = (flags & ORC_DRIFT) ? r : 0.0;
    const int mart = (flags & ORC_MARTINGALE) != 0;
    /* state as in SDE_state<2> + HQEAnderson::log_X (src/inc/SDE.h:9-14, src/HSimulation.tpp:87-94) */
    double X = S0, Xp = S0, V = params[0], logX = log(S0);
    double t = 0.0, tp = 0.0;
    double acc = 0.0;          /* AAsianCallNonAdaptive::accumulated_value */
    double acc_y = 0.0, acc_w = 0.0; /* the same accumulation on ln(X/S0), and its weight sum */
    const double logS0 = log(S0);
    double dt = T[0] / steps;  /* update_earliest(T_0, steps) */
    uint32_t c = 0;
    uint64_t priced = 0;
    uint64_t k = 0; /* steps taken */
    uint32_t blk[4];
    uint64_t have = ~(uint64_t)0;
    const int want_asian = (payoffs & ORC_ASIAN) != 0, want_euro = (payoffs & ORC_EUROPEAN) != 0;
    const int want_cv = (payoffs & ORC_ASIAN_GEO_CV) != 0, want_geo = (payoffs & ORC_GEO_ASIAN) != 0;
    double logXp = logX;
    uint64_t nq = 0, ne = 0;

#define ORC_STEP()                                                                             \
    do {                                                                                       \
        uint64_t w = 2 * k;                                                                    \
        if ((w >> 2) != have) {                                                                \
            have = w >> 2;                                                                     \
            stream_block(seed, 0u, path, have, blk);                                           \
        }                                                                                      \
        uint32_t wv = blk[w & 3], wx = blk[(w + 1) & 3];                                       \
        const double logXp_next = logX;                                                        \
        int qd = orc_qe_step(params, r_drift, dt, mart, &V, &logX, wv, wx);                    \
        nq += qd;                                                                              \
        ne += 1 - qd;                                                                          \
        logXp = logXp_next;                                                                    \
        Xp = X;                                                                                \
        X = exp(logX);                                                                         \
        tp = t;                                                                                \
        t += dt;                                                                               \
        ++k;                                                                                   \
    } while (0)

    ORC_STEP();
    while (priced < n_opts) {
        while (t >= T[c] && priced < n_opts) {
            /* accumulate_final_value of both policies (src/inc/AsianContract.h:29-34,
             * src/inc/VanillaContract.h:28-31) */
            const double xt = Xp + (X - Xp) * (T[c] - tp) / dt;
            double avg;
            if (flags & ORC_TRAPEZOID)
                avg = (acc + (T[c] - tp) * 0.5 * (Xp + xt)) / T[c];
            else
                avg = (acc + (X - Xp) * (T[c] - tp) / dt) / T[c];
            if (finals) {
                finals[(size_t)c * 2 + 0] = avg;
                finals[(size_t)c * 2 + 1] = xt;
            }
            /* geometric counterpart: the same functional on y = ln(X/S0), normalised by its
             * weight sum W; G = W S0 exp(L / W) */
            double geo = 0.0;
            if (want_cv || want_geo) {
                const double y = logX - logS0, yp = logXp - logS0, wf = (T[c] - tp) / dt;
                double L, W;
                if (flags & ORC_TRAPEZOID) {
                    L = (acc_y + (T[c] - tp) * 0.5 * (yp + (yp + (y - yp) * wf))) / T[c];
                    W = (acc_w + (T[c] - tp)) / T[c];
                } else {
                    L = (acc_y + (y - yp) * wf) / T[c];
                    W = acc_w / T[c];
                }
                geo = W * S0 * exp(L / W);
            }
            for (uint32_t j = 0; j < n_strikes[c]; ++j) {
                const double K = strikes[opt_off[c] + j];
                uint64_t o = opt_off[c] + j;
                if (want_asian) {
                    double pay = avg - K > 0.0 ? avg - K : 0.0;
                    acc_add(&sums[o], pay);
                    acc_add(&sumsq[o], pay * pay);
                    o += n_opts;
                }
                if (want_euro) {
                    double pay = xt - K > 0.0 ? xt - K : 0.0;
                    acc_add(&sums[o], pay);
                    acc_add(&sumsq[o], pay * pay);
                    o += n_opts;
                }
                const double pg = geo - K > 0.0 ? geo - K : 0.0;
                if (want_cv) {
                    const double pa = avg - K > 0.0 ? avg - K : 0.0;
                    acc_add(&sums[o], pa - pg);
                    acc_add(&sumsq[o], (pa - pg) * (pa - pg));
                    o += n_opts;
                }
                if (want_geo) {
                    acc_add(&sums[o], pg);
                    acc_add(&sumsq[o], pg * pg);
                }
            }
            priced += n_strikes[c];
            if (priced < n_opts) {
                ++c;
                dt = T[c] / steps; /* update_earliest */
            }
        }
        if (priced >= n_opts) break; /* the reference's one extra accumulate+step has no effect */
        acc += dt * 0.5 * (X + Xp); /* accumulate_value */
        acc_y += dt * 0.5 * ((logX - logS0) + (logXp - logS0));
        acc_w += dt;
        ORC_STEP();
    }
#undef ORC_STEP
    if (finals)
        for (uint32_t cc = c + 1; cc < n_exp; ++cc) finals[(size_t)cc * 2] = finals[(size_t)cc * 2 + 1] = NAN;
    if (counts) {
        counts[0] += nq;
        counts[1] += ne;
    }
}

int orc_price(const double params[5], double S0, double r, uint32_t flags, uint32_t payoffs, uint64_t n_paths,
              uint64_t path_offset, uint64_t seed, uint32_t steps, uint32_t n_exp, const double* T,
              const uint32_t* n_strikes, const double* strikes, double* prices, double* std_errs, double* finals,
              uint64_t* counts, int n_threads) {
    if (n_paths == 0 || steps == 0 || n_exp == 0 || !(payoffs & 15u) || (payoffs & ~15u)) return -1;
    for (uint32_t c = 0; c < n_exp; ++c)
        if (!(T[c] > 0.0) || (c > 0 && !(T[c] > T[c - 1]))) return -2;
    uint32_t* opt_off = (uint32_t*)malloc(sizeof(uint32_t) * (n_exp + 1));
    uint64_t n_opts = 0;
    for (uint32_t c = 0; c < n_exp; ++c) {
        opt_off[c] = (uint32_t)n_opts;
        n_opts += n_strikes[c];
    }
    opt_off[n_exp] = (uint32_t)n_opts;
    if (n_opts == 0) {
        free(opt_off);
        return -1;
    }
    const int n_pay = ((payoffs & ORC_ASIAN) ? 1 : 0) + ((payoffs & ORC_EUROPEAN) ? 1 : 0) +
                      ((payoffs & ORC_ASIAN_GEO_CV) ? 1 : 0) + ((payoffs & ORC_GEO_ASIAN) ? 1 : 0);
    /* block of the control-variate outputs: after the Asian and European blocks */
    const uint64_t cv_block = ((payoffs & ORC_ASIAN) ? 1 : 0) + ((payoffs & ORC_EUROPEAN) ? 1 : 0);
    double* geo_exact = NULL;
    if (payoffs & ORC_ASIAN_GEO_CV) {
        geo_exact = (double*)malloc(sizeof(double) * n_opts);
        if (!geo_exact ||
            orc_geo_asian_price(params, S0, (flags & ORC_DRIFT) ? r : 0.0, T, n_exp, n_strikes, steps, flags, strikes,
                                geo_exact) < 0) {
            free(geo_exact);
            free(opt_off);
            return -3;
        }
    }
    const uint64_t n_out = n_opts * (uint64_t)n_pay;
    const uint64_t n_blocks = (n_paths + ORC_BLOCK - 1) / ORC_BLOCK;
    acc2* part = (acc2*)calloc((size_t)(n_blocks * n_out * 2), sizeof(acc2));
    uint64_t cq = 0, ce = 0;
    if (!part) {
        free(geo_exact);
        free(opt_off);
        return -3;
    }
#ifdef _OPENMP
    if (n_threads <= 0) n_threads = omp_get_max_threads();
#else
    n_threads = 1;
#endif
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(+ : cq, ce)
#endif
    for (int64_t b = 0; b < (int64_t)n_blocks; ++b) {
        acc2* sums = part + (size_t)b * n_out * 2;
        acc2* sumsq = sums + n_out;
        uint64_t lo = (uint64_t)b * ORC_BLOCK, hi = lo + ORC_BLOCK < n_paths ? lo + ORC_BLOCK : n_paths;
        uint64_t cnt[2] = {0, 0};
        for (uint64_t i = lo; i < hi; ++i)
            simulate_path(params, S0, r, flags, payoffs, seed, path_offset + i, steps, n_exp, T, n_strikes, opt_off,
                          strikes, n_opts, sums, sumsq, finals ? finals + (size_t)i * n_exp * 2 : NULL, cnt);
        cq += cnt[0];
        ce += cnt[1];
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    g_last_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (counts) {
        counts[0] = cq;
        counts[1] = ce;
    }
    for (uint64_t o = 0; o < n_out; ++o) {
        acc2 s = {0, 0}, ss = {0, 0};
        for (uint64_t b = 0; b < n_blocks; ++b) {
            const acc2* sums = part + (size_t)b * n_out * 2;
            acc_add(&s, acc_val(&sums[o]));
            acc_add(&ss, acc_val(&sums[n_out + o]));
        }
        const uint64_t opt = o % n_opts;
        uint32_t c = 0;
        while (opt >= opt_off[c + 1]) ++c;
        const double disc = (flags & ORC_DISCOUNT) ? exp(-r * T[c]) : 1.0;
        const double n = (double)n_paths;
        const double mean = acc_val(&s) / n;
        double var = n > 1 ? (acc_val(&ss) - n * mean * mean) / (n - 1.0) : 0.0;
        if (var < 0) var = 0;
        const double exact = (geo_exact && o / n_opts == cv_block) ? geo_exact[opt] : 0.0;
        prices[o] = disc * (mean + exact);
        if (std_errs) std_errs[o] = disc * sqrt(var / n);
    }
    free(geo_exact);
    free(part);
    free(opt_off);
    return 0;
}

/* ------------------------------------------------------------------------------------------
 * Moments of one QE variance step from state V over n independent draws (draw i uses path i of
 * the stream, words 0 and 1). out = {mean, variance, 4th central moment, fraction V' = 0,
 * fraction of quadratic-branch draws, mean of exp(log-price increment)}.
 * ------------------------------------------------------------------------------------------ */
void orc_qe_moments(const double params[5], double delta, double V, uint64_t n, uint64_t seed, double out[6]) {
    double s1 = 0, s2 = 0, s3 = 0, s4 = 0, zeros = 0, quads = 0, ex = 0;
    double vals_mean = 0.0;
    /* two passes for a stable central moment: first the mean */
    for (int pass = 0; pass < 2; ++pass) {
        s1 = s2 = s3 = s4 = zeros = quads = ex = 0;
        for (uint64_t i = 0; i < n; ++i) {
            uint32_t blk[4];
            stream_block(seed, 0u, i, 0, blk);
            double v = V, lx = 0.0;
            int q = orc_qe_step(params, 0.0, delta, 0, &v, &lx, blk[0], blk[1]);
            if (pass == 0) {
                s1 += v;
                continue;
            }
            const double d = v - vals_mean;
            s2 += d * d;
            s4 += d * d * d * d;
            s3 += v;
            zeros += v == 0.0;
            quads += q;
            ex += exp(lx);
        }
        if (pass == 0) vals_mean = s1 / (double)n;
    }
    out[0] = vals_mean;
    out[1] = s2 / (double)(n - 1);
    out[2] = s4 / (double)n;
    out[3] = zeros / (double)n;
    out[4] = quads / (double)n;
    out[5] = ex / (double)n;
}
