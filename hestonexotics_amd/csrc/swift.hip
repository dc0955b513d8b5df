// swift.hip -- SWIFT European pricer on gfx950, FP64 throughout.
//
// The Shannon-wavelet (SWIFT) method of the reference (src/SWIFT.cpp:18-118): payoff
// coefficients of the call on the truncated log-price range, two Fourier transforms to the
// density coefficients c_j, then price = Re sum_j chf(u_j) c_j e^{-rT} K e^{-i u_j x} with
// x = rT + log(S/K) (src/SWIFT.cpp:102-117). The reference uses FFTW for the transforms and
// Eigen for the final 6 x J by J x K product; here each stage is one kernel over all chains of
// the call: the transforms are direct sums with one wavefront per output (J and k2 - k1 are a few
// hundred to a few thousand, and every output is independent), the characteristic function is
// evaluated per (chain, j), and each (chain, strike) is one wavefront that sums over j; every
// wavefront sum is lane-strided in order and combined by a fixed xor butterfly.
//
// Gradients (SWIFT::price_opts_grad, the calibration Jacobian) come from the same
// characteristic-function expression evaluated in forward-mode dual arithmetic carrying all five
// parameter directions, instead of hand-derived partials (src/HDistribution.cpp:47-84).
#include "swift.hpp"
#include "swift_grid.hpp"
#include "cplx_device.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

namespace hexo {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Plain complex versions of the elementary functions, so that the characteristic-function
// expression below instantiates for Cplx (values) and for Dual5 (values and partials).
__device__ __forceinline__ Cplx dsqrt(Cplx a) { return csqrt(a); }
__device__ __forceinline__ Cplx dexp(Cplx a) { return cexp(a); }
__device__ __forceinline__ Cplx dlog(Cplx a) { return clog(a); }
__device__ __forceinline__ Cplx dsinh(Cplx a) { return csinh(a); }
__device__ __forceinline__ Cplx dcosh(Cplx a) { return ccosh(a); }

// Value and the five partial derivatives (forward-mode, all parameter directions at once).
struct Dual5 {
    Cplx v, d[5];
};
__device__ __forceinline__ Dual5 d5const(Cplx v) {
    Dual5 r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.d[k] = cx(0.0);
    return r;
}
__device__ __forceinline__ Dual5 d5var(double v, int dir) {
    Dual5 r = d5const(cx(v));
    r.d[dir] = cx(1.0);
    return r;
}
__device__ __forceinline__ Dual5 operator+(Dual5 a, const Dual5& b) {
    a.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < 5; ++k) a.d[k] = a.d[k] + b.d[k];
    return a;
}
__device__ __forceinline__ Dual5 operator-(Dual5 a, const Dual5& b) {
    a.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < 5; ++k) a.d[k] = a.d[k] - b.d[k];
    return a;
}
__device__ __forceinline__ Dual5 operator*(const Dual5& a, const Dual5& b) {
    Dual5 r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ __forceinline__ Dual5 operator*(Cplx s, Dual5 a) {
    a.v = s * a.v;
#pragma unroll
    for (int k = 0; k < 5; ++k) a.d[k] = s * a.d[k];
    return a;
}
__device__ __forceinline__ Dual5 operator*(double s, Dual5 a) { return cx(s) * a; }
__device__ __forceinline__ Dual5 operator/(const Dual5& a, const Dual5& b) {
    Dual5 r;
    r.v = a.v / b.v;
    const Cplx ib = cx(1.0) / b.v;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
// f(a) with f'(a.v) = df
__device__ __forceinline__ Dual5 d5apply(const Dual5& a, Cplx f, Cplx df) {
    Dual5 r;
    r.v = f;
#pragma unroll
    for (int k = 0; k < 5; ++k) r.d[k] = df * a.d[k];
    return r;
}
__device__ __forceinline__ Dual5 dsqrt(const Dual5& a) {
    const Cplx s = csqrt(a.v);
    return d5apply(a, s, cx(0.5) / s);
}
__device__ __forceinline__ Dual5 dexp(const Dual5& a) {
    const Cplx e = cexp(a.v);
    return d5apply(a, e, e);
}
__device__ __forceinline__ Dual5 dlog(const Dual5& a) { return d5apply(a, clog(a.v), cx(1.0) / a.v); }
__device__ __forceinline__ Dual5 dsinh(const Dual5& a) { return d5apply(a, csinh(a.v), ccosh(a.v)); }
__device__ __forceinline__ Dual5 dcosh(const Dual5& a) { return d5apply(a, ccosh(a.v), csinh(a.v)); }

__device__ __forceinline__ void set_const(Cplx& x, Cplx c) { x = c; }
__device__ __forceinline__ void set_const(Dual5& x, Cplx c) { x = d5const(c); }

// The Heston characteristic function of the reference's parameterisation (HDistribution::Helpers
// and chf, src/HDistribution.cpp:12-38) at real u, generic over the number type: Cplx gives the
// value, Dual5 the value and its five partial derivatives.
template <typename D>
__device__ __forceinline__ D chf_expr(const D& v0, const D& vm, const D& rho, const D& kappa, const D& sigma,
                                      double tau, double u) {
    const Cplx iu = cx(0.0, u);
    const D xi = kappa + iu * (sigma * rho);
    D fac;  // the constant u^2 - i u
    set_const(fac, cx(u * u, -u));
    const D d = dsqrt(xi * xi + (sigma * sigma) * fac);
    const D half = (0.5 * tau) * d;
    const D sh = dsinh(half), ch = dcosh(half);
    const D a1 = fac * sh;
    const D a2 = (d / v0) * ch + (xi / v0) * sh;
    const D A = a1 / a2;
    const D den = (d + xi) + (d - xi) * dexp((-tau) * d);
    const D Dl = dlog((2.0 * d) / den) + (0.5 * tau) * (kappa - d);
    const D t1 = cx(0.0, tau * u) * ((kappa * vm) * rho) / sigma;
    const D t3 = 2.0 * ((Dl * kappa) * vm) / (sigma * sigma);
    return dexp((t1 - A) + t3);
}

// chf(u) and its gradient in (v0, theta, rho, kappa, sigma) in one pass.
__device__ Dual5 heston_chf5(const double (&p)[5], double tau, double u) {
    return chf_expr<Dual5>(d5var(p[0], 0), d5var(p[1], 1), d5var(p[2], 2), d5var(p[3], 3), d5var(p[4], 4), tau, u);
}

__device__ __forceinline__ uint32_t find_chain(const uint64_t* end, uint32_t n, uint64_t t) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (end[mid] > t)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ int64_t pmod(int64_t a, int64_t n) { return ((a % n) + n) % n; }

__device__ __forceinline__ double u_of(const DevSwift& R, uint32_t j) {
    return kPi * (2.0 * (double)j + 1.0) / (2.0 * (double)R.J) * R.exp2_m;
}

// Twiddle tables e^{2 pi i k / 4J}, k < 4J, one per distinct J of the call.
__global__ __launch_bounds__(256) void k_swift_twiddle(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ti = find_chain(a.tab_end, a.n_tables, t);
    const SwiftTable& T = a.tables[ti];
    const uint32_t k = (uint32_t)(t - (ti ? a.tab_end[ti - 1] : 0ull));
    double s, c;
    sincospi((double)k / (2.0 * (double)T.J), &s, &c);
    a.tw[T.off + k] = Cplx{c, s};
}

// Payoff coefficients of the call on [max(lower, 0), upper] (src/SWIFT.cpp:41-58):
// H(upper) - H(max(lower, 0)) with H(y) = -i e^{-i u y} (1/u - e^y / (i + u)).
__global__ __launch_bounds__(256) void k_swift_payoff(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_chain(a.j_end, a.n_req, t);
    const DevSwift& R = a.reqs[ri];
    const uint32_t k = (uint32_t)(t - (ri ? a.j_end[ri - 1] : 0ull));
    const double u = u_of(R, k);
    const double ylo = R.lower > 0.0 ? R.lower : 0.0;
    const Cplx iu1 = cx(u, 1.0);  // i + u
    const Cplx inv_u = cx(1.0 / u);
    const Cplx mi = cx(0.0, -1.0);
    const Cplx hu = mi * cexp(cx(0.0, -u * R.upper)) * (inv_u - cx(exp(R.upper)) / iu1);
    const Cplx hl = mi * cexp(cx(0.0, -u * ylo)) * (inv_u - cx(exp(ylo)) / iu1);
    a.pay[R.pay_off + k] = hu - hl;
}

// In-place radix-2 decimation-in-time FFT of length N = 2^logN in LDS, input in bit-reversed
// order, output in natural order; sign +1: sum x_n e^{+2 pi i n k / N}, -1: e^{-2 pi i n k / N}.
// tw holds e^{2 pi i k / 2N} for k < 2N (the chain's 4J table, N = 2J).
template <int SIGN>
__device__ void fft_lds(Cplx* x, uint32_t N, uint32_t logN, const Cplx* __restrict__ tw) {
    for (uint32_t s = 0; s < logN; ++s) {
        const uint32_t half = 1u << s;
        for (uint32_t b = threadIdx.x; b < N / 2; b += blockDim.x) {
            const uint32_t pos = b & (half - 1u);
            const uint32_t i0 = ((b >> s) << (s + 1)) + pos;
            const uint32_t i1 = i0 + half;
            Cplx w = tw[pos * (N >> s)];  // e^{2 pi i pos / 2 half}
            if (SIGN < 0) w = conj(w);
            const Cplx u = x[i0], v = x[i1] * w;
            x[i0] = u + v;
            x[i1] = u - v;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t bitrev(uint32_t v, uint32_t bits) { return __builtin_bitreverse32(v) >> (32u - bits); }

// First transform as an FFT (src/SWIFT.cpp:59-70), one workgroup per chain: the length-2J
// backward FFT of the zero-padded payoff coefficients, then for each wavelet index i in
// [k1, k2]: dens_i = Re(e^{i pi i/2J} out[i mod 2J]) sqrt(2^m) / J.
__global__ __launch_bounds__(256) void k_swift_fft1(SwiftArgs a) {
    extern __shared__ Cplx lds_c[];
    const DevSwift& R = a.reqs[a.fft_chains[blockIdx.x]];
    const uint32_t N = 2u * R.J, logN = R.log2_2J;
    const Cplx* tw = a.tw + R.tw_off;
    const Cplx* pay = a.pay + R.pay_off;
    for (uint32_t j = threadIdx.x; j < N; j += blockDim.x) lds_c[bitrev(j, logN)] = j < R.J ? pay[j] : cx(0.0);
    __syncthreads();
    fft_lds<+1>(lds_c, N, logN, tw);
    const int64_t n4 = 2 * (int64_t)N;
    const double scale = R.sqrt_exp2_m / (double)R.J;
    for (int64_t i = R.k1 + (int64_t)threadIdx.x; i <= R.k2; i += blockDim.x)
        a.dens[R.dens_off + (uint64_t)(i - R.k1)] = (tw[pmod(i, n4)] * lds_c[pmod(i, N)]).re * scale;
}

// Second transform as an FFT (src/SWIFT.cpp:71-72), one workgroup per chain. The length-4J real
// DFT at the odd bin 2j+1 is a length-2J DFT: X_{2j+1} = sum_{n<2J} z_n e^{-2 pi i n j / 2J} with
// z_n = y_n + y_{n+2J}, y_n = x_n e^{-i pi n / 2J} and x_n the density input at n = i mod 4J.
// c_j = conj(X_{2j+1}) sqrt(2^m) / J.
__global__ __launch_bounds__(256) void k_swift_fft2(SwiftArgs a) {
    extern __shared__ Cplx lds_c[];
    const DevSwift& R = a.reqs[a.fft_chains[blockIdx.x]];
    const uint32_t N = 2u * R.J, logN = R.log2_2J;
    const int64_t n4 = 2 * (int64_t)N;
    const Cplx* tw = a.tw + R.tw_off;
    const double* dens = a.dens + R.dens_off;
    for (uint32_t n = threadIdx.x; n < N; n += blockDim.x) {
        Cplx z = cx(0.0);
        for (uint32_t h = 0; h < 2; ++h) {
            const int64_t nn = (int64_t)n + (int64_t)h * N;
            const int64_t i = R.k1 + pmod(nn - R.k1, n4);  // the wavelet index with i mod 4J = nn
            if (i <= R.k2) z = z + dens[i - R.k1] * conj(tw[nn]);
        }
        lds_c[bitrev(n, logN)] = z;
    }
    __syncthreads();
    fft_lds<-1>(lds_c, N, logN, tw);
    const double scale = R.sqrt_exp2_m / (double)R.J;
    for (uint32_t j = threadIdx.x; j < R.J; j += blockDim.x) a.coeff[R.coeff_off + j] = scale * conj(lds_c[j]);
}

// First transform (src/SWIFT.cpp:59-70): for each wavelet index i in [k1, k2], the length-2J
// backward DFT of the zero-padded payoff coefficients at i mod 2J, rotated by e^{i pi i / 2J}:
// dens_i = Re(e^{i pi i/2J} sum_j pay_j e^{2 pi i j i / 2J}) sqrt(2^m) / J.
// Direct sum for chains above the FFT size: one wavefront per i, lane l sums j = l, l+64, ... in
// order (twiddle tw[(2 j i2) mod 4J]), a fixed xor butterfly combines the lanes.
__global__ __launch_bounds__(256) void k_swift_dens(SwiftArgs a, uint64_t total) {
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64u;  // wave-uniform
    const uint32_t lane = threadIdx.x % 64u;
    if (w >= total) return;
    const uint32_t ri = find_chain(a.dens_end, a.n_req, w);
    const DevSwift& R = a.reqs[ri];
    const int64_t il = (int64_t)(w - (ri ? a.dens_end[ri - 1] : 0ull));
    const int64_t i = R.k1 + il;
    const int64_t n4 = 4 * (int64_t)R.J;
    const Cplx* tw = a.tw + R.tw_off;
    const Cplx* pay = a.pay + R.pay_off;
    const int64_t step = 2 * pmod(i, 2 * (int64_t)R.J);  // index advance per unit of j
    const int64_t stride = (64 * step) % n4;             // per lane iteration
    int64_t idx = ((int64_t)lane * step) % n4;
    double re = 0.0, im = 0.0;
    for (uint32_t j = lane; j < R.J; j += 64u) {
        const Cplx q = pay[j] * tw[idx];
        re += q.re;
        im += q.im;
        idx += stride;
        if (idx >= n4) idx -= n4;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        re += __shfl_xor(re, m);
        im += __shfl_xor(im, m);
    }
    if (lane != 0) return;
    const Cplx rot = tw[pmod(i, n4)];
    a.dens[R.dens_off + (uint64_t)il] = (rot * Cplx{re, im}).re * R.sqrt_exp2_m / (double)R.J;
}

// Second transform (src/SWIFT.cpp:71-72): the length-4J real forward DFT of dens at the odd bin
// 2j+1, conjugated and scaled: c_j = conj(sum_i dens_i e^{-2 pi i (i mod 4J)(2j+1)/4J}) sqrt(2^m)/J.
// Direct sum for chains above the FFT size: one wavefront per j, lanes strided over i as above.
__global__ __launch_bounds__(256) void k_swift_coeff(SwiftArgs a, uint64_t total) {
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64u;  // wave-uniform
    const uint32_t lane = threadIdx.x % 64u;
    if (w >= total) return;
    const uint32_t ri = find_chain(a.dcoef_end, a.n_req, w);
    const DevSwift& R = a.reqs[ri];
    const uint32_t j = (uint32_t)(w - (ri ? a.dcoef_end[ri - 1] : 0ull));
    const int64_t n4 = 4 * (int64_t)R.J;
    const Cplx* tw = a.tw + R.tw_off;
    const double* dens = a.dens + R.dens_off;
    const int64_t k = 2 * (int64_t)j + 1;  // < n4
    const int64_t stride = (64 * k) % n4;
    int64_t idx = (pmod(R.k1 + (int64_t)lane, n4) * k) % n4;
    double re = 0.0, im = 0.0;
    for (int64_t i = R.k1 + (int64_t)lane; i <= R.k2; i += 64) {
        const double d = dens[i - R.k1];
        const Cplx t = tw[idx];
        re += d * t.re;
        im -= d * t.im;  // conj(tw)
        idx += stride;
        if (idx >= n4) idx -= n4;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        re += __shfl_xor(re, m);
        im += __shfl_xor(im, m);
    }
    if (lane != 0) return;
    a.coeff[R.coeff_off + j] = (R.sqrt_exp2_m / (double)R.J) * Cplx{re, -im};
}

// Pricing matrix (src/SWIFT.cpp:106-110): row 0 = chf(u_j) c_j, rows 1..5 = d chf / d p_k c_j.
__global__ __launch_bounds__(256) void k_swift_chf(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_chain(a.j_end, a.n_req, t);
    const DevSwift& R = a.reqs[ri];
    const uint32_t j = (uint32_t)(t - (ri ? a.j_end[ri - 1] : 0ull));
    const double p[5] = {R.p[0], R.p[1], R.p[2], R.p[3], R.p[4]};
    const double u = u_of(R, j);
    const Cplx c = a.coeff[R.coeff_off + j];
    Cplx* pm = a.pm + R.pm_off;
    // the value row comes from the plain evaluation in both modes, so prices do not depend on
    // whether gradients were asked for
    pm[j] = chf_expr<Cplx>(cx(p[0]), cx(p[1]), cx(p[2]), cx(p[3]), cx(p[4]), R.T, u) * c;
    if (R.rows == 6) {
        const Dual5 f = heston_chf5(p, R.T, u);
#pragma unroll
        for (int k = 0; k < 5; ++k) pm[(uint64_t)(k + 1) * R.J + j] = f.d[k] * c;
    }
}

// Final contraction (src/SWIFT.cpp:111-117): out[s][row] = Re sum_j pm[row][j] e^{-rT} K_s e^{-i u_j x_s}.
// One workgroup per (chain, group of kSwiftStrikeBlock = 16 strikes): thread (s, p) = (tid % 16,
// tid / 16) owns strike s and sums j = p, p+16, ... in order; the pricing matrix is staged
// through LDS in chunks of 256 j; the 16 partial sums of a strike are added in a fixed order.
constexpr int kSumChunk = 256;
constexpr uint32_t kSumPhases = 256 / kSwiftStrikeBlock;
__global__ __launch_bounds__(256) void k_swift_sum(SwiftArgs a, uint64_t total) {
    __shared__ Cplx pm_s[6][kSumChunk];
    __shared__ double part_s[6][256];
    const uint64_t blk = blockIdx.x;
    if (blk >= total) return;  // uniform over the workgroup
    const uint32_t ri = find_chain(a.sblk_end, a.n_req, blk);
    const DevSwift& R = a.reqs[ri];
    const uint32_t o0 = (uint32_t)(blk - (ri ? a.sblk_end[ri - 1] : 0ull)) * kSwiftStrikeBlock;
    const uint32_t sl = threadIdx.x % kSwiftStrikeBlock, ph = threadIdx.x / kSwiftStrikeBlock;
    const uint32_t so = o0 + sl;
    const bool on = so < R.n_strikes;
    const double K = a.strikes[R.strike_off + (on ? so : o0)];
    const double x = R.r * R.T + log(R.S0 / K);
    const double scale = on ? R.disc * K : 0.0;
    const uint32_t rows = R.rows;
    // e^{-i u_j x} along j = p, p+16, ...: rotation by e^{-i 16 (u_{j+1} - u_j) x}, re-anchored with
    // a direct evaluation every 16 steps (rounding grows by at most 16 ulps)
    double sr, cr;
    sincos(-((double)kSumPhases * kPi / (double)R.J * R.exp2_m) * x, &sr, &cr);
    const Cplx rot = Cplx{cr, sr};
    const Cplx* pm = a.pm + R.pm_off;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t jc = 0; jc < R.J; jc += kSumChunk) {
        const uint32_t n = min((uint32_t)kSumChunk, R.J - jc);
        __syncthreads();  // the previous chunk has been consumed
        for (uint32_t t = threadIdx.x; t < rows * n; t += blockDim.x) {
            const uint32_t row = t / n, jj = t - row * n;
            pm_s[row][jj] = pm[(uint64_t)row * R.J + jc + jj];
        }
        __syncthreads();
        Cplx z = cx(1.0);
        uint32_t it = 0;
        for (uint32_t jj = ph; jj < n; jj += kSumPhases, ++it) {
            if ((it & 15u) == 0) {
                double sn, cs;
                sincos(-u_of(R, jc + jj) * x, &sn, &cs);
                z = Cplx{cs, sn};
            } else {
                z = z * rot;
            }
            const double er = scale * z.re, ei = scale * z.im;
            acc[0] += pm_s[0][jj].re * er - pm_s[0][jj].im * ei;
            if (rows == 6) {
#pragma unroll
                for (int k = 1; k < 6; ++k) acc[k] += pm_s[k][jj].re * er - pm_s[k][jj].im * ei;
            }
        }
    }
    for (uint32_t k = 0; k < rows; ++k) part_s[k][threadIdx.x] = acc[k];
    __syncthreads();
    // thread (s, row) for the first 16 * rows threads adds the 16 phases of strike s in order
    const uint32_t s_out = threadIdx.x % kSwiftStrikeBlock, k = threadIdx.x / kSwiftStrikeBlock;
    if (k >= rows || o0 + s_out >= R.n_strikes) return;
    double v = 0.0;
    for (uint32_t p = 0; p < kSumPhases; ++p) v += part_s[k][p * kSwiftStrikeBlock + s_out];
    const uint64_t o = R.out_off + o0 + s_out;
    if (k == 0)
        a.prices[o] = v;
    else
        a.grads[o * 5 + (k - 1)] = v;
}

// The dynamic grid of chain blockIdx.x (SwiftParameters ctor, src/SWIFT.cpp:21-35), one wavefront
// per chain. The lanes first find the chain's extreme strikes (and whether every strike is finite
// and positive). The reference loops m = 0, 1, ... while int_error(m) > trunc and keeps one past
// the first m that stops it (src/HDistribution.cpp:8-11: |chf(2^m pi) + chf(-2^m pi)| /
// (4 2^m pi^2 T)); here lane m evaluates scale m for m = 0..30 at once and the lowest lane whose
// error does not exceed trunc (a NaN error stops the loop too) is that first m. Lanes past it
// evaluate scales the loop never reaches; their values are not used.
__global__ __launch_bounds__(64) void k_swift_grid(const SwiftGridIn* in, const double* strikes, SwiftGridOut* out,
                                                   uint32_t n) {
    const uint32_t c = blockIdx.x;
    if (c >= n) return;
    const int lane = (int)threadIdx.x;
    SwiftGridIn q = in[c];
    const double* ks = strikes + q.strike_off;
    double lo = INFINITY, hi = -INFINITY;
    bool bad = false;
    for (uint32_t i = (uint32_t)lane; i < q.n_strikes; i += 64) {
        const double k = ks[i];
        bad |= !(k > 0.0 && k < INFINITY);
        lo = fmin(lo, k);
        hi = fmax(hi, k);
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        lo = fmin(lo, __shfl_xor(lo, m));
        hi = fmax(hi, __shfl_xor(hi, m));
    }
    const bool any_bad = __builtin_amdgcn_ballot_w64(bad) != 0;
    q.kmin = lo;
    q.kmax = hi;
    bool stop = false;
    if (lane <= 30) {
        const double e2m = ::ldexp(1.0, lane);
        const double u = e2m * kPi;
        const Cplx a = chf_expr<Cplx>(cx(q.p[0]), cx(q.p[1]), cx(q.p[2]), cx(q.p[3]), cx(q.p[4]), q.T, u);
        const Cplx b = chf_expr<Cplx>(cx(q.p[0]), cx(q.p[1]), cx(q.p[2]), cx(q.p[3]), cx(q.p[4]), q.T, -u);
        const double err = ::hypot(a.re + b.re, a.im + b.im) / (4.0 * e2m * kPi * kPi * q.T);
        stop = !(err > q.trunc);
    }
    const uint64_t mask = __builtin_amdgcn_ballot_w64(stop);
    if (lane == 0) {
        SwiftGridOut o{};
        if (any_bad) {
            o.status = kSwiftGridBadStrike;
        } else if (mask == 0) {
            o.status = kSwiftGridNoScale;
        } else {
            const uint32_t m = (uint32_t)__builtin_ctzll(mask) + 1u;
            o.status = swift_grid_finish(q, m, o.g, o.J_wanted);
        }
        out[c] = o;
    }
}

// Copies two ranges of 16-byte words from fine-grained pinned host memory to the device (a call's
// tables and its strikes), so the pricing kernels read them from HBM.
__global__ __launch_bounds__(256) void k_swift_stage(const uint4* s0, uint4* d0, uint64_t n0, const uint4* s1,
                                                     uint4* d1, uint64_t n1) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n0 + n1; i += stride) {
        if (i < n0)
            d0[i] = s0[i];
        else
            d1[i - n0] = s1[i - n0];
    }
}

inline dim3 grid_of(uint64_t n, uint32_t block) { return dim3((uint32_t)((n + block - 1) / block)); }

}  // namespace

hipError_t launch_swift(const SwiftArgs& a, const SwiftTotals& t, hipStream_t st) {
    k_swift_twiddle<<<grid_of(t.tw, 256), 256, 0, st>>>(a, t.tw);
    k_swift_payoff<<<grid_of(t.j, 256), 256, 0, st>>>(a, t.j);
    if (t.n_fft) {
        // the FFT kernels' dynamic LDS limit, raised once per device to the largest FFT's need
        static std::atomic<uint64_t> lds_set{0};  // bit d: done on device d
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        const uint64_t bit = 1ull << (dev & 63);
        if (!(lds_set.load() & bit)) {
            const int lds = (int)(sizeof(Cplx) * 2 * kSwiftFftMaxJ);
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_swift_fft1),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            if (e == hipSuccess)
                e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_swift_fft2),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, lds);
            if (e != hipSuccess) return e;
            lds_set.fetch_or(bit);
        }
        k_swift_fft1<<<t.n_fft, 256, t.fft_lds_bytes, st>>>(a);
    }
    if (t.dens) k_swift_dens<<<grid_of(t.dens * 64u, 256), 256, 0, st>>>(a, t.dens);
    if (t.n_fft) k_swift_fft2<<<t.n_fft, 256, t.fft_lds_bytes, st>>>(a);
    if (t.dcoef) k_swift_coeff<<<grid_of(t.dcoef * 64u, 256), 256, 0, st>>>(a, t.dcoef);
    k_swift_chf<<<grid_of(t.j, 256), 256, 0, st>>>(a, t.j);
    k_swift_sum<<<(uint32_t)t.sblocks, 256, 0, st>>>(a, t.sblocks);
    return hipGetLastError();
}

hipError_t launch_swift_stage(const void* s0, void* d0, size_t bytes0, const void* s1, void* d1, size_t bytes1,
                              hipStream_t st) {
    if (bytes0 % 16 || bytes1 % 16) return hipErrorInvalidValue;
    const uint64_t n0 = bytes0 / 16, n1 = bytes1 / 16;
    if (n0 + n1 == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n0 + n1 + 255) / 256, 1024);
    k_swift_stage<<<blocks, 256, 0, st>>>(static_cast<const uint4*>(s0), static_cast<uint4*>(d0), n0,
                                          static_cast<const uint4*>(s1), static_cast<uint4*>(d1), n1);
    return hipGetLastError();
}

hipError_t launch_swift_grid(const SwiftGridIn* in, const double* strikes, SwiftGridOut* out, uint32_t n,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_swift_grid<<<n, 64, 0, st>>>(in, strikes, out, n);
    return hipGetLastError();
}

}  // namespace hexo
