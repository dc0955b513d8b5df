// swift.hip -- SWIFT European pricer on gfx950, FP64 throughout.
//
// The Shannon-wavelet (SWIFT) method of the reference (src/SWIFT.cpp:18-118): payoff
// coefficients of the call on the truncated log-price range, two Fourier transforms to the
// density coefficients c_j, then price = Re sum_j chf(u_j) c_j e^{-rT} K e^{-i u_j x} with
// x = rT + log(S/K) (src/SWIFT.cpp:102-117). The reference uses FFTW for the transforms and
// Eigen for the final 6 x J by J x K product; here each stage is one kernel over all chains of
// the call: the transforms are direct sums (J and k2 - k1 are a few hundred to a few thousand,
// and every output is independent), the characteristic function is evaluated per (chain, j),
// and each (chain, strike) is one wavefront that sums over j with a fixed xor butterfly.
//
// Gradients (SWIFT::price_opts_grad, the calibration Jacobian) come from the same
// characteristic-function expression evaluated in forward-mode dual arithmetic, one parameter
// direction at a time, instead of hand-derived partials (src/HDistribution.cpp:47-84).
#include "swift.hpp"

#include <hip/hip_runtime.h>

namespace hexo {
namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ Cplx cx(double re, double im = 0.0) { return Cplx{re, im}; }
__device__ __forceinline__ Cplx operator+(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cplx operator-(Cplx a, Cplx b) { return Cplx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cplx operator*(Cplx a, Cplx b) {
    return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cplx operator*(double s, Cplx a) { return Cplx{s * a.re, s * a.im}; }
__device__ __forceinline__ Cplx conj(Cplx a) { return Cplx{a.re, -a.im}; }
// Smith's division: no overflow or underflow from forming |b|^2
__device__ __forceinline__ Cplx operator/(Cplx a, Cplx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double t = b.im / b.re, den = b.re + b.im * t;
        return Cplx{(a.re + a.im * t) / den, (a.im - a.re * t) / den};
    }
    const double t = b.re / b.im, den = b.re * t + b.im;
    return Cplx{(a.re * t + a.im) / den, (a.im * t - a.re) / den};
}
__device__ __forceinline__ Cplx cexp(Cplx a) {
    double s, c;
    sincos(a.im, &s, &c);
    const double e = exp(a.re);
    return Cplx{e * c, e * s};
}
__device__ __forceinline__ Cplx clog(Cplx a) { return Cplx{log(hypot(a.re, a.im)), atan2(a.im, a.re)}; }
// principal square root (branch cut on the negative real axis, sign of the imaginary part kept)
__device__ __forceinline__ Cplx csqrt(Cplx a) {
    const double r = hypot(a.re, a.im);
    if (r == 0.0) return Cplx{0.0, a.im};
    if (a.re >= 0.0) {
        const double t = sqrt(0.5 * (r + a.re));
        return Cplx{t, a.im / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (r - a.re));
    return Cplx{fabs(a.im) / (2.0 * t), copysign(t, a.im)};
}
__device__ __forceinline__ Cplx csinh(Cplx a) {
    double s, c;
    sincos(a.im, &s, &c);
    return Cplx{sinh(a.re) * c, cosh(a.re) * s};
}
__device__ __forceinline__ Cplx ccosh(Cplx a) {
    double s, c;
    sincos(a.im, &s, &c);
    return Cplx{cosh(a.re) * c, sinh(a.re) * s};
}

// Value and one directional derivative (forward-mode dual number over the complex field).
struct Dual {
    Cplx v, d;
};
__device__ __forceinline__ Dual dconst(Cplx v) { return Dual{v, cx(0.0)}; }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(double s, Dual a) { return Dual{s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator*(Cplx s, Dual a) { return Dual{s * a.v, s * a.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const Cplx q = a.v / b.v;
    return Dual{q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
    const Cplx s = csqrt(a.v);
    return Dual{s, a.d / (2.0 * s)};
}
__device__ __forceinline__ Dual dexp(Dual a) {
    const Cplx e = cexp(a.v);
    return Dual{e, e * a.d};
}
__device__ __forceinline__ Dual dlog(Dual a) { return Dual{clog(a.v), a.d / a.v}; }
__device__ __forceinline__ Dual dsinh(Dual a) { return Dual{csinh(a.v), ccosh(a.v) * a.d}; }
__device__ __forceinline__ Dual dcosh(Dual a) { return Dual{ccosh(a.v), csinh(a.v) * a.d}; }

// The Heston characteristic function of the reference's parameterisation
// (HDistribution::Helpers and chf, src/HDistribution.cpp:12-38) at real u, as a dual number
// whose tangent is d/d p[dir] (dir < 0: value only).
__device__ Dual heston_chf(const double (&p)[5], double tau, double u, int dir) {
    Dual v0 = dconst(cx(p[0])), vm = dconst(cx(p[1])), rho = dconst(cx(p[2])), kappa = dconst(cx(p[3])),
         sigma = dconst(cx(p[4]));
    if (dir == 0) v0.d = cx(1.0);
    if (dir == 1) vm.d = cx(1.0);
    if (dir == 2) rho.d = cx(1.0);
    if (dir == 3) kappa.d = cx(1.0);
    if (dir == 4) sigma.d = cx(1.0);
    const Cplx iu = cx(0.0, u);
    const Dual xi = kappa + iu * (sigma * rho);
    const Dual fac = dconst(cx(u * u, -u));
    const Dual d = dsqrt(xi * xi + (sigma * sigma) * fac);
    const Dual half = (0.5 * tau) * d;
    const Dual sh = dsinh(half), ch = dcosh(half);
    const Dual a1 = fac * sh;
    const Dual a2 = (d / v0) * ch + (xi / v0) * sh;
    const Dual A = a1 / a2;
    const Dual den = (d + xi) + (d - xi) * dexp((-tau) * d);
    const Dual D = dlog((2.0 * d) / den) + (0.5 * tau) * (kappa - d);
    const Dual t1 = cx(0.0, tau * u) * ((kappa * vm) * rho) / sigma;
    const Dual t3 = 2.0 * ((D * kappa) * vm) / (sigma * sigma);
    return dexp((t1 - A) + t3);
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

// Twiddles e^{2 pi i k / 4J} (k < 4J) and, for k < J, the payoff coefficients
// H(upper) - H(max(lower, 0)) with H(y) = -i e^{-i u y} (1/u - e^y / (i + u)) (src/SWIFT.cpp:41-58).
__global__ __launch_bounds__(256) void k_swift_prep(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_chain(a.tw_end, a.n_req, t);
    const DevSwift& R = a.reqs[ri];
    const uint32_t k = (uint32_t)(t - (ri ? a.tw_end[ri - 1] : 0ull));
    double s, c;
    sincospi((double)k / (2.0 * (double)R.J), &s, &c);
    a.tw[R.tw_off + k] = Cplx{c, s};
    if (k < R.J) {
        const double u = u_of(R, k);
        const double ylo = R.lower > 0.0 ? R.lower : 0.0;
        const Cplx iu1 = cx(u, 1.0);  // i + u
        const Cplx inv_u = cx(1.0 / u);
        const Cplx mi = cx(0.0, -1.0);
        const Cplx hu = mi * cexp(cx(0.0, -u * R.upper)) * (inv_u - cx(exp(R.upper)) / iu1);
        const Cplx hl = mi * cexp(cx(0.0, -u * ylo)) * (inv_u - cx(exp(ylo)) / iu1);
        a.pay[R.pay_off + k] = hu - hl;
    }
}

// First transform (src/SWIFT.cpp:59-70): for each wavelet index i in [k1, k2], the length-2J
// backward DFT of the zero-padded payoff coefficients at i mod 2J, rotated by e^{i pi i / 2J}:
// dens_i = Re(e^{i pi i/2J} sum_j pay_j e^{2 pi i j i / 2J}) sqrt(2^m) / J.
__global__ __launch_bounds__(256) void k_swift_dens(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_chain(a.dens_end, a.n_req, t);
    const DevSwift& R = a.reqs[ri];
    const int64_t il = (int64_t)(t - (ri ? a.dens_end[ri - 1] : 0ull));
    const int64_t i = R.k1 + il;
    const int64_t n4 = 4 * (int64_t)R.J;
    const Cplx* tw = a.tw + R.tw_off;
    const Cplx* pay = a.pay + R.pay_off;
    // e^{2 pi i j i2 / 2J} = tw[(2 j i2) mod 4J]; the index advances by 2 i2 per j
    const int64_t step = 2 * pmod(i, 2 * (int64_t)R.J);
    int64_t idx = 0;
    Cplx acc = cx(0.0);
    for (uint32_t j = 0; j < R.J; ++j) {
        acc = acc + pay[j] * tw[idx];
        idx += step;
        if (idx >= n4) idx -= n4;
    }
    const Cplx rot = tw[pmod(i, n4)];
    a.dens[R.dens_off + (uint64_t)il] = (rot * acc).re * R.sqrt_exp2_m / (double)R.J;
}

// Second transform (src/SWIFT.cpp:71-72): the length-4J real forward DFT of dens at the odd bin
// 2j+1, conjugated and scaled: c_j = conj(sum_i dens_i e^{-2 pi i (i mod 4J)(2j+1)/4J}) sqrt(2^m)/J.
__global__ __launch_bounds__(256) void k_swift_coeff(SwiftArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_chain(a.j_end, a.n_req, t);
    const DevSwift& R = a.reqs[ri];
    const uint32_t j = (uint32_t)(t - (ri ? a.j_end[ri - 1] : 0ull));
    const int64_t n4 = 4 * (int64_t)R.J;
    const Cplx* tw = a.tw + R.tw_off;
    const double* dens = a.dens + R.dens_off;
    const int64_t k = 2 * (int64_t)j + 1;
    int64_t idx = (pmod(R.k1, n4) * k) % n4;
    const int64_t step = k % n4;
    Cplx acc = cx(0.0);
    for (int64_t i = R.k1; i <= R.k2; ++i) {
        acc = acc + dens[i - R.k1] * conj(tw[idx]);
        idx += step;
        if (idx >= n4) idx -= n4;
    }
    a.coeff[R.coeff_off + j] = (R.sqrt_exp2_m / (double)R.J) * conj(acc);
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
    pm[j] = heston_chf(p, R.T, u, -1).v * c;
    for (uint32_t k = 1; k < R.rows; ++k) pm[(uint64_t)k * R.J + j] = heston_chf(p, R.T, u, (int)k - 1).d * c;
}

// One wavefront per (chain, strike): lane l sums j = l, l+64, ... of
// Re(pm[row][j] e^{-rT} K e^{-i u_j x}) in order, a fixed xor butterfly combines the lanes.
__global__ __launch_bounds__(256) void k_swift_sum(SwiftArgs a, uint64_t total) {
    const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64u;  // wave-uniform
    const uint32_t lane = threadIdx.x % 64u;
    if (w >= total) return;
    const uint32_t ri = find_chain(a.strike_end, a.n_req, w);
    const DevSwift& R = a.reqs[ri];
    const uint32_t o = (uint32_t)(w - (ri ? a.strike_end[ri - 1] : 0ull));
    const double K = a.strikes[R.strike_off + o];
    const double x = R.r * R.T + log(R.S0 / K);
    const double scale = R.disc * K;
    const Cplx* pm = a.pm + R.pm_off;
    const bool with_grads = R.rows == 6;  // wave-uniform
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t j = lane; j < R.J; j += 64u) {
        double s, c;
        sincos(-u_of(R, j) * x, &s, &c);
        const double er = scale * c, ei = scale * s;
        acc[0] += pm[j].re * er - pm[j].im * ei;
        if (with_grads) {
#pragma unroll
            for (int k = 1; k < 6; ++k) {
                const Cplx q = pm[(uint64_t)k * R.J + j];
                acc[k] += q.re * er - q.im * ei;
            }
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += __shfl_xor(acc[k], m);
    }
    if (lane != 0) return;
    a.prices[R.out_off + o] = acc[0];
    if (with_grads)
        for (int k = 0; k < 5; ++k) a.grads[(R.out_off + o) * 5 + k] = acc[k + 1];
}

inline dim3 grid_of(uint64_t n, uint32_t block) { return dim3((uint32_t)((n + block - 1) / block)); }

}  // namespace

hipError_t launch_swift(const SwiftArgs& a, const SwiftTotals& t, hipStream_t st) {
    k_swift_prep<<<grid_of(t.tw, 256), 256, 0, st>>>(a, t.tw);
    k_swift_dens<<<grid_of(t.dens, 256), 256, 0, st>>>(a, t.dens);
    k_swift_coeff<<<grid_of(t.j, 256), 256, 0, st>>>(a, t.j);
    k_swift_chf<<<grid_of(t.j, 256), 256, 0, st>>>(a, t.j);
    k_swift_sum<<<grid_of(t.strikes * 64u, 256), 256, 0, st>>>(a, t.strikes);
    return hipGetLastError();
}

}  // namespace hexo
