// geo.hip -- exact price of the geometric control variate of the Asian call, FP64 on gfx950.
//
// For an expiry's Asian functional A = sum_j w_j X_{t_j} (weights from the run schedule, sum W)
// the control variate is G = c e^L with c = W S0 and L = sum_j v_j ln(X_{t_j}/S0), v = w / W.
// Under the Heston model
//   phi(s) = E[e^{sL}] = exp(A_0 + B_0 v0)
// by a backward recursion over the grid points: at t_j the log-price coefficient a grows by s v_j;
// over (t_{j-1}, t_j] the Riccati equations of exp(A + a y + B V),
//   B' = sigma^2 B^2 / 2 + (rho sigma a - kappa) B + (a^2 - a) / 2,   A' = kappa theta B + r a,
// are solved in closed form from the running B (no cancellation as sigma -> 0, see geo_step).
// The call price is Lewis' inversion
//   E[(c e^L - K)^+] = c (phi(1) - e^{k/2}/pi int_0^inf Re(e^{iuk} phi(1/2 - iu)) / (u^2 + 1/4) du),
// k = ln(K / c), by 64-node Gauss-Legendre panels of width h (host-chosen from the scale of L and
// the strikes) until the characteristic function has decayed below 1e-15 on a whole panel.
//
// One wavefront per (request, expiry, block of <= 64 strikes): lane q evaluates node q of the
// current panel (a recursion over the expiry's grid, a few hundred to a few thousand steps); the
// panel's values go through LDS and lane i accumulates strike i's integral over the 64 nodes in
// node order, so every price is a fixed-order sum. Runs once per batch, before the path kernels.
#include <hip/hip_runtime.h>

#include "cplx_device.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace hexo {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kGeoWaves = 4;  // wavefronts per workgroup

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double cabs_(Cplx z) { return hypot(z.re, z.im); }

// phi(s) over points 0..n of pts = (t_j, v_j) pairs.
// With beta = kappa - rho sigma a, d = sqrt(beta^2 - sigma^2 (a^2 - a)), the stable root
// B- = (a^2 - a)/(beta + d) (or (beta - d)/sigma^2 when beta + d is the smaller) and
// e = e^{-d tau}, om = 1 - e:
//   B(tau) = B- + (B - B-) e 2d / D,   D = (beta + d - sigma^2 B) - sigma^2 (B- - B) e,
//   A(tau) = A + r a tau + kappa theta (B- tau - (2/sigma^2) log(1 + sigma^2 (B- - B) om / (2d))).
__device__ Cplx geo_cf(const DevReq& R, const double* pts, uint32_t n, Cplx s) {
    const double rho = R.rho, kap = R.kappa, sig = R.sigma, s2 = sig * sig, kth = R.kappa * R.theta;
    Cplx a = cx(0.0), A = cx(0.0), B = cx(0.0);
    for (uint32_t j = n; j >= 1; --j) {
        const double tau = pts[2 * j] - pts[2 * (j - 1)];
        a = a + pts[2 * j + 1] * s;
        const Cplx bt = cx(kap) - (rho * sig) * a;
        const Cplx q = a * a - a;
        const Cplx d = csqrt(bt * bt - s2 * q);
        const Cplx bp = bt + d, bm = bt - d;
        const Cplx Bm = cabs_(bp) >= cabs_(bm) ? q / bp : (1.0 / s2) * bm;
        const Cplx om = -1.0 * cexpm1(-tau * d);
        const Cplx e = cx(1.0) - om;
        const Cplx D = (bp - s2 * B) - s2 * ((Bm - B) * e);
        const Cplx Bn = Bm + ((B - Bm) * e) * ((2.0 * d) / D);
        const Cplx z = (s2 * ((Bm - B) * om)) / (2.0 * d);
        A = A + (R.r_drift * tau) * a + kth * (tau * Bm - (2.0 / s2) * clog1p(z));
        B = Bn;
    }
    return cexp(A + R.v0 * B);
}

__global__ __launch_bounds__(256) void k_geo_exact(GeoArgs a) {
    __shared__ double sh[kGeoWaves][3][64];  // per wave: Re phi, Im phi (times the node weight), u
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const uint32_t item = blockIdx.x * kGeoWaves + (uint32_t)w;
    if (item >= a.n_items) return;  // whole wavefronts only
    const DevGeo& G = a.items[item];
    const DevReq& R = a.reqs[G.req];
    const double* pts = a.pts + 2 * (size_t)G.pts_begin;
    const bool has_k = (uint32_t)lane < G.n_strikes;
    const double K = has_k ? a.opts[G.opt + (uint32_t)lane].k_abs : 0.0;
    const double k = K > 0.0 ? log(K / G.c) : 0.0;
    const Cplx e1 = geo_cf(R, pts, G.n_pts, cx(1.0));  // E[e^L]
    const double xq = a.gl[lane], wq = a.gl[64 + lane];
    double integ = 0.0;
    bool converged = false;
    for (uint32_t p = 0; p < G.max_panels; ++p) {
        const double u = ((double)p + 0.5 * (xq + 1.0)) * G.h;
        const Cplx ph = geo_cf(R, pts, G.n_pts, cx(0.5, -u));
        const double f = 0.5 * G.h * wq / (u * u + 0.25);
        sh[w][0][lane] = f * ph.re;
        sh[w][1][lane] = f * ph.im;
        sh[w][2][lane] = u;
        double amax = cabs_(ph);
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) amax = fmax(amax, __shfl_xor(amax, m));
        wave_sync_lds();
        if (K > 0.0) {
            for (int q = 0; q < 64; ++q) {
                double sn, cs;
                sincos(sh[w][2][q] * k, &sn, &cs);
                integ += cs * sh[w][0][q] - sn * sh[w][1][q];
            }
        }
        wave_sync_lds();
        if (p > 0 && amax < 1e-15) {  // wave-uniform: every lane holds the wave's maximum
            converged = true;
            break;
        }
    }
    if (!has_k) return;
    double price;
    if (!converged)
        price = __builtin_nan("");
    else if (K > 0.0)
        price = G.c * (e1.re - exp(0.5 * k) / kPi * integ);
    else
        price = G.c * e1.re - K;  // max(G - K, 0) = G - K
    a.geo[G.out + (uint64_t)lane] = price;
}

}  // namespace

hipError_t launch_geo(const GeoArgs& a, hipStream_t st) {
    if (a.n_items == 0) return hipSuccess;
    k_geo_exact<<<(a.n_items + kGeoWaves - 1) / kGeoWaves, 64 * kGeoWaves, 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace hexo
