// geo.hip -- exact price of the geometric control variate of the Asian call, FP64 on gfx950.
//
// For an expiry's Asian functional A = sum_j w_j X_{t_j} (weights from the run schedule, sum W)
// the control variate is G = c e^L with c = W S0 and L = sum_j v_j ln(X_{t_j}/S0), v = w / W.
// Under the Heston model
//   phi(s) = E[e^{sL}] = exp(A_0 + B_0 v0)
// by a backward recursion over the grid points: at t_j the log-price coefficient a grows by s v_j;
// over (t_{j-1}, t_j] the Riccati equations of exp(A + a y + B V),
//   B' = sigma^2 B^2 / 2 + (rho sigma a - kappa) B + (a^2 - a) / 2,   A' = kappa theta B + r a,
// are solved in closed form from the running B (no cancellation as sigma -> 0, see geo_cf).
// The call price is Lewis' inversion
//   E[(c e^L - K)^+] = c (phi(1) - e^{k/2}/pi int_0^inf Re(e^{iuk} phi(1/2 - iu)) / (u^2 + 1/4) du),
// k = ln(K / c), by 64-node Gauss-Legendre panels of width h (host-chosen from the scale of L and
// the strikes) until the characteristic function has decayed below 1e-15 u0 on a whole panel
// starting at u0 (the integral's remaining tail is then below 1e-15).
//
// Work items are (request, expiry, block of <= 64 strikes), in up to three rounds. Pass 1 of a
// round evaluates the next p0 panels of every unconverged item at once, one wavefront per panel
// (lane q: node q, a recursion over the expiry's grid, a few hundred to a few thousand steps;
// lane i: strike i's panel sum over the 64 nodes in node order), plus E[e^L] in the first round;
// p0 is the host's estimate of where the characteristic function has decayed. Pass 2 gives each
// item one wavefront: lane i adds strike i's panel sums in panel order, keeps them for the next
// round if the decay was not reached, and in the last round evaluates further panels itself;
// every price is therefore the same fixed-order sum whatever p0 is. Runs once per batch, before
// the path kernels.
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

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) x = fmax(x, __shfl_xor(x, m));
    return x;  // every lane holds the wave's maximum
}

__device__ __forceinline__ uint32_t find_item(const uint32_t* end, uint32_t n, uint32_t t) {
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

// Panel p of an item, by one wavefront: lane q evaluates node q, u_q = (p + (x_q + 1)/2) h, and
// phi(1/2 - i u_q) weighted by w_q (h/2) / (u_q^2 + 1/4); the 64 values go through LDS (sh: this
// wave's 3 x 64 doubles) and lane i returns strike i's panel sum, node by node, of
// Re(e^{i u_q k_i} weighted phi). *mag is the wave's max |phi| over the panel.
// Panel 0 holds the integrand's 1/(u^2 + 1/4) peak at u = 0, which 64 nodes integrate to 1e-13
// over [0, 16] but only to 1e-4 over [0, 120]: it is summed as n_sub = min(64, ceil(h / 16))
// sub-panels of 64 nodes (wave-uniform, in order). The other panels keep width h.
__device__ __forceinline__ uint32_t first_panel_parts(double h) {
    return (uint32_t)fmin(64.0, ceil(h / 16.0));
}

__device__ double panel_sum(const GeoArgs& a, const DevGeo& G, const DevReq& R, const double* pts, uint32_t p, int lane,
                            double k, bool has_k, double (*sh)[64], double* mag) {
    const uint32_t n_sub = p == 0 ? first_panel_parts(G.h) : 1u;
    const double hs = G.h / (double)n_sub;
    double sum = 0.0, mg = 0.0;
    for (uint32_t j = 0; j < n_sub; ++j) {
        const double u = ((double)(p + j) + 0.5 * (a.gl[lane] + 1.0)) * hs;  // p = 0 or j = 0
        const Cplx ph = geo_cf(R, pts, G.n_pts, cx(0.5, -u));
        const double f = 0.5 * hs * a.gl[64 + lane] / (u * u + 0.25);
        sh[0][lane] = f * ph.re;
        sh[1][lane] = f * ph.im;
        sh[2][lane] = u;
        mg = fmax(mg, wave_max(cabs_(ph)));
        wave_sync_lds();
        if (has_k) {
            for (int q = 0; q < 64; ++q) {
                double sn, cs;
                sincos(sh[2][q] * k, &sn, &cs);
                sum += cs * sh[0][q] - sn * sh[1][q];
            }
        }
        wave_sync_lds();  // sh is rewritten by the next sub-panel or panel
    }
    *mag = mg;
    return sum;
}

__device__ __forceinline__ void strike_of(const GeoArgs& a, const DevGeo& G, int lane, double& K, double& k,
                                          bool& has_k) {
    const bool in = (uint32_t)lane < G.n_strikes;
    K = in ? a.opts[G.opt + (uint32_t)lane].k_abs : 0.0;
    has_k = K > 0.0;  // K <= 0: max(G - K, 0) = G - K needs no integral
    k = has_k ? log(K / G.c) : 0.0;
}

// Per-item state between rounds (kGeoState doubles): each strike's integral so far, E[e^L], the
// panels summed so far, whether the sum has converged and whether the last panel summed was below
// the stop rule's bound.
constexpr int kGeoState = 64 + 4;
constexpr int kGeoRounds = 3;  // parallel rounds of p0 panels before the serial continuation

// Pass 1 of round r: every wavefront evaluates one panel of one unconverged item (round r covers
// that item's panels done .. done + p0 - 1) and stores each strike's panel sum; in the first round the
// item's last wavefront evaluates E[e^L] instead.
__global__ __launch_bounds__(256) void k_geo_panels(GeoArgs a, uint32_t round) {
    __shared__ double sh[kGeoWaves][3][64];
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const uint32_t wv = blockIdx.x * kGeoWaves + (uint32_t)w;
    if (wv >= a.n_waves) return;  // whole wavefronts only
    const uint32_t item = __builtin_amdgcn_readfirstlane(find_item(a.wave_end, a.n_items, wv));
    const DevGeo& G = a.items[item];
    const DevReq& R = a.reqs[G.req];
    const double* pts = a.pts + 2 * (size_t)G.pts_begin;
    const uint32_t j = wv - G.wave_begin;
    const double* st = a.state + (size_t)item * kGeoState;
    uint32_t base = 0;
    if (round > 0) {
        if (j == G.p0 || st[66] != 0.0) return;  // E[e^L] is known; or the item has converged
        base = (uint32_t)st[65];
    } else if (j == G.p0) {
        const Cplx e1 = geo_cf(R, pts, G.n_pts, cx(1.0));
        if (lane == 0) a.panel_max[wv] = e1.re;
        return;
    }
    const uint32_t p = base + j;
    if (p >= G.max_panels) return;
    double K, k, mag;
    bool has_k;
    strike_of(a, G, lane, K, k, has_k);
    a.panel[(size_t)wv * 64 + (uint32_t)lane] = panel_sum(a, G, R, pts, p, lane, k, has_k, sh[w], &mag);
    if (lane == 0) a.panel_max[wv] = mag;
}

// Pass 2 of round r: one wavefront per item; lane i adds strike i's panel sums in panel order up
// to the second of the first two consecutive panels (after the first) whose |phi| stays below
// 1e-15 u0, u0 a panel's start: beyond u0 the integrand |phi| / (u^2 + 1/4) leaves at most
// sup|phi| / u0 to the integral, and two panels in a row guard against a dip of a non-monotone
// |phi|. An unconverged item keeps its sums for the next round; in the last round the wavefront
// evaluates further panels itself. The price is the same ordered sum whatever p0 and the number of
// rounds are.
__global__ __launch_bounds__(256) void k_geo_sum(GeoArgs a, uint32_t round) {
    __shared__ double sh[kGeoWaves][3][64];
    const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63u);
    const uint32_t item = blockIdx.x * kGeoWaves + (uint32_t)w;
    if (item >= a.n_items) return;  // whole wavefronts only
    const DevGeo& G = a.items[item];
    const DevReq& R = a.reqs[G.req];
    const double* pts = a.pts + 2 * (size_t)G.pts_begin;
    double* st = a.state + (size_t)item * kGeoState;
    if (round > 0 && st[66] != 0.0) return;
    double K, k;
    bool has_k;
    strike_of(a, G, lane, K, k, has_k);
    const double e1 = round == 0 ? a.panel_max[G.wave_begin + G.p0] : st[64];  // E[e^L]
    double integ = round == 0 ? 0.0 : st[lane];
    const uint32_t base = round == 0 ? 0u : (uint32_t)st[65];
    bool below = round == 0 ? false : st[67] != 0.0;  // the previous panel was below the bound
    bool converged = false;
    uint32_t p = base;
    for (uint32_t j = 0; j < G.p0 && p < G.max_panels && !converged; ++j, ++p) {
        integ += a.panel[(size_t)(G.wave_begin + j) * 64 + (uint32_t)lane];
        const bool now = p > 0 && a.panel_max[G.wave_begin + j] < 1e-15 * (double)p * G.h;
        converged = now && below;
        below = now;
    }
    const bool last = round + 1 == kGeoRounds;
    if (!converged && !last && p < G.max_panels) {
        st[lane] = integ;
        if (lane == 0) {
            st[64] = e1;
            st[65] = (double)p;
            st[66] = 0.0;
            st[67] = below ? 1.0 : 0.0;
        }
        return;
    }
    for (; p < G.max_panels && !converged; ++p) {
        double mag;
        integ += panel_sum(a, G, R, pts, p, lane, k, has_k, sh[w], &mag);
        const bool now = p > 0 && mag < 1e-15 * (double)p * G.h;  // wave-uniform
        converged = now && below;
        below = now;
    }
    if (lane == 0) st[66] = 1.0;  // later rounds skip the item
    if ((uint32_t)lane >= G.n_strikes) return;
    double price;
    if (!converged)
        price = __builtin_nan("");
    else if (has_k)
        price = G.c * (e1 - exp(0.5 * k) / kPi * integ);
    else
        price = G.c * e1 - K;  // max(G - K, 0) = G - K
    a.geo[G.out + (uint64_t)lane] = price;
}

}  // namespace

size_t geo_state_doubles(uint32_t n_items) { return (size_t)n_items * kGeoState; }

hipError_t launch_geo(const GeoArgs& a, hipStream_t st) {
    if (a.n_items == 0) return hipSuccess;
    for (uint32_t r = 0; r < (uint32_t)kGeoRounds; ++r) {
        k_geo_panels<<<(a.n_waves + kGeoWaves - 1) / kGeoWaves, 64 * kGeoWaves, 0, st>>>(a, r);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        k_geo_sum<<<(a.n_items + kGeoWaves - 1) / kGeoWaves, 64 * kGeoWaves, 0, st>>>(a, r);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace hexo
