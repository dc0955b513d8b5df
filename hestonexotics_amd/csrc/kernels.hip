// kernels.hip -- HIP kernels of the Heston QE Monte-Carlo path for gfx950.
//
// k_path_fp32  fused FP32 path kernel: Philox4x32-10 -> inverse normal -> QE step -> Asian/European
//              finalisation -> per-tile payoff sums. Replaces the reference's OpenMP path loop
//              and scheme (src/HSimulation.tpp:23-49, :52-86), the option policies
//              (src/inc/AsianContract.h:25-47, src/inc/VanillaContract.h:24-39) and the RNG ring
//              refills (src/RNG.cpp:28-43, src/as241.f90:85-118).
// k_path_fp64  the plain FP64 kernel: the same path in the reference's own FP64 form.
// k_reduce_*   cross-tile reduction in a fixed order and the estimator (price, standard error);
//              replaces the racy `prices[...] +=` of src/HSimulation.tpp:40,47-48.
//
// Work distribution: a persistent grid of independent wavefronts. Each wave takes a work item
// (request, PPT consecutive 64-path tiles) from an atomic queue; items are ordered longest request
// first. Lane l of the wave runs path tile*64 + l of each of its PPT tiles. At every expiry the
// wave transposes its paths' (A, X_T) through LDS so that lanes work over (option, 8-path group)
// pairs; each tile's sums therefore depend only on its 64 paths, never on the launch shape.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace hexo {

namespace {

constexpr uint32_t kEuropean = 2u;
constexpr uint32_t kAsianGeoCV = 4u;
constexpr uint32_t kGeoAsian = 8u;
constexpr uint32_t kTrapezoid = 4u;
constexpr uint32_t kMartingale = 8u;
constexpr uint32_t kCvRegression = 16u;
constexpr uint32_t kPrecFP64 = 0u;
constexpr uint32_t kPrecComp = 2u;

__device__ __forceinline__ void wave_sync_lds() {
    // LDS traffic of one wave is processed in order; this keeps the compiler from moving the
    // transpose's reads above its writes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Work item -> (sorted request position). item_end is a prefix sum over the sorted requests.
__device__ __forceinline__ uint32_t find_sorted(const uint32_t* item_end, uint32_t n, uint32_t item) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (item_end[mid] > item)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Takes the wave's next work item. Every lane executes the atomic (lane 0 adds 1, the others 0)
// and the wave keeps lane 0's ticket, so there is no divergent region around the atomic: with
// `if (lane == 0) atomicAdd(...)` followed by a broadcast, the loop structurizer may move the
// broadcast into a loop that lanes 1..63 re-enter without lane 0, repeating one item forever.
__device__ __forceinline__ uint32_t next_item(uint32_t* counter, int lane) {
    const uint32_t old = atomicAdd(counter, lane == 0 ? 1u : 0u);
    return uniform(old);
}

// Compensated (Neumaier) FP32 accumulation.
struct CompF {
    float s, c;
    __device__ __forceinline__ void add(float x) {
        const float t = s + x;
        c += (__builtin_fabsf(s) >= __builtin_fabsf(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ void merge(float s2, float c2) {
        const float t = s + s2;
        const float e = (__builtin_fabsf(s) >= __builtin_fabsf(s2)) ? (s - t) + s2 : (s2 - t) + s;
        s = t;
        c = c + c2 + e;
    }
};

// Per-expiry tile reduction. vals[j*64 + p] holds (A, X_T) of path p of sub-tile j (in the
// kernel's units); paths beyond n_paths hold values whose payoff is 0 (FP32: -inf; FP64: masked).
// Output o of the request is payoff-major: o = pay * n_opts + option. The summation order of a
// tile's 64 payoffs depends only on the number of strikes of the expiry (a property of the
// request), never on the launch shape:
//  * many strikes (>= kWideStrikes): lane i takes strike i and sums the 64 paths in path order,
//    reading each path's values as an LDS broadcast;
//  * few strikes: lane (o = lane/8, g = lane%8) sums paths 8g..8g+7 of output o in path order and
//    the 8 group sums are combined by a fixed xor tree.
constexpr uint32_t kWideStrikes = 48;

template <typename Acc>
struct TileSums;

template <>
struct TileSums<double> {  // FP64 sums (HEXO_FP32 and the FP64 kernel)
    double s = 0.0, ss = 0.0;
    __device__ __forceinline__ void add(double x) {
        s += x;
        ss = __builtin_fma(x, x, ss);
    }
    __device__ __forceinline__ void combine(const TileSums& o) {
        s += o.s;
        ss += o.ss;
    }
    __device__ __forceinline__ void poison() { s = ss = __builtin_nan(""); }
    __device__ __forceinline__ void load(const void* partials, size_t slot) {
        const double2 v = reinterpret_cast<const double2*>(partials)[slot];
        s = v.x;
        ss = v.y;
    }
    __device__ __forceinline__ void tree8() {
#pragma unroll
        for (int m = 1; m < 8; m <<= 1) {
            s += __shfl_xor(s, m);
            ss += __shfl_xor(ss, m);
        }
    }
    __device__ __forceinline__ void store(void* partials, size_t slot) const {
        reinterpret_cast<double2*>(partials)[slot] = make_double2(s, ss);
    }
};

template <>
struct TileSums<float> {  // compensated FP32 sums (HEXO_FP32_COMPENSATED)
    CompF s{0.f, 0.f}, ss{0.f, 0.f};
    __device__ __forceinline__ void add(float x) {
        s.add(x);
        ss.add(x * x);
    }
    __device__ __forceinline__ void combine(const TileSums& o) {
        s.merge(o.s.s, o.s.c);
        ss.merge(o.ss.s, o.ss.c);
    }
    __device__ __forceinline__ void poison() { s = ss = CompF{__builtin_nanf(""), 0.0f}; }
    __device__ __forceinline__ void load(const void* partials, size_t slot) {
        const float4 v = reinterpret_cast<const float4*>(partials)[slot];
        s = CompF{v.x, v.y};
        ss = CompF{v.z, v.w};
    }
    __device__ __forceinline__ void tree8() {
#pragma unroll
        for (int m = 1; m < 8; m <<= 1) {
            const float s2 = __shfl_xor(s.s, m), c2 = __shfl_xor(s.c, m);
            const float q2 = __shfl_xor(ss.s, m), d2 = __shfl_xor(ss.c, m);
            s.merge(s2, c2);
            ss.merge(q2, d2);
        }
    }
    __device__ __forceinline__ void store(void* partials, size_t slot) const {
        reinterpret_cast<float4*>(partials)[slot] = make_float4(s.s, s.c, ss.s, ss.c);
    }
};

// The level of the chunk tree ((t0 + t1) + (t2 + t3)) that N consecutive tiles of one chunk form
// (layout.hpp, kChunkTiles): t0 for one tile, t0 + t1 for two, the whole tree for four.
template <int N, typename Acc>
__device__ __forceinline__ TileSums<Acc> chunk_level(TileSums<Acc> (&t)[N]) {
    static_assert(N == 1 || N == 2 || N == kChunkTiles, "tiles per stored partial: 1, 2 or 4");
    if constexpr (N >= 2) t[0].combine(t[1]);
    if constexpr (N == 4) {
        t[2].combine(t[3]);
        t[0].combine(t[2]);
    }
    return t[0];
}

// Payoff kind of output block `pay` of a request: the pay-th set bit of its payoffs (outputs are
// payoff-major in bit order: Asian, European, Asian with the geometric control variate, the
// geometric call).
__device__ __forceinline__ uint32_t payoff_kind(uint32_t payoffs, uint32_t pay) {
    uint32_t b = payoffs;
    for (uint32_t i = 0; i < pay; ++i) b &= b - 1u;
    return b & (~b + 1u);
}

// Payoff of one path for strike K and payoff kind. Path values: the FP32 kernel's float2 (A, X_T)
// or float4 (A, X_T, G, -) in units of S0, paths beyond n_paths holding -inf; the FP64 kernel's
// double4 (A, X_T, G, -) with an explicit validity flag. The control-variate payoff is the
// arithmetic minus the geometric call payoff. max(., 0) turns a NaN value into 0: tiles holding a
// NaN value are marked by the path kernel instead (NanBits) and their sums set to NaN.
__device__ __forceinline__ float payoff_of(float2 v, uint32_t kind, float K, bool) {
    return __builtin_fmaxf((kind == kEuropean ? v.y : v.x) - K, 0.0f);
}
__device__ __forceinline__ float payoff_of(float4 v, uint32_t kind, float K, bool) {
    const float pg = __builtin_fmaxf(v.z - K, 0.0f);
    const float pa = __builtin_fmaxf((kind == kEuropean ? v.y : v.x) - K, 0.0f);
    return kind == kGeoAsian ? pg : kind == kAsianGeoCV ? pa - pg : pa;
}
__device__ __forceinline__ double payoff_of(double4 v, uint32_t kind, double K, bool valid) {
    const double pg = __builtin_fmax(v.z - K, 0.0);
    const double pa = __builtin_fmax((kind == kEuropean ? v.y : v.x) - K, 0.0);
    const double pay = kind == kGeoAsian ? pg : kind == kAsianGeoCV ? pa - pg : pa;
    return valid ? pay : 0.0;
}

// NaN marks of a wave's tiles at one expiry, wave-uniform: bit 3j + c is set when a path of tile j
// (within n_paths) has a NaN value in component c (0: A, 1: X_T, 2: G). A path whose price left the
// floating-point range (inf - inf in the finalisation) has one; the sums of every output of the
// tile that reads that component are then NaN, as the NaN payoff would have made them, instead of
// silently dropping the path.
struct NanBits {
    uint32_t bits = 0;
    __device__ __forceinline__ void mark(int j, bool valid, float A, float xt, float G) {
        const bool na = valid && A != A, nx = valid && xt != xt, ng = valid && G != G;
        bits |= (__builtin_amdgcn_ballot_w64(na) ? 1u : 0u) << (3 * j);
        bits |= (__builtin_amdgcn_ballot_w64(nx) ? 2u : 0u) << (3 * j);
        bits |= (__builtin_amdgcn_ballot_w64(ng) ? 4u : 0u) << (3 * j);
    }
    __device__ __forceinline__ void mark(int j, bool valid, double A, double xt, double G) {
        const bool na = valid && A != A, nx = valid && xt != xt, ng = valid && G != G;
        bits |= (__builtin_amdgcn_ballot_w64(na) ? 1u : 0u) << (3 * j);
        bits |= (__builtin_amdgcn_ballot_w64(nx) ? 2u : 0u) << (3 * j);
        bits |= (__builtin_amdgcn_ballot_w64(ng) ? 4u : 0u) << (3 * j);
    }
    // whether outputs of payoff `kind` of tile j read a NaN value
    __device__ __forceinline__ bool hit(int j, uint32_t kind) const {
        const uint32_t c = kind == kEuropean ? 2u : kind == kGeoAsian ? 4u : kind == kAsianGeoCV ? 5u : 1u;
        return ((bits >> (3 * j)) & c) != 0u;
    }
};

// Stores a wave's PPT tile sums of output o: the chunk tree's level for PPT tiles (1, 2 or 4), or,
// for 3 paths per thread (which does not divide a chunk), every tile's own sum, from which the
// group reduction builds the same tree.
template <int PPT, typename Acc>
__device__ __forceinline__ void store_tile_sums(const PathArgs& a, const DevReq& R, uint32_t tile0, uint32_t n_out,
                                                uint32_t o, bool active, TileSums<Acc> (&t)[PPT]) {
    if constexpr (kChunkTiles % PPT == 0) {
        const TileSums<Acc> c = chunk_level<PPT, Acc>(t);
        if (active) c.store(a.partials, (size_t)R.partial_base + (size_t)(tile0 / PPT) * n_out + o);
    } else {
#pragma unroll
        for (int j = 0; j < PPT; ++j)
            if (active && tile0 + (uint32_t)j < R.n_tiles)
                t[j].store(a.partials, (size_t)R.partial_base + (size_t)(tile0 + (uint32_t)j) * n_out + o);
    }
}

template <int PPT, typename ValT, typename Acc>
__device__ __forceinline__ void reduce_tiles_impl(const PathArgs& a, const DevReq& R, const DevFin& F,
                                                  uint32_t tile0, int lane, const ValT* vals, NanBits nan) {
    using KT = decltype(ValT{}.x);
    const uint32_t nk = F.n_strikes;
    const uint32_t n_out = R.n_opts * R.n_pay;
    for (uint32_t pay = 0; pay < R.n_pay; ++pay) {
        const uint32_t kind = payoff_kind(R.payoffs, pay);  // wave-uniform
        if (nk >= kWideStrikes) {
            for (uint32_t kb = 0; kb < nk; kb += kTile) {
                const uint32_t kj = kb + (uint32_t)lane;
                const bool active = kj < nk;
                const DevOpt& opt = a.opts[R.opt_begin + F.opt_begin + (active ? kj : 0u)];
                const KT K = sizeof(KT) == 4 ? (KT)opt.k_rel : (KT)opt.k_abs;
                const uint32_t o = pay * R.n_opts + F.opt_begin + kj;
                TileSums<Acc> t[PPT];
#pragma unroll
                for (int j = 0; j < PPT; ++j) {
                    const uint32_t tile = tile0 + (uint32_t)j;
                    if (tile >= R.n_tiles) continue;  // absent tiles are zero sums
                    const uint64_t nv = R.n_paths - (uint64_t)tile * kTile;
                    for (int p = 0; p < kTile; ++p)
                        t[j].add((Acc)payoff_of(vals[j * kTile + p], kind, K, (uint64_t)p < nv));
                    if (nan.hit(j, kind)) t[j].poison();
                }
                store_tile_sums<PPT, Acc>(a, R, tile0, n_out, o, active, t);
            }
        } else {
            const int g = lane & 7;
            for (uint32_t kb = 0; kb < nk; kb += 8) {
                const uint32_t kj = kb + (uint32_t)(lane >> 3);
                const bool active = kj < nk;
                const DevOpt& opt = a.opts[R.opt_begin + F.opt_begin + (active ? kj : 0u)];
                const KT K = sizeof(KT) == 4 ? (KT)opt.k_rel : (KT)opt.k_abs;
                const uint32_t o = pay * R.n_opts + F.opt_begin + kj;
                TileSums<Acc> t[PPT];
#pragma unroll
                for (int j = 0; j < PPT; ++j) {
                    const uint32_t tile = tile0 + (uint32_t)j;
                    if (tile >= R.n_tiles) continue;  // absent tiles are zero sums
                    const uint64_t nv = R.n_paths - (uint64_t)tile * kTile;
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int p = g * 8 + i;
                        t[j].add((Acc)payoff_of(vals[j * kTile + p], kind, K, (uint64_t)p < nv));
                    }
                    t[j].tree8();  // every lane of the 8-lane group ends with the same bits
                    if (nan.hit(j, kind)) t[j].poison();
                }
                store_tile_sums<PPT, Acc>(a, R, tile0, n_out, o, g == 0 && active, t);
            }
        }
    }
}

template <int PPT, typename ValT>
__device__ __forceinline__ void reduce_tiles(const PathArgs& a, const DevReq& R, const DevFin& F, uint32_t tile0,
                                             int lane, const ValT* vals, NanBits nan) {
    if (std::is_same<decltype(ValT{}.x), float>::value && R.precision == kPrecComp)
        reduce_tiles_impl<PPT, ValT, float>(a, R, F, tile0, lane, vals, nan);
    else
        reduce_tiles_impl<PPT, ValT, double>(a, R, F, tile0, lane, vals, nan);
}

__device__ __forceinline__ void count_flush(unsigned long long* counts, uint32_t c0, uint32_t c1, uint32_t c2,
                                            uint32_t c3, int lane) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        c0 += __shfl_xor(c0, m);
        c1 += __shfl_xor(c1, m);
        c2 += __shfl_xor(c2, m);
        c3 += __shfl_xor(c3, m);
    }
    // all lanes hold the wave total after the butterfly; only lane 0 contributes it (no divergent region)
    const bool first = lane == 0;
    atomicAdd(&counts[0], first ? (unsigned long long)c0 : 0ull);
    atomicAdd(&counts[1], first ? (unsigned long long)c1 : 0ull);
    atomicAdd(&counts[2], first ? (unsigned long long)c2 : 0ull);
    atomicAdd(&counts[3], first ? (unsigned long long)c3 : 0ull);
}

// ------------------------------------------------------------------------------------------
// FP32 path kernel
// ------------------------------------------------------------------------------------------
// No implicit FMA contraction in the FP32 path: every fused multiply-add below is written out.
// With contraction the compiler fuses a product into a later add only where both land in one basic
// block, which depends on how each paths-per-thread instance is laid out (one path per thread
// flattens the quadratic branch and turned tt - 1 into fma(m^2, 1/s2h, -1)), so the paths would
// change by an ulp with the launch shape.
#pragma clang fp contract(off)
struct SegF {
    float omd, th, c1h, c2h, g0, ka, kb, k3;
    float mA, ma, mh, mr;
};

// Per-wave LDS scratch of the FP32 kernel: the transpose of the per-path values at expiries, 4 words
// per path ((A, X_T) or, with the geometric payoffs, (A, X_T, G, -)).
template <int PPT>
struct WaveLds {
    static constexpr int kWords = 4 * PPT * kTile;  // 32-bit words
};

struct StepCount {
    uint32_t quad = 0, exp = 0, central = 0, tail = 0;
};

// The step's conditional moments from V (src/HSimulation.tpp:58-62): d0 = th - V, m = V + omd d0
// (exactly th at V = th, whatever the roundings), s2h = s2 / 2 (from the halved constants: the
// same bits as s2, scaled), m^2, tt = 2 / psi = m^2 / s2h and the branch: quadratic when
// psi < 1.5, i.e. tt > 4/3. tt is formed before the branch because the quadratic branch starts
// from it; the exponential branch does not use it.
template <int PPT>
__device__ __forceinline__ void qe_moments_fp32(const float (&V)[PPT], const SegF& c, float (&d0)[PPT],
                                                float (&m)[PPT], float (&s2h)[PPT], float (&m2)[PPT],
                                                float (&tt)[PPT], bool (&quad)[PPT]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        d0[j] = c.th - V[j];
        m[j] = __builtin_fmaf(c.omd, d0[j], V[j]);
        s2h[j] = __builtin_fmaf(V[j], c.c1h, c.c2h);
        m2[j] = m[j] * m[j];
        tt[j] = m2[j] * __builtin_amdgcn_rcpf(s2h[j]);  // finite by the FP32 validation bound
        quad[j] = tt[j] > 4.0f / 3.0f;
    }
}

// The update of one step given its moments and normals: the variance by the quadratic branch
// a (b + Z_V)^2 or the exponential branch, then the log-price step (MART: Andersen's
// martingale-corrected form, falling back per path and step to the plain step where
// E[exp(A V')] does not exist; GEO: also Y before the step and the run's sum of Y).
template <int PPT, bool MART, bool GEO>
__device__ __forceinline__ void qe_update_fp32(float (&V)[PPT], float (&Y)[PPT], float (&X)[PPT], float (&Xp)[PPT],
                                               float (&P)[PPT], float (&Yp)[PPT], float (&PY)[PPT],
                                               const uint32_t (&wv)[PPT], const float (&zv)[PPT],
                                               const float (&zx)[PPT], const float (&d0)[PPT], const float (&m)[PPT],
                                               const float (&s2h)[PPT], const float (&m2)[PPT], const float (&tt)[PPT],
                                               const bool (&quad)[PPT], const SegF& c) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        float dm;  // V' - m
        // MART: A m - ln E[exp(A V')] = fs f(ft) + corr with f(t) = t - ln(1 + t), in natural
        // log units; mok: the expectation exists
        float ft = 0.0f, fs = 0.0f, corr = 0.0f;
        bool mok = false;
        if (quad[j]) {
            const float tm1 = tt[j] - 1.0f;
            const float b2 = tm1 + __builtin_amdgcn_sqrtf(tt[j] * tm1);
            const float b = __builtin_amdgcn_sqrtf(b2);
            // V' = a (b + Z)^2 with a = m / (1 + b^2), taken as V' - m = a (Z (Z + 2 b) - 1): its
            // mean is 0 whatever the rounding of a and b, so the scheme's E[V'] = m holds exactly
            // (a biased mean would move the variance's mean-reversion level by (bias) / (kappa dt))
            const float aq = m[j] * __builtin_amdgcn_rcpf(__builtin_fmaf(b, b, 1.0f));
            // near b + Z = 0, m + (V' - m) can round a few ulp of m below 0 where a (b + Z)^2 is
            // >= 0: harmless to the moments, and the diffusion below takes |k3 (V + V')|
            dm = aq * __builtin_fmaf(zv[j], __builtin_fmaf(2.0f, b, zv[j]), -1.0f);
            if constexpr (MART) {
                // x = 2 A a < 1: A m - ln E[exp(A V')] = -(f(-x) + x^2 b^2 / (1 - x)) / 2
                const float x = (c.mA + c.mA) * aq;
                mok = x < 1.0f;
                ft = -x;
                fs = -0.5f;
                corr = -0.5f * (x * x) * b2 * __builtin_amdgcn_rcpf(1.0f - x);
            }
        } else {
            // p = (psi-1)/(psi+1), beta = 2/(m (psi+1)); V' = ln((1-p)/(1-U))/beta if U > p, else 0.
            // With sm = s2 + m^2: 1-p = 2m^2/sm and 1/beta = sm/(2m).
            const float sm = __builtin_fmaf(2.0f, s2h[j], m2[j]);
            // (1-U) sm, 1 - U from ~k; the complement is taken here (asm: not hoisted in front of the
            // branch, where every lane would pay for it; the wait state as at the asm helpers in
            // device.hpp)
            uint32_t nk;
            asm("s_nop 0\n\tv_not_b32 %0, %1" : "=v"(nk) : "v"(wv[j]));
            const float lhs = uniform_f32(nk) * sm;
            const float tm2 = m2[j] + m2[j];
            const float lg = __builtin_amdgcn_logf(tm2 * rcp_refined(lhs));
            const float ib = sm * rcp_refined(m[j]);  // 2 / beta
            const float v = ib * (0.5f * 0.693147180559945309f) * lg;
            // U > p exactly when (1-p)/(1-U) > 1, i.e. when the log is positive: V' = max(v, 0)
            dm = __builtin_fmaxf(v, 0.0f) - m[j];
            if constexpr (MART) {
                // cA = A / beta < 1, r1 = (1-p) / (1 - cA) = 2 m^2 / (sm (1 - cA)), y = r1 cA:
                // A m - ln E[exp(A V')] = f(y) - r1 cA^2
                const float cA = (0.5f * c.mA) * ib;
                mok = cA < 1.0f;
                const float r1 = tm2 * __builtin_amdgcn_rcpf(__builtin_fmaf(-cA, sm, sm));
                ft = r1 * cA;
                fs = 1.0f;
                corr = -ft * cA;
            }
        }
        // The reference's log-price step multiplies V' - V by rho/sigma and V + V' - 2 th by
        // kappa rho dt / (2 sigma). As differences of FP32 variances they would carry ulp(V) errors
        // amplified by 1/sigma (prices tens of percent off at sigma = 1e-6, or at sigma = 1e-4
        // with dt = 2e-7). Written on V' - m and th - V, each term has its own size:
        // K1 V + K2 V' = (K1 + K2) th + K2 (V' - m) - (K1 + K2 e^{-kappa dt}) (th - V), the first
        // part in g0, the coefficients ka = K2 and kb (O(dt), formed in FP64) in the segment.
        const float Vn = m[j] + dm;
        const float sum = V[j] + Vn;
        float inc = __builtin_fmaf(c.ka, dm, __builtin_fmaf(c.kb, d0[j], c.g0));
        if constexpr (MART) {
            // f(t) = t - ln(1 + t): its series through t^7 for |t| < 1/16 (relative error < 2e-8),
            // else from the rounded 1 + t. The direct form alone would return f = t wherever 1 + t
            // rounds to 1, a one-signed error of up to 6e-8 per step where t stays small.
            const float t2 = ft * ft;
            float h = __builtin_fmaf(ft, -1.0f / 7.0f, 1.0f / 6.0f);
            h = __builtin_fmaf(ft, h, -0.2f);
            h = __builtin_fmaf(ft, h, 0.25f);
            h = __builtin_fmaf(ft, h, -1.0f / 3.0f);
            h = __builtin_fmaf(ft, h, 0.5f);
            const float fdir = __builtin_fmaf(-0.693147180559945309f, __builtin_amdgcn_logf(1.0f + ft), ft);
            const float f = __builtin_fabsf(ft) < 0.0625f ? t2 * h : fdir;
            const float cl = __builtin_fmaf(fs, f, corr) * 1.44269504088896341f;
            const float im = __builtin_fmaf(c.ma, dm, __builtin_fmaf(-c.mh, sum, c.mr)) + cl;
            inc = mok ? im : inc;
        }
        const float sq = __builtin_amdgcn_sqrtf(__builtin_fabsf(c.k3 * sum));
        if constexpr (GEO) Yp[j] = Y[j];
        Y[j] = __builtin_fmaf(sq, zx[j], Y[j] + inc);
        Xp[j] = X[j];
        X[j] = __builtin_amdgcn_exp2f(Y[j]);
        P[j] += X[j];
        if constexpr (GEO) PY[j] += Y[j];
        V[j] = Vn;
    }
}

// Branch statistics (COUNT): quad / exp per step; central / tail per normal drawn (Z_V only where
// the variance branch is quadratic).
__device__ __forceinline__ void count_step(StepCount& cnt, bool valid, bool quad, bool tv, bool tx) {
    if (!valid) return;
    cnt.quad += quad ? 1u : 0u;
    cnt.exp += quad ? 0u : 1u;
    cnt.central += (quad && !tv ? 1u : 0u) + (tx ? 0u : 1u);
    cnt.tail += (quad && tv ? 1u : 0u) + (tx ? 1u : 0u);
}

// One QE step for the PPT paths of a lane (Andersen's QE scheme as in src/HSimulation.tpp:52-86):
// the normals Z_V (quadratic lanes only: the exponential branch reads the word as a uniform) and
// Z_X, then the update. Y is log2(X/S0); X = 2^Y; P accumulates X for the Asian sum.
template <int PPT, bool COUNT, bool MART, bool GEO>
__device__ __forceinline__ void qe_step_fp32(float (&V)[PPT], float (&Y)[PPT], float (&X)[PPT], float (&Xp)[PPT],
                                             float (&P)[PPT], float (&Yp)[PPT], float (&PY)[PPT],
                                             const uint32_t (&wv)[PPT], const uint32_t (&wx)[PPT],
                                             const SegF& c, const bool (&valid)[PPT], StepCount& cnt,
                                             const NormalVregs& nv) {
    float d0[PPT], m[PPT], s2h[PPT], m2[PPT], tt[PPT], zv[PPT], zx[PPT];
    bool quad[PPT];
    qe_moments_fp32<PPT>(V, c, d0, m, s2h, m2, tt, quad);
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        bool tv, tx;
        // Z_V with its tail polynomial wherever the word needs it, also on the rare exponential-
        // branch lanes that do not read it: the branch on the quadratic test then comes once, at
        // the variance update, instead of also in front of the tail polynomial
        zv[j] = normal_word_f32(wv[j], true, nv, tv);
        zx[j] = normal_word_f32(wx[j], true, nv, tx);
        if constexpr (COUNT) count_step(cnt, valid[j], quad[j], tv, tx);
    }
    qe_update_fp32<PPT, MART, GEO>(V, Y, X, Xp, P, Yp, PY, wv, zv, zx, d0, m, s2h, m2, tt, quad, c);
}

// Philox4x32-10 (device.hpp philox4x32_10) of the path counter {blk, plo, phi, 0} with the step
// pair blk, the path's high word phi and the key wave-uniform. The rounds are the same; the values
// that stay wave-uniform through the first three rounds are marked so (readfirstlane), which keeps
// their products on the scalar unit: both of round 1's, one each of rounds 2 and 3.
__device__ __forceinline__ u32x4 philox_path(uint32_t blk, uint32_t plo, uint32_t phi, uint32_t k0, uint32_t k1) {
    // round 1: c0 = blk, c2 = phi, c3 = 0 uniform
    uint64_t p0 = (uint64_t)kPhiloxM0 * blk, p1 = (uint64_t)kPhiloxM1 * phi;
    uint32_t c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), plo, k0, 0x96);
    uint32_t c1 = uniform((uint32_t)p1);
    uint32_t c2 = uniform((uint32_t)(p0 >> 32)) ^ uniform(k1);
    uint32_t c3 = uniform((uint32_t)p0);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 2: c2 uniform; the new c0 and c1 are uniform
    p0 = (uint64_t)kPhiloxM0 * c0;
    p1 = (uint64_t)kPhiloxM1 * c2;
    c0 = uniform((uint32_t)(p1 >> 32)) ^ c1 ^ uniform(k0);
    c1 = uniform((uint32_t)p1);
    c2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c3 = (uint32_t)p0;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
    // round 3: c0 uniform; the new c3 is uniform
    p0 = (uint64_t)kPhiloxM0 * c0;
    p1 = (uint64_t)kPhiloxM1 * c2;
    c0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
    c1 = (uint32_t)p1;
    c2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
    c3 = uniform((uint32_t)p0);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
#pragma unroll
    for (int r = 3; r < 10; ++r) {
        p0 = (uint64_t)kPhiloxM0 * c0;
        p1 = (uint64_t)kPhiloxM1 * c2;
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), c3, k1, 0x96);
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return u32x4{c0, c1, c2, c3};
}

template <int PPT>
__device__ __forceinline__ void philox_block(uint32_t blk, const uint32_t (&plo)[PPT], const uint32_t (&phi)[PPT],
                                             uint32_t key0, uint32_t key1, uint32_t (&w0)[PPT], uint32_t (&w1)[PPT],
                                             uint32_t (&w2)[PPT], uint32_t (&w3)[PPT]) {
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const u32x4 r = philox_path(blk, plo[j], phi[j], key0, key1);
        w0[j] = r.x;
        w1[j] = r.y;
        w2[j] = r.z;
        w3[j] = r.w;
    }
}

template <int PPT, bool DEBUG, bool COUNT, bool MART, bool GEO>
__device__ __forceinline__ void path_fp32(const PathArgs& a, uint32_t ri, uint32_t tile0, int lane, uint32_t* lds) {
    const DevReq& R = a.reqs[ri];
    const uint64_t n_paths = R.n_paths;
    const uint32_t key0 = R.seed_lo, key1 = R.seed_hi;
    float V[PPT], Y[PPT], X[PPT], Xp[PPT], P[PPT], acc[PPT], Xs[PPT];
    float Yp[PPT], PY[PPT], accY[PPT], Ys[PPT];  // GEO: the same accumulation on Y = log2(X/S0)
    uint32_t plo[PPT], phi[PPT], sv[PPT], sx[PPT];
    bool valid[PPT];
    StepCount cnt;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const uint64_t pl = (uint64_t)(tile0 + (uint32_t)j) * kTile + (uint64_t)lane;
        valid[j] = pl < n_paths;
        const uint64_t gp = R.path_offset + pl;
        plo[j] = (uint32_t)gp;
        // path_offset is a multiple of 64, so a tile never straddles 2^32: the high word is
        // wave-uniform and the first Philox rounds run partly on the scalar unit
        phi[j] = uniform((uint32_t)((R.path_offset + (uint64_t)(tile0 + (uint32_t)j) * kTile) >> 32));
        V[j] = R.v0_f;
        Y[j] = 0.0f;
        X[j] = 1.0f;
        Xp[j] = 1.0f;
        P[j] = 0.0f;
        acc[j] = 0.0f;
        Xs[j] = 1.0f;
        Yp[j] = PY[j] = accY[j] = Ys[j] = 0.0f;
        sv[j] = sx[j] = 0u;
    }
    const NormalVregs nv = normal_vregs();
    uint32_t k = 0;  // global step index; step k uses words 2k, 2k+1 of the path's stream
    const uint32_t n_segs = R.n_segs;
    for (uint32_t s = 0; s < n_segs; ++s) {
        const DevSeg& S = a.segs[R.seg_begin + s];
        const SegF c{S.omd, S.th, S.c1h, S.c2h, S.g0, S.ka, S.kb, S.k3, S.mA, S.ma, S.mh, S.mr};
        const uint32_t n = S.n_steps;
        uint32_t i = 0;
        if ((k & 1u) && n > 0) {  // second half of a block opened by the previous run
            qe_step_fp32<PPT, COUNT, MART, GEO>(V, Y, X, Xp, P, Yp, PY, sv, sx, c, valid, cnt, nv);
            ++i;
            ++k;
        }
        for (; i + 2 <= n; i += 2, k += 2) {
            uint32_t w0[PPT], w1[PPT], w2[PPT], w3[PPT];
            philox_block<PPT>(k >> 1, plo, phi, key0, key1, w0, w1, w2, w3);
            qe_step_fp32<PPT, COUNT, MART, GEO>(V, Y, X, Xp, P, Yp, PY, w0, w1, c, valid, cnt, nv);
            qe_step_fp32<PPT, COUNT, MART, GEO>(V, Y, X, Xp, P, Yp, PY, w2, w3, c, valid, cnt, nv);
        }
        if (i < n) {
            uint32_t w0[PPT], w1[PPT];
            philox_block<PPT>(k >> 1, plo, phi, key0, key1, w0, w1, sv, sx);
            qe_step_fp32<PPT, COUNT, MART, GEO>(V, Y, X, Xp, P, Yp, PY, w0, w1, c, valid, cnt, nv);
            ++k;
        }
        // Asian accumulation of the run's interior steps b..e-1:
        // delta * (X_{b-1}/2 + X_b + ... + X_{e-2} + X_{e-1}/2) = delta * (P - X_e + (Xs - Xp)/2)
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            acc[j] = __builtin_fmaf(S.delta_f, (P[j] - X[j]) + 0.5f * (Xs[j] - Xp[j]), acc[j]);
            if constexpr (GEO) accY[j] = __builtin_fmaf(S.delta_f, (PY[j] - Y[j]) + 0.5f * (Ys[j] - Yp[j]), accY[j]);
        }
        // per-path values at an expiry: (A, X_T), with GEO (A, X_T, G, -); the wave's LDS region
        // holds 4 words per path either way
        using Val = typename std::conditional<GEO, float4, float2>::type;
        Val* vals = reinterpret_cast<Val*>(lds);
        for (uint32_t f = 0; f < S.fin_count; ++f) {
            const DevFin& F = a.fins[S.fin_begin + f];
            const float w = F.w_f, invT = F.invT_f;
            NanBits nan;
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const float dX = X[j] - Xp[j];
                const float xt = __builtin_fmaf(dX, w, Xp[j]);
                const float A = (R.flags & kTrapezoid) ? (acc[j] + F.hwdt_f * (Xp[j] + xt)) * invT
                                                       : __builtin_fmaf(dX, w, acc[j]) * invT;
                // paths beyond n_paths carry -inf: every payoff max(0, . - K) is then 0
                if constexpr (GEO) {
                    // the same functional on Y, normalised by its weight sum W: G/S0 = W 2^(L/W)
                    const float dY = Y[j] - Yp[j];
                    const float L = (R.flags & kTrapezoid)
                                        ? (accY[j] + F.hwdt_f * (Yp[j] + __builtin_fmaf(dY, w, Yp[j]))) * invT
                                        : __builtin_fmaf(dY, w, accY[j]) * invT;
                    const float G = F.gW_f * __builtin_amdgcn_exp2f(L * F.gInvW_f);
                    vals[j * kTile + lane] =
                        valid[j] ? make_float4(A, xt, G, 0.0f) : make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
                    nan.mark(j, valid[j], A, xt, G);
                } else {
                    vals[j * kTile + lane] = valid[j] ? make_float2(A, xt) : make_float2(-INFINITY, -INFINITY);
                    nan.mark(j, valid[j], A, xt, 0.0f);
                }
                if constexpr (DEBUG) {
                    if (valid[j] && a.finals) {
                        const uint64_t pl = (uint64_t)(tile0 + (uint32_t)j) * kTile + (uint64_t)lane;
                        double* out = a.finals + (pl * R.n_exp + F.chain) * 2;
                        out[0] = (double)A * R.S0;
                        out[1] = (double)xt * R.S0;
                    }
                }
            }
            wave_sync_lds();
            reduce_tiles<PPT>(a, R, F, tile0, lane, vals, nan);
            wave_sync_lds();
        }
        const float hn = 0.5f * S.dnext_f;
#pragma unroll
        for (int j = 0; j < PPT; ++j) {
            acc[j] = __builtin_fmaf(hn, X[j] + Xp[j], acc[j]);
            Xs[j] = X[j];
            P[j] = 0.0f;
            if constexpr (GEO) {
                accY[j] = __builtin_fmaf(hn, Y[j] + Yp[j], accY[j]);
                Ys[j] = Y[j];
                PY[j] = 0.0f;
            }
        }
    }
    if constexpr (COUNT) count_flush(a.counts, cnt.quad, cnt.exp, cnt.central, cnt.tail, lane);
}

template <int PPT, bool DEBUG, bool COUNT, bool MART, bool GEO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PPT == 2 ? 5 : 4, 8))) void k_path_fp32(PathArgs a) {
    extern __shared__ uint32_t lds_u32[];
    const int lane = (int)(threadIdx.x & 63u);
    // the wave's LDS region, at a wave-uniform address
    uint32_t* lds = lds_u32 + uniform(threadIdx.x >> 6) * WaveLds<PPT>::kWords;
    for (;;) {
        const uint32_t item = next_item(a.counter, lane);
        if (item >= a.total_items) break;
        const uint32_t sp = uniform(find_sorted(a.item_end, a.n_sorted, item));
        const uint32_t first = sp ? a.item_end[sp - 1] : 0u;
        const uint32_t ri = uniform(a.order[sp]);
        path_fp32<PPT, DEBUG, COUNT, MART, GEO>(a, ri, uniform((item - first) * (uint32_t)PPT), lane, lds);
    }
}
#pragma clang fp contract(fast)
// ------------------------------------------------------------------------------------------
// Plain FP64 path kernel: the reference's step and policies in FP64, recomputing the step's
// moments and K0..K4 every step exactly as written in src/HSimulation.tpp:52-86.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ bool qe_step_fp64(const DevReq& R, double delta, double& V, double& logX, uint32_t wv,
                                             uint32_t wx) {
    const double theta = R.theta, rho = R.rho, kappa = R.kappa, eps = R.sigma;
    const double e = ::exp(-kappa * delta);
    const double m = theta + (V - theta) * e;
    const double s2 = __builtin_fabs(V * eps * eps * e / kappa * (1.0 - e) +
                                     theta * eps * eps / (2.0 * kappa) * (1.0 - e) * (1.0 - e));
    const double psi = s2 / (m * m);
    const double Vp = V;
    const bool quad = psi < 1.5;
    double aq = 0.0, bp2 = 0.0, p = 0.0, beta = 0.0;
    if (quad) {
        bp2 = 2.0 / psi - 1.0 + ::sqrt(2.0 / psi * (2.0 / psi - 1.0));
        const double b = ::sqrt(bp2);
        aq = m / (1.0 + bp2);
        const double zv = normal_f64(wv);
        V = aq * (b + zv) * (b + zv);
    } else {
        p = (psi - 1.0) / (psi + 1.0);
        beta = 2.0 / (m * (psi + 1.0));
        const double u = uniform_f64(wv);
        V = p < u ? ::log((1.0 - p) / (1.0 - u)) / beta : 0.0;
    }
    double k0 = -rho * kappa * theta / eps * delta + R.r_drift * delta;
    const double k1 = 0.5 * delta * (kappa * rho / eps - 0.5) - rho / eps;
    const double k2 = 0.5 * delta * (kappa * rho / eps - 0.5) + rho / eps;
    const double k3 = 0.5 * delta * (1.0 - rho * rho);
    const double k4 = 0.5 * delta * (1.0 - rho * rho);
    if (R.flags & kMartingale) {
        // Andersen's martingale-corrected K0* = -ln E[exp(A V') | V] - (K1 + K3/2) V, A = K2 + K4/2
        const double A = k2 + 0.5 * k4;
        if (quad && 2.0 * A * aq < 1.0) {
            const double lnM = A * bp2 * aq / (1.0 - 2.0 * A * aq) - 0.5 * ::log1p(-2.0 * A * aq);
            k0 = -lnM - (k1 + 0.5 * k3) * Vp + R.r_drift * delta;
        } else if (!quad && A < beta) {
            const double lnM = ::log(p + beta * (1.0 - p) / (beta - A));
            k0 = -lnM - (k1 + 0.5 * k3) * Vp + R.r_drift * delta;
        }
    }
    logX = logX + k0 + k1 * Vp + k2 * V + ::sqrt(k3 * Vp + k4 * V) * normal_f64(wx);
    return quad;
}

__device__ __forceinline__ bool central_f64(uint32_t k) { return __builtin_fabs(uniform_f64(k) - 0.5) <= 0.425; }

template <bool DEBUG, bool COUNT>
__device__ __forceinline__ void path_fp64(const PathArgs& a, uint32_t ri, uint32_t tile, int lane, double4* vals) {
    const DevReq& R = a.reqs[ri];
    const bool geo = (R.payoffs & (kAsianGeoCV | kGeoAsian)) != 0;  // wave-uniform
    double accY = 0.0, yp = 0.0;  // the Asian accumulation on y = ln(X/S0); y before the last step
    const uint64_t pl = (uint64_t)tile * kTile + (uint64_t)lane;
    const bool valid = pl < R.n_paths;
    const uint64_t gp = R.path_offset + pl;
    const uint32_t plo = (uint32_t)gp, phi = uniform((uint32_t)((R.path_offset + (uint64_t)tile * kTile) >> 32));
    double V = R.v0, logX = R.log_S0, X = R.S0, Xp = R.S0, acc = 0.0;
    uint32_t wb0 = 0, wb1 = 0, k = 0;
    uint32_t cq = 0, ce = 0, cc = 0, ct = 0;
    for (uint32_t s = 0; s < R.n_segs; ++s) {
        const DevSeg& S = a.segs[R.seg_begin + s];
        const double delta = S.delta;
        const uint32_t n = S.n_steps;
        for (uint32_t i = 0; i < n; ++i, ++k) {
            uint32_t wv, wx;
            if ((k & 1u) == 0u) {
                const u32x4 w = philox4x32_10(k >> 1, plo, phi, 0u, R.seed_lo, R.seed_hi);
                wv = w.x;
                wx = w.y;
                wb0 = w.z;
                wb1 = w.w;
            } else {
                wv = wb0;
                wx = wb1;
            }
            yp = logX - R.log_S0;
            const bool quad = qe_step_fp64(R, delta, V, logX, wv, wx);
            Xp = X;
            X = ::exp(logX);
            if constexpr (COUNT) {
                if (valid) {
                    cq += quad ? 1u : 0u;
                    ce += quad ? 0u : 1u;
                    const bool cv = central_f64(wv), cx = central_f64(wx);
                    cc += (quad && cv ? 1u : 0u) + (cx ? 1u : 0u);
                    ct += (quad && !cv ? 1u : 0u) + (cx ? 0u : 1u);
                }
            }
            if (i + 1u < n) {
                acc += delta * 0.5 * (X + Xp);  // AAsianCallNonAdaptive::accumulate_value
                if (geo) accY += delta * 0.5 * ((logX - R.log_S0) + yp);
            }
        }
        for (uint32_t f = 0; f < S.fin_count; ++f) {
            const DevFin& F = a.fins[S.fin_begin + f];
            const double xt = Xp + (X - Xp) * (F.T - F.tp) / F.dt;
            const double A = (R.flags & kTrapezoid) ? (acc + (F.T - F.tp) * 0.5 * (Xp + xt)) / F.T
                                                    : (acc + (X - Xp) * (F.T - F.tp) / F.dt) / F.T;
            double G = 0.0;
            if (geo) {
                // the same functional on y, normalised by its weight sum W: G = W S0 exp(L / W)
                const double y = logX - R.log_S0, wf = (F.T - F.tp) / F.dt;
                const double L = (R.flags & kTrapezoid) ? (accY + (F.T - F.tp) * 0.5 * (yp + (yp + (y - yp) * wf))) / F.T
                                                        : (accY + (y - yp) * wf) / F.T;
                G = F.gW * R.S0 * ::exp(L / F.gW);
            }
            vals[lane] = make_double4(A, xt, G, 0.0);
            NanBits nan;
            nan.mark(0, valid, A, xt, G);
            if constexpr (DEBUG) {
                if (valid && a.finals) {
                    double* out = a.finals + (pl * R.n_exp + F.chain) * 2;
                    out[0] = A;
                    out[1] = xt;
                }
            }
            wave_sync_lds();
            reduce_tiles<1>(a, R, F, tile, lane, vals, nan);
            wave_sync_lds();
        }
        acc += S.dnext * 0.5 * (X + Xp);
        if (geo) accY += S.dnext * 0.5 * ((logX - R.log_S0) + yp);
    }
    if constexpr (COUNT) count_flush(a.counts, cq, ce, cc, ct, lane);
}

template <bool DEBUG, bool COUNT>
__global__ __launch_bounds__(256) void k_path_fp64(PathArgs a) {
    extern __shared__ double4 lds_d4[];
    const int lane = (int)(threadIdx.x & 63u);
    double4* vals = lds_d4 + (threadIdx.x >> 6) * kTile;
    for (;;) {
        const uint32_t item = next_item(a.counter, lane);
        if (item >= a.total_items) break;
        const uint32_t sp = uniform(find_sorted(a.item_end, a.n_sorted, item));
        const uint32_t first = sp ? a.item_end[sp - 1] : 0u;
        const uint32_t ri = uniform(a.order[sp]);
        path_fp64<DEBUG, COUNT>(a, ri, uniform(item - first), lane, vals);
    }
}

// ------------------------------------------------------------------------------------------
// Reduction: pass A sums each 64-tile group of a (request, output) in tile order; pass B sums
// the groups in order and forms the estimator. Both orders depend only on the request.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t find_req(const uint64_t* end, uint32_t n, uint64_t t) {
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

// Stored partial `slot` of output o, or a zero sum past the request's last stored one.
template <typename Acc>
__device__ __forceinline__ TileSums<Acc> stored_sum(const ReduceArgs& a, const DevReq& R, uint32_t n_out,
                                                    uint32_t n_slots, uint32_t slot, uint32_t o) {
    TileSums<Acc> t;
    if (slot < n_slots) t.load(a.partials, (size_t)R.partial_base + (size_t)slot * n_out + o);
    return t;
}

// Sums the chunks of one 64-tile group in chunk order. Each chunk's sum completes the tree
// ((t0 + t1) + (t2 + t3)) from the level the path kernel stored (store_tiles tiles per partial),
// so a chunk's bits do not depend on the paths per thread the request ran with.
template <typename Acc>
__device__ __forceinline__ TileSums<Acc> group_sum(const ReduceArgs& a, const DevReq& R, uint32_t n_out, uint32_t g,
                                                   uint32_t o) {
    const uint32_t st = R.store_tiles;
    const uint32_t n_slots = (R.n_tiles + st - 1) / st;
    const uint32_t t0 = g * kGroupTiles;
    const uint32_t t1 = min(t0 + (uint32_t)kGroupTiles, R.n_tiles);
    TileSums<Acc> acc;
    for (uint32_t c = t0 / kChunkTiles; c < (t1 + kChunkTiles - 1) / kChunkTiles; ++c) {
        TileSums<Acc> chunk;
        if (st == kChunkTiles) {
            chunk = stored_sum<Acc>(a, R, n_out, n_slots, c, o);
        } else if (st == 2) {
            TileSums<Acc> t[2] = {stored_sum<Acc>(a, R, n_out, n_slots, 2 * c, o),
                                  stored_sum<Acc>(a, R, n_out, n_slots, 2 * c + 1, o)};
            t[0].combine(t[1]);
            chunk = t[0];
        } else {
            TileSums<Acc> t[kChunkTiles];
#pragma unroll
            for (int i = 0; i < kChunkTiles; ++i) t[i] = stored_sum<Acc>(a, R, n_out, n_slots, kChunkTiles * c + i, o);
            chunk = chunk_level<kChunkTiles, Acc>(t);
        }
        acc.combine(chunk);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_reduce_groups(ReduceArgs a, uint64_t total) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ri = find_req(a.group_end, a.n_req, t);
    const uint64_t local = t - (ri ? a.group_end[ri - 1] : 0ull);
    const DevReq& R = a.reqs[ri];
    const uint32_t n_out = R.n_opts * R.n_pay;
    const uint32_t g = (uint32_t)(local / n_out);
    const uint32_t o = (uint32_t)(local - (uint64_t)g * n_out);
    const size_t gslot = (size_t)R.group_base + (size_t)g * n_out + o;
    if (R.precision == kPrecComp)
        group_sum<float>(a, R, n_out, g, o).store(a.groups, gslot);
    else
        group_sum<double>(a, R, n_out, g, o).store(a.groups, gslot);
}

// Sum and sum of squares of output o over the request's groups, on one wavefront: lane l sums
// groups l, l + 64, l + 128, ... in order and the 64 lane sums are combined by a fixed xor
// butterfly (every lane ends with the same bits). The order depends only on the request's group
// count, never on the launch.
__device__ __forceinline__ void output_sums(const ReduceArgs& a, const DevReq& R, uint32_t o, uint32_t lane,
                                            double& s, double& ss) {
    const uint32_t n_out = R.n_opts * R.n_pay;
    if (R.precision == kPrecComp) {
        const float4* p = reinterpret_cast<const float4*>(a.groups);
        CompF cs{0.f, 0.f}, css{0.f, 0.f};
        for (uint32_t g = lane; g < R.n_groups; g += kTile) {
            const float4 v = p[(size_t)R.group_base + (size_t)g * n_out + o];
            cs.merge(v.x, v.y);
            css.merge(v.z, v.w);
        }
#pragma unroll
        for (int m = 1; m < kTile; m <<= 1) {
            const float s2 = __shfl_xor(cs.s, m), c2 = __shfl_xor(cs.c, m);
            const float q2 = __shfl_xor(css.s, m), d2 = __shfl_xor(css.c, m);
            cs.merge(s2, c2);
            css.merge(q2, d2);
        }
        s = (double)cs.s + (double)cs.c;
        ss = (double)css.s + (double)css.c;
    } else {
        const double2* p = reinterpret_cast<const double2*>(a.groups);
        s = 0.0;
        ss = 0.0;
        for (uint32_t g = lane; g < R.n_groups; g += kTile) {
            const double2 v = p[(size_t)R.group_base + (size_t)g * n_out + o];
            s += v.x;
            ss += v.y;
        }
#pragma unroll
        for (int m = 1; m < kTile; m <<= 1) {
            s += __shfl_xor(s, m);
            ss += __shfl_xor(ss, m);
        }
    }
}

// Sample mean and variance (n - 1 denominator, clamped at 0) from a sum and a sum of squares.
__device__ __forceinline__ void mean_var(double s, double ss, double n, double& mean, double& var) {
    mean = s / n;
    var = n > 1.0 ? (ss - n * mean * mean) / (n - 1.0) : 0.0;
    var = var > 0.0 ? var : 0.0;
}

// One wavefront per output. A control-variate output adds the exact geometric price to the mean of
// A - G; with kCvRegression it is A - beta (G - E[G]) instead, beta = Cov(A, G) / Var G from the
// same paths' Asian (block 0) and geometric outputs of the option: Cov from
// Var(A - G) = Var A + Var G - 2 Cov(A, G), variance Var A - beta Cov(A, G).
__global__ __launch_bounds__(256) void k_reduce_final(ReduceArgs a, uint64_t total) {
    const uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / (uint64_t)kTile;  // wave-uniform
    const uint32_t lane = threadIdx.x % (uint32_t)kTile;
    if (t >= total) return;
    const uint32_t ri = find_req(a.out_end, a.n_req, t);
    const uint32_t o = (uint32_t)(t - (ri ? a.out_end[ri - 1] : 0ull));
    const DevReq& R = a.reqs[ri];
    const uint32_t opt = o % R.n_opts;
    const bool cv = a.geo && payoff_kind(R.payoffs, o / R.n_opts) == kAsianGeoCV;
    const bool reg = cv && (R.flags & kCvRegression);  // request-uniform
    double s, ss, sa = 0.0, ssa = 0.0, sg = 0.0, ssg = 0.0;
    output_sums(a, R, o, lane, s, ss);
    if (reg) {
        output_sums(a, R, opt, lane, sa, ssa);
        output_sums(a, R, (uint32_t)__popc(R.payoffs & (kGeoAsian - 1u)) * R.n_opts + opt, lane, sg, ssg);
    }
    if (lane != 0) return;
    const double disc = a.opts[R.opt_begin + opt].disc;
    const double scale = (R.precision == kPrecFP64) ? 1.0 : R.S0;
    const double n = (double)R.n_paths;
    double mean, var;
    mean_var(s, ss, n, mean, var);
    double price = scale * mean + (cv ? a.geo[R.out_base + o] : 0.0);
    if (reg) {
        double ma, va, mg, vg;
        mean_var(sa, ssa, n, ma, va);
        mean_var(sg, ssg, n, mg, vg);
        const double cov = 0.5 * (va + vg - var);
        const double beta = vg > 0.0 ? cov / vg : 0.0;
        price = scale * ma - beta * (scale * mg - a.geo[R.out_base + o]);
        var = va - beta * cov;
        var = var > 0.0 ? var : 0.0;
    }
    a.prices[R.out_base + o] = disc * price;
    if (a.std_errs) a.std_errs[R.out_base + o] = scale * disc * ::sqrt(var / n);
}

// ------------------------------------------------------------------------------------------
// Component kernels for the parity tests and measurement.
// ------------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_rng_fill(uint32_t* out, uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths,
                           uint64_t word0, uint64_t n_words) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_paths * n_words) return;
    const uint64_t i = t / n_words, j = t - i * n_words;
    const uint64_t w = word0 + j, path = path0 + i;
    const u32x4 b = philox4x32_10((uint32_t)(w >> 2), (uint32_t)path, (uint32_t)(path >> 32), stream, (uint32_t)seed,
                                  (uint32_t)(seed >> 32));
    const uint32_t e = (uint32_t)(w & 3u);
    out[t] = e == 0 ? b.x : e == 1 ? b.y : e == 2 ? b.z : b.w;
}

__global__ void k_philox(const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32x4 b = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], key[2 * i], key[2 * i + 1]);
    out[4 * i] = b.x;
    out[4 * i + 1] = b.y;
    out[4 * i + 2] = b.z;
    out[4 * i + 3] = b.w;
}

__global__ void k_normals(const uint32_t* words, uint64_t n, uint32_t precision, double* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = precision == kPrecFP64 ? normal_f64(words[i]) : (double)normal_f32(words[i]);
}

__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

inline dim3 grid_for(uint64_t n, uint32_t block) { return dim3((uint32_t)((n + block - 1) / block)); }

}  // namespace

// ------------------------------------------------------------------------------------------
// Launch wrappers (host)
// ------------------------------------------------------------------------------------------
const void* path_kernel_symbol(uint32_t precision, uint32_t ppt, PathMode mode, bool mart, bool geo) {
    if (precision == kPrecFP64) {
        switch (mode) {
            case PathMode::Normal: return reinterpret_cast<const void*>(&k_path_fp64<false, false>);
            case PathMode::Debug: return reinterpret_cast<const void*>(&k_path_fp64<true, false>);
            case PathMode::Count: return reinterpret_cast<const void*>(&k_path_fp64<false, true>);
        }
        return nullptr;
    }
#define HEXO_PPT_MODES(P, M, G)                                                                          \
    switch (mode) {                                                                                      \
        case PathMode::Normal: return reinterpret_cast<const void*>(&k_path_fp32<P, false, false, M, G>); \
        case PathMode::Debug: return reinterpret_cast<const void*>(&k_path_fp32<P, true, false, M, G>);   \
        case PathMode::Count: return reinterpret_cast<const void*>(&k_path_fp32<P, false, true, M, G>);   \
    }                                                                                                    \
    return nullptr;
#define HEXO_PPT_CASE(P)                    \
    case P:                                 \
        if (mart && geo) {                  \
            HEXO_PPT_MODES(P, true, true)   \
        }                                   \
        if (mart) {                         \
            HEXO_PPT_MODES(P, true, false)  \
        }                                   \
        if (geo) {                          \
            HEXO_PPT_MODES(P, false, true)  \
        }                                   \
        HEXO_PPT_MODES(P, false, false)
    switch (ppt) {
        HEXO_PPT_CASE(1)
        HEXO_PPT_CASE(2)
        HEXO_PPT_CASE(3)
        HEXO_PPT_CASE(4)
        default: return nullptr;
    }
#undef HEXO_PPT_CASE
#undef HEXO_PPT_MODES
}

size_t path_lds_bytes(uint32_t precision, uint32_t ppt, uint32_t block) {
    const size_t waves = block / kTile;
    // FP32: per wave the transpose of the per-path values at expiries (4 words per path, WaveLds);
    // FP64: 64 double4 per wave
    return precision == kPrecFP64 ? waves * kTile * sizeof(double4) : waves * 4 * ppt * kTile * sizeof(uint32_t);
}

hipError_t launch_paths(uint32_t precision, uint32_t ppt, PathMode mode, uint32_t grid, uint32_t block,
                        hipStream_t st, const PathArgs& a, bool mart, bool geo) {
    const size_t lds = path_lds_bytes(precision, ppt, block);
    if (precision == kPrecFP64) {
        switch (mode) {
            case PathMode::Normal: k_path_fp64<false, false><<<grid, block, lds, st>>>(a); break;
            case PathMode::Debug: k_path_fp64<true, false><<<grid, block, lds, st>>>(a); break;
            case PathMode::Count: k_path_fp64<false, true><<<grid, block, lds, st>>>(a); break;
        }
        return hipGetLastError();
    }
#define HEXO_LAUNCH_M(P, M, G)                                                                      \
    switch (mode) {                                                                                 \
        case PathMode::Normal: k_path_fp32<P, false, false, M, G><<<grid, block, lds, st>>>(a); break; \
        case PathMode::Debug: k_path_fp32<P, true, false, M, G><<<grid, block, lds, st>>>(a); break;   \
        case PathMode::Count: k_path_fp32<P, false, true, M, G><<<grid, block, lds, st>>>(a); break;   \
    }
#define HEXO_LAUNCH(P)                  \
    if (mart && geo) {                  \
        HEXO_LAUNCH_M(P, true, true)    \
    } else if (mart) {                  \
        HEXO_LAUNCH_M(P, true, false)   \
    } else if (geo) {                   \
        HEXO_LAUNCH_M(P, false, true)   \
    } else {                            \
        HEXO_LAUNCH_M(P, false, false)  \
    }
    switch (ppt) {
        case 1: HEXO_LAUNCH(1) break;
        case 2: HEXO_LAUNCH(2) break;
        case 3: HEXO_LAUNCH(3) break;
        case 4: HEXO_LAUNCH(4) break;
        default: return hipErrorInvalidValue;
    }
#undef HEXO_LAUNCH
#undef HEXO_LAUNCH_M
    return hipGetLastError();
}

hipError_t launch_reduce(const ReduceArgs& a, uint64_t n_group_threads, uint64_t n_out_threads, hipStream_t st) {
    if (n_group_threads) k_reduce_groups<<<grid_for(n_group_threads, 256), 256, 0, st>>>(a, n_group_threads);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // one wavefront per output
    if (n_out_threads) k_reduce_final<<<grid_for(n_out_threads * kTile, 256), 256, 0, st>>>(a, n_out_threads);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st) {
    k_fill_u32<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st) {
    k_fill_u64<<<grid_for(n, 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_rng_fill(uint32_t* out, uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths,
                           uint64_t word0, uint64_t n_words, hipStream_t st) {
    k_rng_fill<<<grid_for(n_paths * n_words, 256), 256, 0, st>>>(out, seed, stream, path0, n_paths, word0, n_words);
    return hipGetLastError();
}

hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out, hipStream_t st) {
    k_philox<<<grid_for(n, 256), 256, 0, st>>>(ctr, key, n, out);
    return hipGetLastError();
}

hipError_t launch_normals(const uint32_t* words, uint64_t n, uint32_t precision, double* out, hipStream_t st) {
    k_normals<<<grid_for(n, 256), 256, 0, st>>>(words, n, precision, out);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, uint32_t grid, hipStream_t st) {
    k_copy<<<grid, 256, 0, st>>>(reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), bytes / 16);
    return hipGetLastError();
}

}  // namespace hexo
