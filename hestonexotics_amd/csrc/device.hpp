// device.hpp -- device building blocks of the Heston QE path kernels (gfx950).
//
//  * Philox4x32-10 counter-based generator (Salmon et al., SC'11). Replaces the reference's
//    shishua ring buffers (src/inc/RNG.h:17-52, src/RNG.cpp:8-47): any (seed, path, step) draw is
//    computed directly, so paths can be split across waves, launches or GPUs freely.
//  * Uniforms u = (k + 1/2) 2^-32 from a 32-bit word k (the reference maps a 64-bit word,
//    src/RNG.cpp:31). u is never 0 or 1, and 1 - u(k) = u(~k) exactly.
//  * AS241 inverse normal (Wichura 1988; src/as241.f90:15-119): an FP64 version in the
//    reference's form and an FP32 version that forms q = u - 1/2 and min(u, 1-u) from the
//    integer, evaluates the central and tail rationals without branching and selects.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hexo {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

struct u32x4 {
    uint32_t x, y, z, w;
};

// One counter block: ctr = {c0, c1, c2, c3}, key = {k0, k1}. When c0, c2, c3 and the key are
// wave-uniform (step index, path high word, stream, seed) the first round runs on the scalar unit.
__device__ __forceinline__ u32x4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                               uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)kPhiloxM0 * c0;
        const uint64_t p1 = (uint64_t)kPhiloxM1 * c2;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96)
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

// ---- AS241 coefficients (src/as241.f90:31-44 A/B, :49-63 C/D, :68-82 E/F) ----
#define HEXO_AS241_A 3.3871328727963666080e0, 1.3314166789178437745e+2, 1.9715909503065514427e+3, \
    1.3731693765509461125e+4, 4.5921953931549871457e+4, 6.7265770927008700853e+4, 3.3430575583588128105e+4, \
    2.5090809287301226727e+3
#define HEXO_AS241_B 1.0, 4.2313330701600911252e+1, 6.8718700749205790830e+2, 5.3941960214247511077e+3, \
    2.1213794301586595867e+4, 3.9307895800092710610e+4, 2.8729085735721942674e+4, 5.2264952788528545610e+3
#define HEXO_AS241_C 1.42343711074968357734e0, 4.63033784615654529590e0, 5.76949722146069140550e0, \
    3.64784832476320460504e0, 1.27045825245236838258e0, 2.41780725177450611770e-1, 2.27238449892691845833e-2, \
    7.74545014278341407640e-4
#define HEXO_AS241_D 1.0, 2.05319162663775882187e0, 1.67638483018380384940e0, 6.89767334985100004550e-1, \
    1.48103976427480074590e-1, 1.51986665636164571966e-2, 5.47593808499534494600e-4, 1.05075007164441684324e-9
#define HEXO_AS241_E 6.65790464350110377720e0, 5.46378491116411436990e0, 1.78482653991729133580e0, \
    2.96560571828504891230e-1, 2.65321895265761230930e-2, 1.24266094738807843860e-3, 2.71155556874348757815e-5, \
    2.01033439929228813265e-7
#define HEXO_AS241_F 1.0, 5.99832206555887937690e-1, 1.36929880922735805310e-1, 1.48753612908506148525e-2, \
    7.86869131145613259100e-4, 1.84631831751005468180e-5, 1.42151175831644588870e-7, 2.04426310338993978564e-15

template <typename T>
struct As241 {
    static constexpr T a[8] = {HEXO_AS241_A};
    static constexpr T b[8] = {HEXO_AS241_B};
    static constexpr T c[8] = {HEXO_AS241_C};
    static constexpr T d[8] = {HEXO_AS241_D};
    static constexpr T e[8] = {HEXO_AS241_E};
    static constexpr T f[8] = {HEXO_AS241_F};
};

// ((c7 r + c6) r + ...) r + c0 with fused multiply-adds; a leading coefficient of 1 (B, D, F
// have c0 = 1) is handled by the same chain.
template <typename T>
__device__ __forceinline__ T horner8(const T (&c)[8], T r) {
    T acc = c[7];
#pragma unroll
    for (int i = 6; i >= 0; --i) acc = __builtin_fma(acc, r, c[i]);
    return acc;
}

// 1/x from the hardware estimate and one Newton step (error ~ulp^2 before the final rounding): the
// estimate alone is not correctly rounded, and a systematic error in it biases scheme means.
__device__ __forceinline__ float rcp_refined(float x) {
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

__device__ __forceinline__ double uniform_f64(uint32_t k) { return ((double)k + 0.5) * 0x1p-32; }
// (k + 1/2) 2^-32 and 1 - that, in FP32, from the integer.
__device__ __forceinline__ float uniform_f32(uint32_t k) { return __builtin_fmaf((float)k, 0x1p-32f, 0x1p-33f); }
__device__ __forceinline__ float one_minus_uniform_f32(uint32_t k) { return uniform_f32(~k); }

// AS241 in FP64, the reference's region logic (src/as241.f90:85-118) on u = (k + 1/2) 2^-32.
// u never reaches 0 or 1, so the IFAULT branch cannot trigger; 32-bit inputs never reach the
// far tail (r <= 4.78 < 5), which is kept for completeness.
__device__ __forceinline__ double normal_f64(uint32_t k) {
    const double p = uniform_f64(k);
    const double q = p - 0.5;
    if (__builtin_fabs(q) <= 0.425) {
        const double r = 0.180625 - q * q;
        return q * horner8(As241<double>::a, r) / horner8(As241<double>::b, r);
    }
    double r = (q < 0.0) ? p : 1.0 - p;
    r = ::sqrt(-::log(r));
    double z;
    if (r <= 5.0) {
        r -= 1.6;
        z = horner8(As241<double>::c, r) / horner8(As241<double>::d, r);
    } else {
        r -= 5.0;
        z = horner8(As241<double>::e, r) / horner8(As241<double>::f, r);
    }
    return (q < 0.0) ? -z : z;
}

// ---- AS241 PPND7 coefficients: the single-precision routine of the same algorithm (Wichura,
// Appl. Statist. 37(3), 1988), accurate to about 1 part in 10^7 -- the resolution of FP32.
// Central region (|q| <= 0.425): q A(r)/B(r), r = 0.180625 - q^2; tail (r = sqrt(-ln min(u,1-u))
// <= 5): C(r - 1.6)/D(r - 1.6). The far-tail pair (r > 5) is unreachable from 32-bit words:
// min(u, 1-u) >= 2^-33 gives r <= 4.78.
struct Ppnd7 {
    static constexpr float a[4] = {3.3871327179e0f, 5.0434271938e1f, 1.5929113202e2f, 5.9109374720e1f};
    static constexpr float b[4] = {1.0f, 1.7895169469e1f, 7.8757757664e1f, 6.7187563600e1f};
    static constexpr float c[4] = {1.4234372777e0f, 2.7568153900e0f, 1.3067284816e0f, 1.7023821103e-1f};
    static constexpr float d[3] = {1.0f, 7.3700164250e-1f, 1.2021132975e-1f};
};

template <int N>
__device__ __forceinline__ float hornerf(const float (&c)[N], float r) {
    float acc = c[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; --i) acc = __builtin_fmaf(acc, r, c[i]);
    return acc;
}

// a * K + b (v_fmamk_f32) and a * b + K (v_fmaak_f32) with K a 32-bit literal and the other operands
// in vector registers: the VOP2 forms that encode a literal. A VOP3 v_fma_f32 cannot, so for a
// constant operand the compiler reads a scalar register instead, and on gfx950 an FP32 instruction
// with a scalar-register operand issues at about half the rate of one without (measured). The
// result is the same fused multiply-add, bit for bit.
template <uint32_t K>
__device__ __forceinline__ float fmamk_lit(float a, float b) {
    float d;
    asm("v_fmamk_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "i"(K), "v"(b));
    return d;
}
template <uint32_t K>
__device__ __forceinline__ float fmaak_lit(float a, float b) {
    float d;
    asm("v_fmaak_f32 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "i"(K));
    return d;
}
#define HEXO_F32_BITS(x) __builtin_bit_cast(uint32_t, (float)(x))

// A value the compiler must keep in a vector register (it cannot prove the result wave-uniform and
// fold it back into a scalar operand).
__device__ __forceinline__ float in_vgpr(float x) {
    float d;
    asm("v_mov_b32 %0, %1" : "=v"(d) : "v"(x));
    return d;
}

// The PPND7 central rational's constants that its literal-form Horner steps cannot encode: the
// region's 0.180625 (in r = 0.180625 - q^2, a VOP3 fused multiply-add) and the second coefficients
// of A and B (the first step, r a3 + a2, takes a3 as the literal). Made once per thread.
struct Ppnd7Vregs {
    float c, a2, b2;  // central region
    float c2, d1;     // tail region: the second coefficients of C and D
};
__device__ __forceinline__ Ppnd7Vregs ppnd7_vregs() {
    return Ppnd7Vregs{in_vgpr(0.180625f), in_vgpr(Ppnd7::a[2]), in_vgpr(Ppnd7::b[2]), in_vgpr(Ppnd7::c[2]),
                      in_vgpr(Ppnd7::d[1])};
}

// N(0,1) in FP32 from a 32-bit word through AS241 PPND7 (the FP32 kernel's inverse normal).
// q = (k - 2^31 + 1/2) 2^-32 and the tail argument min(u, 1-u) = (min(k, ~k) + 1/2) 2^-32 are
// formed from the integer, so both tails keep full relative precision (no 1 - u cancellation).
__device__ __forceinline__ float central_q_f32(uint32_t k) {
    const int32_t sk = (int32_t)(k ^ 0x80000000u);
    return __builtin_fmaf((float)sk, 0x1p-32f, 0x1p-33f);
}
// Central region: q A(r)/B(r) with r = 0.180625 - q^2, meaningful for |q| <= 0.425. The Horner
// steps of A and B, ((a3 r + a2) r + a1) r + a0 and likewise B, are the same fused multiply-adds as
// hornerf's, in the literal forms above.
__device__ __forceinline__ float normal_central_rational_f32(float q, const Ppnd7Vregs& k) {
    const float rc = __builtin_fmaf(-q, q, k.c);
    float A = fmamk_lit<HEXO_F32_BITS(Ppnd7::a[3])>(rc, k.a2);
    A = fmaak_lit<HEXO_F32_BITS(Ppnd7::a[1])>(A, rc);
    A = fmaak_lit<HEXO_F32_BITS(Ppnd7::a[0])>(A, rc);
    float B = fmamk_lit<HEXO_F32_BITS(Ppnd7::b[3])>(rc, k.b2);
    B = fmaak_lit<HEXO_F32_BITS(Ppnd7::b[1])>(B, rc);
    B = fmaak_lit<HEXO_F32_BITS(Ppnd7::b[0])>(B, rc);
    return (q * A) * __builtin_amdgcn_rcpf(B);
}
// The central value of word k; `central` = |q| <= 0.425 (the value is meaningless otherwise).
__device__ __forceinline__ float normal_central_f32(uint32_t k, bool& central) {
    const float q = central_q_f32(k);
    central = __builtin_fabsf(q) <= 0.425f;
    return normal_central_rational_f32(q, ppnd7_vregs());
}
// Tail region (|q| > 0.425): r = sqrt(-ln(min(u, 1-u))) - 1.6, C(r)/D(r) with the sign of q; the
// Horner steps of C and D in the literal forms, as the central rational's.
__device__ __forceinline__ float normal_tail_f32(uint32_t k, const Ppnd7Vregs& pk) {
    const int32_t sk = (int32_t)(k ^ 0x80000000u);
    const uint32_t mn = k ^ ~(uint32_t)(sk >> 31);
    const float rr = __builtin_fmaf((float)mn, 0x1p-32f, 0x1p-33f);
    const float rt = __builtin_amdgcn_sqrtf(-0.693147180559945309f * __builtin_amdgcn_logf(rr)) - 1.6f;
    float C = fmamk_lit<HEXO_F32_BITS(Ppnd7::c[3])>(rt, pk.c2);
    C = fmaak_lit<HEXO_F32_BITS(Ppnd7::c[1])>(C, rt);
    C = fmaak_lit<HEXO_F32_BITS(Ppnd7::c[0])>(C, rt);
    float D = fmamk_lit<HEXO_F32_BITS(Ppnd7::d[2])>(rt, pk.d1);
    D = fmaak_lit<HEXO_F32_BITS(Ppnd7::d[0])>(D, rt);
    const float z = C * __builtin_amdgcn_rcpf(D);
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, z) ^ ((uint32_t)sk & 0x80000000u));
}
__device__ __forceinline__ float normal_f32(uint32_t k) {
    bool central;
    const float zc = normal_central_f32(k, central);
    return central ? zc : normal_tail_f32(k, ppnd7_vregs());
}

}  // namespace hexo
