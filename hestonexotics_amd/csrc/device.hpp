// device.hpp -- device building blocks of the Heston QE path kernels (gfx950).
//
//  * Philox4x32-10 counter-based generator (Salmon et al., SC'11). Replaces the reference's
//    shishua ring buffers (src/inc/RNG.h:17-52, src/RNG.cpp:8-47): any (seed, path, step) draw is
//    computed directly, so paths can be split across waves, launches or GPUs freely.
//  * Uniforms u = (k + 1/2) 2^-32 from a 32-bit word k (the reference maps a 64-bit word,
//    src/RNG.cpp:31). u is never 0 or 1, and 1 - u(k) = u(~k) exactly.
//  * Inverse normals: AS241 PPND16 (Wichura 1988; src/as241.f90:15-119) in FP64 in the reference's
//    form, and the FP32 kernel's own: polynomials in log2(s (1 - s)) with s = min(u, 1-u) formed
//    from the integer, the tail polynomial evaluated only where a lane needs it.
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

// A value the compiler must keep in a vector register (it cannot prove the result wave-uniform and
// fold it back into a scalar operand): on gfx950 an FP32 or bitwise instruction with a
// scalar-register operand issues at about half the rate of one without (measured). The compiler
// does not insert wait states in front of inline asm, and a transcendental (v_log, v_exp, v_sqrt,
// v_rcp, ...) needs one before the next instruction reads its result or overwrites its source
// register, whichever instruction the scheduler puts before the asm; so the asm opens with its own
// wait state (s_nop 0).
__device__ __forceinline__ float in_vgpr(float x) {
    float d;
    asm("s_nop 0\n\tv_mov_b32 %0, %1" : "=v"(d) : "v"(x));
    return d;
}
__device__ __forceinline__ uint32_t in_vgpr_u32(uint32_t x) {
    uint32_t d;
    asm("s_nop 0\n\tv_mov_b32 %0, %1" : "=v"(d) : "v"(x));
    return d;
}

// ---- The FP32 kernel's inverse normal ----
// z = Phi^-1(u) for u = (k + 1/2) 2^-32, the uniform of word k (as uniform_f64). The word is folded
// onto the lower half, j = min(k, ~k), so that s = (j + 1/2) 2^-32 = min(u, 1 - u) comes from the
// integer with only the conversion's rounding and both tails keep full relative precision. With
// t = s (1 - s) and l = log2 t (in [-33, -2] for 32-bit words), |z| = (1 - 2s) h(l), where h is smooth
// in l: Giles' variable for erfinv (M. Giles, "Approximating the erfinv function", GPU Computing
// Gems Jade Edition, 2011: w = -ln(1 - x^2) with x = 1 - 2s, i.e. w = -(l + 2) ln 2). h is a degree-9
// polynomial in l where l >= -11 (s >= 4.9e-4: all but about 1 draw in 1024) and a degree-8
// polynomial in sqrt(-l) - 4.5306 below. The coefficients are weighted least-squares fits of
// -ndtri(s) / (1 - 2s) by normal_f32_fit.py (this directory; the central fit is made in l + 6.5 and
// expanded in powers of l in FP64). Evaluated in FP32 (hardware log2 and square root, fused
// multiply-adds), the result is within 8e-7 of Phi^-1(u) for l >= -11 and within 3e-7 relative
// below (tests/test_oracle_normal_f32.py pins the restatement in the oracle).
struct NormalF32 {
    static constexpr float lc = -11.0f;  // the tail polynomial applies below
    static constexpr float c[10] = {0.8369221091270447f,     -0.18630073964595795f,   0.011675846762955189f,
                                    1.11724120870349e-05f,   -0.0002003998088184744f, -9.236554433300626e-06f,
                                    1.7887376770886476e-06f, 2.4262888587145426e-07f, 1.1781071229677309e-08f,
                                    2.1350940460074241e-10f};  // central: powers of l
    static constexpr float yc = 4.5305938720703125f;  // tail variable y = sqrt(-l) - yc
    static constexpr float d[9] = {4.8360700607299805f,      1.2490957975387573f,     -0.011836512014269829f,
                                   0.0019035495351999998f,   -0.00013401612523011863f, -0.00016207653970923275f,
                                   0.00019050169794354588f,  -0.00012945415801368654f, 3.693158214446157e-05f};
};

// The constants of the evaluation that the literal forms cannot carry (the second operand of the
// first Horner step, 2^-33, the sign mask of a three-input bitwise op), made once per thread in
// vector registers.
struct NormalVregs {
    float c8, h33;
    uint32_t sign;
};
__device__ __forceinline__ NormalVregs normal_vregs() {
    return NormalVregs{in_vgpr(NormalF32::c[8]), in_vgpr(0x1p-33f), in_vgpr_u32(0x80000000u)};
}

// The FP32 normal of word k. `tail` reports l < lc; the tail polynomial is evaluated only where
// `need` is also set (a caller that does not use the value passes false: it then holds the central
// polynomial's value, meaningless in the tail).
__device__ __forceinline__ float normal_word_f32(uint32_t k, bool need, const NormalVregs& nv, bool& tail) {
    const uint32_t j = k ^ (uint32_t)((int32_t)k >> 31);
    const float s = __builtin_fmaf((float)j, 0x1p-32f, nv.h33);
    const float t = __builtin_fmaf(-s, s, s);
    const float l = __builtin_amdgcn_logf(t);
    // Horner steps the compiler forms as v_fmamk_f32 / v_fmaak_f32 with the coefficient as the
    // instruction's literal (no scalar-register operand)
    float h = __builtin_fmaf(l, NormalF32::c[9], nv.c8);
    h = __builtin_fmaf(h, l, NormalF32::c[7]);
    h = __builtin_fmaf(h, l, NormalF32::c[6]);
    h = __builtin_fmaf(h, l, NormalF32::c[5]);
    h = __builtin_fmaf(h, l, NormalF32::c[4]);
    h = __builtin_fmaf(h, l, NormalF32::c[3]);
    h = __builtin_fmaf(h, l, NormalF32::c[2]);
    h = __builtin_fmaf(h, l, NormalF32::c[1]);
    h = __builtin_fmaf(h, l, NormalF32::c[0]);
    tail = l < NormalF32::lc;
    if (need && tail) {
        const float y = __builtin_amdgcn_sqrtf(-l) - NormalF32::yc;
        float g = __builtin_fmaf(y, NormalF32::d[8], NormalF32::d[7]);
        g = __builtin_fmaf(g, y, NormalF32::d[6]);
        g = __builtin_fmaf(g, y, NormalF32::d[5]);
        g = __builtin_fmaf(g, y, NormalF32::d[4]);
        g = __builtin_fmaf(g, y, NormalF32::d[3]);
        g = __builtin_fmaf(g, y, NormalF32::d[2]);
        g = __builtin_fmaf(g, y, NormalF32::d[1]);
        h = __builtin_fmaf(g, y, NormalF32::d[0]);
    }
    const float pr = h * __builtin_fmaf(-2.0f, s, 1.0f);  // |z| = h (1 - 2s)
    // negative for k < 2^31 (u < 1/2)
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, pr) ^ (~k & nv.sign));
}
__device__ __forceinline__ float normal_f32(uint32_t k) {
    bool tail;
    return normal_word_f32(k, true, normal_vregs(), tail);
}

}  // namespace hexo
