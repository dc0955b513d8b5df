// cplx_device.hpp -- FP64 complex arithmetic for the device code (SWIFT pricer, geometric
// control variate): plain struct Cplx with the operators and elementary functions they need.
#pragma once

#include <hip/hip_runtime.h>

#include "cplx.hpp"

namespace hexo {

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

// log(1 + z) and exp(w) - 1 without cancellation for small arguments (short series below 1e-3,
// whose truncation error is below 1e-15 relative)
__device__ __forceinline__ Cplx clog1p(Cplx z) {
    if (hypot(z.re, z.im) < 1e-3) {
        // z (1 - z (1/2 - z (1/3 - z/4)))
        Cplx t = cx(1.0 / 3.0) - 0.25 * z;
        t = cx(0.5) - z * t;
        t = cx(1.0) - z * t;
        return z * t;
    }
    return clog(cx(1.0 + z.re, z.im));
}
__device__ __forceinline__ Cplx cexpm1(Cplx w) {
    if (hypot(w.re, w.im) < 1e-3) {
        // w (1 + w (1/2 + w (1/6 + w/24)))
        Cplx t = cx(1.0 / 6.0) + (1.0 / 24.0) * w;
        t = cx(0.5) + w * t;
        t = cx(1.0) + w * t;
        return w * t;
    }
    const Cplx e = cexp(w);
    return Cplx{e.re - 1.0, e.im};
}

}  // namespace hexo
