"""Derives the coefficients of the FP32 kernel's inverse normal (device.hpp, NormalF32).

With s = min(u, 1 - u), t = s (1 - s) and l = log2 t, |Phi^-1(u)| = (1 - 2s) h(l) with h smooth in l
(Giles' erfinv variable, M. Giles, "Approximating the erfinv function", GPU Computing Gems Jade
Edition, 2011). h is fitted as
  central: a degree-9 polynomial on l in [-11, -2], fitted in x = l + 6.5 and expanded in powers
           of l in FP64 (the kernel evaluates it in l, without the shift),
  tail:    a degree-8 polynomial in y = sqrt(-l) - yc on l in [log2(2^-33 (1 - 2^-33)), -11],
by least squares on Chebyshev nodes, reweighted (Lawson) towards the minimax of the absolute error
of z = (1 - 2s) h. The target is scipy's ndtri. Prints the coefficients as FP32 values and the
error of their FP32 evaluation (float32 rounding after every step, as the kernel's fused
multiply-adds) on a sample of 32-bit words.

  python normal_f32_fit.py
"""
from math import comb

import numpy as np
from scipy.special import ndtri

L_C = -11.0       # central / tail switch in l
X_C = 6.5         # central variable shift
DEG_C, DEG_T = 9, 8


def h_of_l(l):
    t = np.exp2(l)
    r = np.sqrt(1.0 - 4.0 * t)  # 1 - 2s
    s = 2.0 * t / (1.0 + r)     # (1 - r) / 2 without cancellation
    return -ndtri(s) / r, r


def nodes(lo, hi, n=6000):
    k = np.arange(n)
    return 0.5 * (hi - lo) * np.cos(np.pi * (k + 0.5) / n) + 0.5 * (hi + lo)


def lawson(x, h, r, deg, iters=80):
    V = np.vander(x, deg + 1, increasing=True)
    w = np.ones_like(x)
    for _ in range(iters):
        W = np.sqrt(w) * r
        coef, *_ = np.linalg.lstsq(V * W[:, None], h * W, rcond=None)
        e = np.abs((V @ coef - h) * r)
        w = w * (e / e.max() + 1e-4)
        w /= w.mean()
    return coef


def fit():
    l = nodes(L_C, -2.0)
    h, r = h_of_l(l)
    cx = lawson(l + X_C, h, r, DEG_C)
    central = np.zeros(DEG_C + 1)
    for k, ck in enumerate(cx):  # p(l + X_C) in powers of l
        for i in range(k + 1):
            central[i] += ck * comb(k, i) * X_C ** (k - i)
    central = central.astype(np.float32)
    l_min = np.log2(2.0 ** -33 * (1.0 - 2.0 ** -33))
    v = nodes(np.sqrt(-L_C), np.sqrt(-l_min))
    yc = np.float32(0.5 * (np.sqrt(-L_C) + np.sqrt(-l_min)))
    h, r = h_of_l(-v * v)
    tail = lawson(v - float(yc), h, r, DEG_T).astype(np.float32)
    return central, tail, float(yc)


def f32_fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def horner(c, x):
    acc = np.full_like(x, c[-1])
    for ci in c[-2::-1]:
        acc = f32_fma(acc, x, ci)
    return acc


def normal_f32(k, central, tail, yc):
    """The kernel's evaluation in float32 arithmetic (log2 and sqrt correctly rounded here)."""
    k = np.asarray(k, np.uint64)
    j = np.where(k >= 2 ** 31, 2 ** 32 - 1 - k, k)
    s = f32_fma(j.astype(np.float32), np.float32(2.0 ** -32), np.float32(2.0 ** -33))
    t = f32_fma(-s, s, s)
    l = np.log2(t.astype(np.float64)).astype(np.float32)
    hc = horner(central, l)
    y = (np.sqrt(-l.astype(np.float64)).astype(np.float32).astype(np.float64) - yc).astype(np.float32)
    ht = horner(tail, y)
    h = np.where(l < np.float32(L_C), ht, hc)
    pr = (h.astype(np.float64) * f32_fma(np.full_like(s, -2.0), s, 1.0).astype(np.float64)).astype(np.float32)
    return np.where(k >= 2 ** 31, pr, -pr)


def main():
    central, tail, yc = fit()
    rng = np.random.default_rng(5)
    k = np.concatenate([rng.integers(0, 2 ** 32, 2 ** 22), np.arange(0, 100000), np.arange(2 ** 32 - 100000, 2 ** 32),
                        np.arange(2 ** 21 - 50000, 2 ** 21 + 50000)]).astype(np.uint64)
    z = normal_f32(k, central, tail, yc).astype(np.float64)
    zt = ndtri((k.astype(np.float64) + 0.5) * 2.0 ** -32)
    e = np.abs(z - zt)
    big = np.abs(zt) >= 1.0
    print(f"max abs error {e.max():.3e}; max relative error for |z| >= 1: {(e[big] / np.abs(zt[big])).max():.3e}")
    print("central (powers of l), c0..c9:")
    for c in central:
        print(f"    {float(c)!r}f,")
    print(f"tail (y = sqrt(-l) - {yc!r}), d0..d8:")
    for c in tail:
        print(f"    {float(c)!r}f,")


if __name__ == "__main__":
    main()
