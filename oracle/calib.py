"""Calibration oracle (TEST INFRASTRUCTURE ONLY: the checker for hexo_calibrate, never shipped).

Restates the reference's calibration, src/HCalibration.cpp:62-117: the initial point from the
volume-weighted average Black-Scholes implied volatility (src/BSM.cpp:22-66), then
Levenberg-Marquardt with analytic Jacobian over SWIFT prices (get_prices_for_levmar /
get_jacobian_for_levmar, src/HCalibration.cpp:37-59), SWIFT grids re-chosen at every parameter
point (update_adata, :10-36).

The reference's Levenberg-Marquardt is levmar 2.6 (M. Lourakis, fetched and patched by
Makefile.am:102-108, not vendored): dlevmar_der, built without LAPACK. Its published algorithm
(Madsen, Nielsen and Tingleff, "Methods for non-linear least squares problems", 2004, Alg. 3.16,
with levmar's stopping rules) is restated here: damping mu = tau max diag(J^T J) at the first
iterate; a step solves (J^T J + mu I) dp = J^T e; it is taken when the gain ratio is positive,
with mu *= max(1/3, 1 - (2 rho - 1)^3) and nu = 2, else mu *= nu, nu *= 2. Stop codes:
1 small gradient, 2 small step, 3 iteration limit, 4 singular, 5 no further reduction,
6 small error, 7 invalid values. Trial points outside the Heston domain, or whose SWIFT grid
exceeds the pricer's limit, count as rejected steps.
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as O

YEARLY_RISK_FREE = 0.0005  # src/inc/BSM.h:10
EPS = np.finfo(np.float64).eps


def bs_call(S, K, sigma, T, r=YEARLY_RISK_FREE):
    """src/BSM.cpp:12-21."""
    d1 = (math.log(S / K) + (r + 0.5 * sigma * sigma) * T) / (sigma * math.sqrt(T))
    d2 = (math.log(S / K) + (r - 0.5 * sigma * sigma) * T) / (sigma * math.sqrt(T))
    n = lambda v: 0.5 * math.erfc(-v / math.sqrt(2.0))
    return S * n(d1) - K * math.exp(-r * T) * n(d2)


def imp_vol(S, K, price, T):
    """Bisection of src/BSM.cpp:22-50 (bracket by doubling/halving from 1/64, to 1e-6)."""
    if price <= S - math.exp(-YEARLY_RISK_FREE * T) * K:
        return -1.0
    if price >= S:
        return -2.0
    i = 1.0 / 64
    if bs_call(S, K, i, T) < price:
        lo = i
        i *= 2
        while bs_call(S, K, i, T) < price:
            i *= 2
        hi = i
    else:
        hi = i
        i *= 0.5
        while bs_call(S, K, i, T) >= price:
            i *= 0.5
        lo = i
    while True:
        mid = (lo + hi) / 2
        if bs_call(S, K, mid, T) < price:
            lo = mid
        else:
            hi = mid
        if abs(hi - lo) < 1e-6:
            return mid


def initial_point(S, chains):
    """src/HCalibration.cpp:67-78 with avg_imp_vol (src/BSM.cpp:51-66)."""
    num = den = 0.0
    for T, K, prices, vols in chains:
        for k, p, v in zip(K, prices, vols):
            iv = imp_vol(S, k, p, T)
            if iv >= 0:
                num += v * iv
                den += v
    iv = num / den if den else 0.3
    return np.array([iv * iv, iv * iv, -0.2, 1.0, 1.0])


def valid(p):
    return p[0] > 0 and p[1] > 0 and p[3] > 0 and p[4] > 0 and -1 < p[2] < 1


SWIFT_MAX_J = 16384  # largest SWIFT grid the pricer accepts (include/hexo.h HEXO_SWIFT_MAX_J)


class Unpriceable(ValueError):
    """The SWIFT grid for a parameter set exceeds SWIFT_MAX_J."""


def model(p, S, chains, r, trunc, grads):
    prices, jac = [], []
    for T, K, _, _ in chains:
        try:
            out = O.swift_price(list(p), S, r, T, K, trunc=trunc, grads=grads)
        except ValueError as exc:
            raise Unpriceable(str(exc)) from exc
        if out[1][7] > SWIFT_MAX_J:
            raise Unpriceable(f"SWIFT grid J={int(out[1][7])} exceeds {SWIFT_MAX_J}")
        prices.append(out[0])
        if grads:
            jac.append(out[2])
    return np.concatenate(prices), (np.vstack(jac) if grads else None)


def levenberg_marquardt(S, chains, p0, r=YEARLY_RISK_FREE, trunc=1e-10, itmax=100, tau=1e-3, eps1=1e-10,
                        eps2=1e-10, eps3=1e-10):
    """Returns dict(params, err0, err, iterations, reason, nfev, njev)."""
    x = np.concatenate([np.asarray(c[2], np.float64) for c in chains])
    p = np.array(p0, np.float64)
    hx, _ = model(p, S, chains, r, trunc, False)
    e = x - hx
    eL2 = float(e @ e)
    err0 = eL2
    nfev, njev = 1, 0
    mu, nu = 0.0, 2
    stop = 0
    k = 0
    while k < itmax:
        if eL2 <= eps3:
            stop = 6
            break
        _, J = model(p, S, chains, r, trunc, True)
        njev += 1
        JtJ = J.T @ J
        Jte = J.T @ e
        if np.max(np.abs(Jte)) <= eps1:
            stop = 1
            break
        pL2 = float(p @ p)
        if k == 0:
            mu = tau * float(np.max(np.diag(JtJ)))
        while True:
            A = JtJ + mu * np.eye(5)
            try:
                dp = np.linalg.solve(A, Jte)
                solved = bool(np.all(np.isfinite(dp)))
            except np.linalg.LinAlgError:
                solved = False
            if solved:
                dpL2 = float(dp @ dp)
                pn = p + dp
                if dpL2 <= eps2 * eps2 * pL2:
                    stop = 2
                    break
                if dpL2 >= (pL2 + eps2) / (EPS * EPS):
                    stop = 4
                    break
                hn = None
                if valid(pn):
                    try:
                        hn, _ = model(pn, S, chains, r, trunc, False)
                    except Unpriceable:
                        hn = None
                if hn is not None:
                    nfev += 1
                    en = x - hn
                    enL2 = float(en @ en)
                    if not math.isfinite(enL2):
                        stop = 7
                        break
                    dL = float(dp @ (mu * dp + Jte))
                    dF = eL2 - enL2
                    if dL > 0 and dF > 0:
                        t = 2.0 * dF / dL - 1.0
                        mu *= max(1.0 / 3.0, 1.0 - t * t * t)
                        nu = 2
                        p, e, eL2 = pn, en, enL2
                        break
            mu *= nu
            nu2 = nu * 2
            if nu2 > 2 ** 62:
                stop = 5
                break
            nu = nu2
        k += 1
        if stop:
            break
    if not stop:
        stop = 3
    return {"params": p, "err0": err0, "err": eL2, "iterations": k, "reason": stop, "nfev": nfev, "njev": njev}
