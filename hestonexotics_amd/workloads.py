"""Seeded synthetic workloads for the pricing path (configurations C1-C5).

C1  one arithmetic-Asian call: S0=K=100, r=0.02, T=1, kappa=2, theta=0.04, sigma=0.3, rho=-0.7,
    v0=0.04; 252 steps, 100,000 paths, seed 42; FP64.
C2  strike/maturity grid: same Heston parameters; 64 strikes linspace(70, 130) x 8 maturities
    {0.1, 0.25, 0.5, 1, 2, 3, 4, 5}; 1,000,000 paths per maturity shared by all strikes and by the
    Asian and European payoffs; round(52 T) steps; FP32 stepping, FP64 reduction.
C3  512 requests: kappa~U[0.5,20], theta~U[0.01,0.2], sigma~U[0.1,1], rho~U[-0.95,0],
    v0~U[0.01,0.2], S0=100, K~U[80,120], T~U[0.1,5], steps=round(252 T), paths log-uniform in
    [1e4, 1e7]; seed per request.
C4  Poisson arrivals at 200 req/s, each request drawn as in C3 with payoff Asian or European w.p. 1/2.
C5  stiff Asian call: kappa=0.5, theta=0.04, sigma=0.9, rho=-0.9, v0=0.04, S0=K=100, r=0.02, T=1;
    2^27 paths x 2048 steps.

The rate r = 0.02 is applied through the HEXO_DRIFT and HEXO_DISCOUNT extensions (the reference's
scheme has neither); flags=0 reproduces the reference's undiscounted, driftless estimator.
"""
from __future__ import annotations

import numpy as np

from .api import ASIAN, DISCOUNT, DRIFT, EUROPEAN, FP32, FP64, Request

BASE_PARAMS = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3)
C2_MATURITIES = [0.1, 0.25, 0.5, 1.0, 2.0, 3.0, 4.0, 5.0]


def request_seed(base: int, idx: int) -> int:
    """64-bit per-request seed: splitmix64 of (base, idx)."""
    z = (base * 0x9E3779B97F4A7C15 + idx + 1) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def c1(n_paths: int = 100_000, precision: int = FP64, payoffs: int = ASIAN, seed: int = 42) -> Request:
    return Request(**BASE_PARAMS, S0=100.0, expiries=[1.0], strikes=[[100.0]], steps=252, n_paths=n_paths, seed=seed,
                   r=0.02, flags=DRIFT | DISCOUNT, payoffs=payoffs, precision=precision)


def c2(n_paths: int = 1_000_000, precision: int = FP32, seed: int = 2024, n_strikes: int = 64,
       maturities=None) -> list[Request]:
    strikes = list(np.linspace(70.0, 130.0, n_strikes))
    mats = C2_MATURITIES if maturities is None else maturities
    return [Request(**BASE_PARAMS, S0=100.0, expiries=[T], strikes=[strikes], steps=int(round(52 * T)),
                    n_paths=n_paths, seed=request_seed(seed, i), r=0.02, flags=DRIFT | DISCOUNT,
                    payoffs=ASIAN | EUROPEAN, precision=precision) for i, T in enumerate(mats)]


def c3_params(n: int = 512, seed: int = 7, path_scale: float = 1.0, min_paths: int = 64):
    rng = np.random.default_rng(seed)
    kappa = rng.uniform(0.5, 20.0, n)
    theta = rng.uniform(0.01, 0.2, n)
    sigma = rng.uniform(0.1, 1.0, n)
    rho = rng.uniform(-0.95, 0.0, n)
    v0 = rng.uniform(0.01, 0.2, n)
    K = rng.uniform(80.0, 120.0, n)
    T = rng.uniform(0.1, 5.0, n)
    paths = np.round(10.0 ** rng.uniform(4.0, 7.0, n))
    paths = np.maximum(np.round(paths * path_scale), min_paths).astype(np.int64)
    pay = rng.integers(0, 2, n)
    return dict(kappa=kappa, theta=theta, sigma=sigma, rho=rho, v0=v0, K=K, T=T, paths=paths, pay=pay)


def c3(n: int = 512, seed: int = 7, precision: int = FP32, path_scale: float = 1.0, payoffs=None) -> list[Request]:
    p = c3_params(n, seed, path_scale)
    out = []
    for i in range(n):
        T = float(p["T"][i])
        pay = (ASIAN | EUROPEAN) if payoffs is None else payoffs
        out.append(Request(v0=float(p["v0"][i]), theta=float(p["theta"][i]), rho=float(p["rho"][i]),
                           kappa=float(p["kappa"][i]), sigma=float(p["sigma"][i]), S0=100.0, expiries=[T],
                           strikes=[[float(p["K"][i])]], steps=max(1, int(round(252 * T))),
                           n_paths=int(p["paths"][i]), seed=request_seed(seed, i), r=0.02,
                           flags=DRIFT | DISCOUNT, payoffs=pay, precision=precision))
    return out


def c4(rate: float = 200.0, duration_s: float = 60.0, seed: int = 11, precision: int = FP32,
       path_scale: float = 1.0):
    """Poisson arrival trace: returns (requests, arrival times in seconds)."""
    rng = np.random.default_rng(seed)
    t, times = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            break
        times.append(t)
    n = len(times)
    p = c3_params(n, seed + 1, path_scale)
    reqs = []
    for i in range(n):
        T = float(p["T"][i])
        reqs.append(Request(v0=float(p["v0"][i]), theta=float(p["theta"][i]), rho=float(p["rho"][i]),
                            kappa=float(p["kappa"][i]), sigma=float(p["sigma"][i]), S0=100.0, expiries=[T],
                            strikes=[[float(p["K"][i])]], steps=max(1, int(round(252 * T))),
                            n_paths=int(p["paths"][i]), seed=request_seed(seed, i), r=0.02,
                            flags=DRIFT | DISCOUNT, payoffs=ASIAN if p["pay"][i] == 0 else EUROPEAN,
                            precision=precision))
    return reqs, np.asarray(times)


def c5(n_paths: int = 2 ** 27, steps: int = 2048, precision: int = FP32, seed: int = 5) -> Request:
    return Request(v0=0.04, theta=0.04, rho=-0.9, kappa=0.5, sigma=0.9, S0=100.0, expiries=[1.0], strikes=[[100.0]],
                   steps=steps, n_paths=n_paths, seed=seed, r=0.02, flags=DRIFT | DISCOUNT, payoffs=ASIAN,
                   precision=precision)


def path_steps(reqs) -> float:
    """Nominal path-steps: sum of n_paths * steps * n_expiries."""
    return float(sum(r.n_paths * r.steps * len(r.expiries) for r in reqs))
