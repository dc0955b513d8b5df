"""GPU: Andersen's martingale correction (HEXO_MARTINGALE) through the C ABI.

(1) The FP64 kernel reproduces the oracle's corrected paths to 1e-12 relative and its prices to
    1e-12 (identical draws).
(2) The FP32 kernel with the corrected step (its own kernel class, a re-associated form without
    K0/K1 cancellation) agrees with FP64 on identical draws within the FP32 tolerance.
(3) The property itself at full scale: E[X_T] = S0 exp(r T) within the standard error with the
    correction, tens of standard errors off without it.
(4) Batches mixing FP64, FP32 and corrected-FP32 requests return what each returns alone.
"""
import dataclasses
import math

import numpy as np
import pytest

import hestonexotics_amd as H
from hestonexotics_amd import workloads as W

pytestmark = pytest.mark.gpu

FLAGS = H.DRIFT | H.MARTINGALE


def say(*a):
    print(*a, flush=True)


def requests():
    base = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, r=0.02)
    return {
        "c1_like": H.Request(**base, expiries=[1.0], strikes=[[90.0, 100.0, 110.0]], steps=252, n_paths=2000,
                             seed=42, flags=FLAGS, precision=H.FP64),
        "coarse_multi": H.Request(**base, expiries=[0.5, 2.0], strikes=[[100.0], [0.0, 100.0]], steps=3,
                                  n_paths=1000, seed=7, flags=FLAGS | H.DISCOUNT, precision=H.FP64),
        "stiff_c5": H.Request(v0=0.2, theta=0.04, rho=-0.9, kappa=0.5, sigma=0.9, S0=100.0, r=0.02,
                              expiries=[1.0], strikes=[[0.0, 100.0]], steps=64, n_paths=640, seed=5, flags=FLAGS,
                              precision=H.FP64),
        "positive_rho": H.Request(v0=0.2, theta=0.04, rho=0.5, kappa=2.0, sigma=0.3, S0=100.0, r=0.02,
                                  expiries=[2.0], strikes=[[0.0, 110.0]], steps=2, n_paths=1000, seed=3,
                                  flags=FLAGS, precision=H.FP64),
    }


def oracle_for(oracle, req, want_finals=True, kind="strict"):
    return oracle.price(req.params, req.S0, req.expiries, req.n_strikes(), req.flat_strikes(), steps=req.steps,
                        n_paths=req.n_paths, seed=req.seed, r=req.r, flags=req.flags, payoffs=req.payoffs,
                        path_offset=req.path_offset, want_finals=want_finals, kind=kind)


@pytest.mark.parametrize("name", sorted(requests()))
def test_pathwise_fp64_martingale(ctx, oracle, name):
    req = requests()[name]
    say("gpu martingale pathwise fp64", name)
    g = ctx.paths(req, 0, req.n_paths)
    o = oracle_for(oracle, req)["finals"]
    o_fma = oracle_for(oracle, req, kind="fma")["finals"]
    mask = ~np.isnan(o)
    assert np.array_equal(mask, ~np.isnan(g))
    # 1e-12, or twice the spread of two valid FP64 evaluations where the recursion amplifies it
    spread = (np.abs(o_fma[mask] - o[mask]) / np.abs(o[mask])).max()
    tol = max(1e-12, 2.0 * spread)
    rel = np.abs(g[mask] - o[mask]) / np.abs(o[mask])
    say(f"  max rel {rel.max():.3e} (tol {tol:.1e})")
    assert rel.max() < tol
    if name != "stiff_c5":
        assert tol == 1e-12


@pytest.mark.parametrize("name", sorted(requests()))
def test_prices_fp64_martingale_equal_oracle(ctx, oracle, name):
    req = requests()[name]
    say("gpu martingale prices fp64", name)
    p, s = ctx.price(req)
    o = oracle_for(oracle, req, want_finals=False)
    np.testing.assert_allclose(p, o["prices"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(s, o["std_errs"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name", sorted(requests()))
def test_pathwise_fp32_martingale_close_to_fp64(ctx, name):
    req = requests()[name]
    say("gpu martingale pathwise fp32", name)
    g64 = ctx.paths(req, 0, req.n_paths)
    req.precision = H.FP32
    g32 = ctx.paths(req, 0, req.n_paths)
    mask = ~np.isnan(g64)
    rel = np.abs(g32[mask] - g64[mask]) / np.abs(g64[mask])
    say(f"  median {np.median(rel):.2e}, 99.9% {np.quantile(rel, 0.999):.2e}")
    assert np.median(rel) < 2e-6 and np.quantile(rel, 0.999) < 1e-3


def test_fp32_martingale_vs_fp64_identical_draws(ctx):
    say("gpu martingale fp32 vs fp64, C2 subset and the stiff case")
    reqs = W.c2(n_paths=200_000, precision=H.FP64, maturities=[0.25, 1.0, 3.0], n_strikes=16)
    reqs.append(H.Request(v0=0.04, theta=0.04, rho=-0.9, kappa=0.5, sigma=0.9, S0=100.0, expiries=[1.0],
                          strikes=[[90.0, 100.0]], steps=2048, n_paths=20_000, seed=77, r=0.02,
                          precision=H.FP64))
    for r in reqs:
        r.flags |= FLAGS
    r64 = ctx.price_batch(reqs)
    for r in reqs:
        r.precision = H.FP32
    r32 = ctx.price_batch(reqs)
    worst = 0.0
    for q, (p64, _), (p32, _) in zip(reqs, r64, r32):
        big = p64 > 1e-2
        e = (np.abs(p32[big] - p64[big]) / p64[big]).max()
        say(f"  T {q.expiries[0]} steps {q.steps}: max rel err fp32 vs fp64 {e:.3e}")
        worst = max(worst, e)
    assert worst < 1e-4


@pytest.mark.parametrize("precision", [H.FP32, H.FP64])
def test_terminal_mean_is_forward_at_scale(ctx, precision):
    """v0 = 0.2 with C5's sigma and rho, four one-year steps: the uncorrected scheme's E[X_T] is
    about 1% low; the corrected one hits the forward within the standard error."""
    n = (1 << 22) if precision == H.FP32 else (1 << 21)
    S0, r, T = 100.0, 0.02, 4.0
    fwd = S0 * math.exp(r * T)
    z = {}
    for mart in (False, True):
        req = H.Request(v0=0.2, theta=0.04, rho=-0.9, kappa=0.5, sigma=0.9, S0=S0, r=r, expiries=[T],
                        strikes=[[0.0]], steps=4, n_paths=n, seed=21, payoffs=H.EUROPEAN,
                        flags=H.DRIFT | (H.MARTINGALE if mart else 0), precision=precision)
        p, s = ctx.price(req)
        z[mart] = (p[0] - fwd) / s[0]
    say("  z-scores of E[X_T] - forward (plain, corrected):", z[False], z[True])
    assert abs(z[True]) < 4.0
    assert abs(z[False]) > 20.0


def test_mixed_batch_equals_single(ctx):
    say("gpu martingale mixed batch == single")
    reqs = list(requests().values())
    mixed = []
    for i, r in enumerate(reqs):
        a = dataclasses.replace(r, precision=(H.FP64, H.FP32, H.FP32, H.FP32_COMPENSATED)[i])
        if i == 1:
            a.flags &= ~H.MARTINGALE
        mixed.append(a)
    together = ctx.price_batch(mixed)
    for r, (p, s) in zip(mixed, together):
        p1, s1 = ctx.price(r)
        np.testing.assert_array_equal(p, p1)
        np.testing.assert_array_equal(s, s1)


def test_all_variants_over_the_uneven_request_distribution(ctx):
    """64 requests drawn from the uneven-request distribution (kappa up to 20, sigma up to 1, rho
    down to -0.95, T from 0.1 to 5 on 252 steps a year), each with the martingale correction, the
    Asian, European, control-variate (regression coefficient) and geometric blocks: every FP32
    price is finite and agrees with the FP64 kernel on identical draws within 1e-4 relative (prices
    above 0.01) or 0.05 standard errors."""
    say("gpu all variants over the uneven request distribution")
    reqs = W.c3(n=64, seed=19, precision=H.FP32, path_scale=2e-3,
                payoffs=H.ASIAN | H.EUROPEAN | H.ASIAN_GEO_CV | H.GEO_ASIAN)
    for r in reqs:
        r.flags |= H.MARTINGALE | H.CV_REGRESSION
    out32 = ctx.price_batch(reqs)
    for r in reqs:
        r.precision = H.FP64
    out64 = ctx.price_batch(reqs)
    worst_rel = worst_se = 0.0
    for (p32, s32), (p64, s64) in zip(out32, out64):
        assert np.all(np.isfinite(p32)) and np.all(np.isfinite(s32))
        big = p64 > 1e-2
        rel = np.abs(p32 - p64)[big] / p64[big]
        se = np.abs(p32 - p64) / np.maximum(s64, 1e-300)
        worst_rel = max(worst_rel, float(rel.max(initial=0.0)))
        worst_se = max(worst_se, float(np.where(s64 > 0, se, 0.0).max()))
        assert np.all((rel < 1e-4) | (se[big] < 0.05))
    say(f"  worst FP32 vs FP64: {worst_rel:.2e} relative, {worst_se:.3f} SE")
