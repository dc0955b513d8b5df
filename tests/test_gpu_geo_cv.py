"""GPU: the geometric-average control variate of the Asian call (payoff ASIAN_GEO_CV).

(1) The exact geometric prices computed on the GPU (hexo_geo_asian_price, the kernel every
    control-variate batch runs once) equal the oracle's to 1e-9 S0: short and long maturities,
    several expiries with an empty chain, the corrected trapezoid, the stiff and the peaked
    (slowly decaying characteristic function) cases, positive correlation, more than 64 strikes,
    a zero strike.
(2) FP64 kernel: the Monte-Carlo part of the control-variate price, mean(max(A-K,0) - max(G-K,0)),
    equals the oracle's on the same draws to 1e-12 relative; whole prices to 2e-9 S0.
(3) FP32 against FP64 on identical draws within 1e-4; the standard error is a small fraction of the
    plain Asian estimator's, and the two estimators agree within the plain standard error.
(4) Mixed batches of all kernel classes return what each request returns alone.
(5) The regression-coefficient estimator formed on the GPU (flag CV_REGRESSION) equals the host's
    cv_optimal over the same paths' blocks.
"""
import dataclasses

import numpy as np
import pytest

import hestonexotics_amd as H
from hestonexotics_amd import workloads as W

pytestmark = pytest.mark.gpu

BASE = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0)


def say(*a):
    print(*a, flush=True)


def geo_requests():
    return {
        "c1_like": H.Request(**BASE, expiries=[1.0], strikes=[[80.0, 100.0, 120.0]], steps=252, n_paths=2000,
                             seed=42, r=0.02, flags=H.DRIFT | H.DISCOUNT),
        "multi_expiry": H.Request(**BASE, expiries=[0.3, 0.31, 0.7, 1.2],
                                  strikes=[[95, 100], [], [100, 105, 120], [0.0, 100]], steps=7, n_paths=777,
                                  seed=11, r=0.02, flags=H.DRIFT | H.TRAPEZOID),
        "stiff_c5": H.Request(v0=0.04, theta=0.04, rho=-0.9, kappa=0.5, sigma=0.9, S0=100.0, expiries=[1.0],
                              strikes=[[90.0, 100.0, 110.0]], steps=128, n_paths=640, seed=5, r=0.02,
                              flags=H.DRIFT),
        "peaked": H.Request(v0=0.01, theta=0.01, rho=-0.5, kappa=0.5, sigma=1.0, S0=100.0, expiries=[0.1],
                            strikes=[[97.0, 100.0, 103.0]], steps=25, n_paths=500, seed=8),
        "positive_rho": H.Request(v0=0.04, theta=0.09, rho=0.4, kappa=1.5, sigma=0.4, S0=100.0, expiries=[2.0],
                                  strikes=[[90.0, 110.0]], steps=104, n_paths=500, seed=3, r=0.03,
                                  flags=H.DRIFT | H.DISCOUNT),
        "wide_strikes": H.Request(**BASE, expiries=[0.5], strikes=[list(np.linspace(60.0, 140.0, 70))], steps=26,
                                  n_paths=300, seed=4),
    }


def cv(req, precision, payoffs=H.ASIAN_GEO_CV, flags=0):
    return dataclasses.replace(req, precision=precision, payoffs=payoffs, flags=req.flags | flags)


def oracle_price(oracle, req, kind="strict"):
    return oracle.price(req.params, req.S0, req.expiries, req.n_strikes(), req.flat_strikes(), steps=req.steps,
                        n_paths=req.n_paths, seed=req.seed, r=req.r, flags=req.flags, payoffs=req.payoffs,
                        path_offset=req.path_offset, kind=kind)


def discounts(req):
    T = np.repeat(np.asarray(req.expiries, np.float64), req.n_strikes())
    return np.exp(-req.r * T) if req.flags & H.DISCOUNT else np.ones_like(T)


@pytest.mark.parametrize("name", sorted(geo_requests()))
def test_exact_geometric_price_matches_oracle(ctx, oracle, name):
    req = geo_requests()[name]
    say("gpu geo exact", name)
    g = ctx.geo_asian_price(req)
    o = oracle.geo_asian_price(req.params, req.S0, req.expiries, req.n_strikes(), req.flat_strikes(),
                               steps=req.steps, r=req.r if req.flags & H.DRIFT else 0.0, flags=req.flags)
    o = o * discounts(req)
    say(f"  max |gpu - oracle| / S0 = {np.abs(g - o).max() / req.S0:.2e}")
    assert np.all(np.isfinite(g))
    np.testing.assert_allclose(g, o, rtol=0, atol=1e-9 * req.S0)


@pytest.mark.parametrize("name", sorted(geo_requests()))
def test_fp64_control_variate_prices_equal_oracle(ctx, oracle, name):
    req = cv(geo_requests()[name], H.FP64, payoffs=H.ASIAN | H.ASIAN_GEO_CV)
    say("gpu geo cv fp64 prices", name)
    p, s = ctx.price(req)
    o = oracle_price(oracle, req)
    n = req.n_opts
    np.testing.assert_allclose(p[:n], o["prices"][:n], rtol=1e-12, atol=1e-12)   # plain Asian block
    # the Monte-Carlo part of the control-variate block
    exact_g = ctx.geo_asian_price(req)
    exact_o = oracle.geo_asian_price(req.params, req.S0, req.expiries, req.n_strikes(), req.flat_strikes(),
                                     steps=req.steps, r=req.r if req.flags & H.DRIFT else 0.0,
                                     flags=req.flags) * discounts(req)
    mc_g, mc_o = p[n:] - exact_g, o["prices"][n:] - exact_o
    # a - g cancels most of a: the tolerance is absolute, 1e-11 S0, or twice the spread of two
    # valid FP64 evaluations of the oracle (with and without FMA contraction)
    o_fma = oracle_price(oracle, req, kind="fma")
    spread = np.abs(o_fma["prices"][n:] - o["prices"][n:]).max()
    tol = max(1e-11 * req.S0, 2.0 * spread)
    say(f"  MC part max |gpu - oracle| {np.abs(mc_g - mc_o).max():.2e} (tol {tol:.1e})")
    np.testing.assert_allclose(mc_g, mc_o, rtol=0, atol=tol)
    np.testing.assert_allclose(s[n:], o["std_errs"][n:], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(p[n:], o["prices"][n:], rtol=0, atol=2e-9 * req.S0)


def test_fp32_control_variate_vs_fp64_and_variance_reduction(ctx):
    say("gpu geo cv fp32 vs fp64, grid subset")
    reqs = W.c2(n_paths=200_000, precision=H.FP64, maturities=[0.25, 1.0, 3.0], n_strikes=16)
    for r in reqs:
        r.payoffs = H.ASIAN | H.ASIAN_GEO_CV
    r64 = ctx.price_batch(reqs)
    for r in reqs:
        r.precision = H.FP32
    r32 = ctx.price_batch(reqs)
    worst, ratios = 0.0, []
    for q, (p64, s64), (p32, s32) in zip(reqs, r64, r32):
        n = q.n_opts
        big = p64[n:] > 1e-2
        worst = max(worst, (np.abs(p32[n:][big] - p64[n:][big]) / p64[n:][big]).max())
        atm = np.argmin(np.abs(q.flat_strikes() - 100.0))
        ratios.append(s32[n + atm] / s32[atm])
        # the two estimators of the same price agree within the plain one's standard error
        assert np.all(np.abs(p32[n:] - p32[:n]) < 4 * s32[:n] + 2e-3), (q.expiries, p32[n:] - p32[:n])
    say("  max rel err fp32 vs fp64:", worst, " ATM SE ratio cv/plain:", ratios)
    assert worst < 1e-4
    assert max(ratios) < 0.1


@pytest.mark.parametrize("name", sorted(geo_requests()))
def test_fp32_control_variate_close_to_fp64(ctx, name):
    """FP32 kernels (plain and compensated reduction, with and without the martingale
    correction) against FP64 on identical draws, every request shape: multi-expiry runs exercise
    the log-price accumulation across step-width changes."""
    say("gpu geo cv fp32 vs fp64", name)
    for flags in (0, H.MARTINGALE):
        p64, _ = ctx.price(cv(geo_requests()[name], H.FP64, H.ASIAN | H.ASIAN_GEO_CV, flags))
        for prec in (H.FP32, H.FP32_COMPENSATED):
            p32, _ = ctx.price(cv(geo_requests()[name], prec, H.ASIAN | H.ASIAN_GEO_CV, flags))
            big = np.abs(p64) > 1e-2
            rel = np.abs(p32[big] - p64[big]) / np.abs(p64[big])
            say(f"  flags {flags} precision {prec}: max rel {rel.max():.2e}")
            assert rel.max() < 1e-4


@pytest.mark.parametrize("name", sorted(geo_requests()))
def test_geometric_block_and_regression_estimator(ctx, oracle, name):
    """The plain geometric-call block (GEO_ASIAN) equals the oracle's on the same draws (FP64
    kernel); with it the regression-coefficient control variate is never worse than the plain
    estimator or the coefficient-1 one (FP32 kernel)."""
    say("gpu geo block", name)
    req = cv(geo_requests()[name], H.FP64, H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN)
    p, s = ctx.price(req)
    o = oracle_price(oracle, req)
    n = req.n_opts
    o_fma = oracle_price(oracle, req, kind="fma")
    tol = max(1e-12 * req.S0, 2.0 * np.abs(o_fma["prices"][2 * n:] - o["prices"][2 * n:]).max())
    np.testing.assert_allclose(p[2 * n:], o["prices"][2 * n:], rtol=1e-12, atol=tol)
    req.precision = H.FP32
    p, s = ctx.price(req)
    exact = ctx.geo_asian_price(req)
    price, se, beta = H.cv_optimal((p[:n], s[:n]), (p[2 * n:], s[2 * n:]), (p[n:2 * n], s[n:2 * n]), exact)
    say(f"  beta {np.round(beta, 3)}, SE plain {s[:n]}, cv {s[n:2 * n]}, optimal {se}")
    assert np.all(se <= s[:n] * (1 + 1e-9)) and np.all(se <= s[n:2 * n] * (1 + 1e-9))
    assert np.all(np.abs(price - p[:n]) < 4 * s[:n] + 1e-12)


@pytest.mark.parametrize("precision", [H.FP64, H.FP32, H.FP32_COMPENSATED])
@pytest.mark.parametrize("name", sorted(geo_requests()))
def test_regression_flag_equals_host_estimator(ctx, name, precision):
    """CV_REGRESSION forms A - beta (G - E[G]) in the final reduction on the GPU: its block equals
    cv_optimal applied to the same paths' Asian, control-variate and geometric blocks priced
    without the flag; the other blocks are unchanged bit for bit."""
    say("gpu cv regression flag", name, precision)
    req = cv(geo_requests()[name], precision, H.ASIAN | H.EUROPEAN | H.ASIAN_GEO_CV | H.GEO_ASIAN)
    p, s = ctx.price(req)
    pr, sr = ctx.price(cv(req, precision, req.payoffs, H.CV_REGRESSION))
    n = req.n_opts
    exact = ctx.geo_asian_price(req)
    price, se, _ = H.cv_optimal((p[:n], s[:n]), (p[3 * n:], s[3 * n:]), (p[2 * n:3 * n], s[2 * n:3 * n]), exact)
    for blk in (0, 1, 3):
        np.testing.assert_array_equal(pr[blk * n:(blk + 1) * n], p[blk * n:(blk + 1) * n])
        np.testing.assert_array_equal(sr[blk * n:(blk + 1) * n], s[blk * n:(blk + 1) * n])
    np.testing.assert_allclose(pr[2 * n:3 * n], price, rtol=1e-12, atol=1e-12 * req.S0)
    # the standard error is a difference of variances: compared on the variance scale
    np.testing.assert_allclose(sr[2 * n:3 * n] ** 2, se ** 2, rtol=1e-9, atol=1e-12 * s[:n].max() ** 2)
    assert np.all(sr[2 * n:3 * n] <= s[:n] * (1 + 1e-9))


def test_exact_prices_over_the_uneven_request_distribution(ctx, oracle):
    """The 512 uneven requests' distribution (kappa up to 20, sigma up to 1, T from 0.1 to 5 on
    252 steps a year) with the control variate, in one batch: every exact geometric price is
    finite, and a sample of them equals the oracle's to 1e-9 S0."""
    say("gpu geo exact over the uneven request distribution")
    reqs = W.c3(precision=H.FP32, path_scale=1e-4, payoffs=H.ASIAN_GEO_CV)
    out = ctx.price_batch(reqs)
    assert all(np.all(np.isfinite(p)) for p, _ in out)
    rng = np.random.default_rng(1)
    worst = 0.0
    for i in rng.choice(len(reqs), 8, replace=False):
        q = reqs[int(i)]
        g = ctx.geo_asian_price(q)
        o = oracle.geo_asian_price(q.params, q.S0, q.expiries, q.n_strikes(), q.flat_strikes(), steps=q.steps,
                                   r=q.r, flags=q.flags) * discounts(q)
        worst = max(worst, float(np.abs(g - o).max() / q.S0))
    say(f"  max |gpu - oracle| / S0 over the sample: {worst:.2e}")
    assert worst < 1e-9


def test_mixed_batch_equals_single(ctx):
    say("gpu geo cv mixed batch == single")
    base = geo_requests()["c1_like"]
    variants = [
        cv(base, H.FP64, H.ASIAN | H.ASIAN_GEO_CV),
        cv(base, H.FP32, H.ASIAN),
        cv(base, H.FP32, H.ASIAN_GEO_CV, H.MARTINGALE),
        cv(base, H.FP32, H.ASIAN | H.EUROPEAN | H.ASIAN_GEO_CV),
        cv(geo_requests()["multi_expiry"], H.FP32_COMPENSATED, H.EUROPEAN | H.ASIAN_GEO_CV),
        cv(geo_requests()["stiff_c5"], H.FP32, H.ASIAN_GEO_CV, H.MARTINGALE),
        cv(geo_requests()["peaked"], H.FP32, H.GEO_ASIAN),
        cv(geo_requests()["positive_rho"], H.FP64, H.EUROPEAN | H.GEO_ASIAN),
        cv(geo_requests()["wide_strikes"], H.FP32, H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN, H.CV_REGRESSION),
    ]
    together = ctx.price_batch(variants)
    for r, (p, s) in zip(variants, together):
        p1, s1 = ctx.price(r)
        assert np.all(np.isfinite(p))
        np.testing.assert_array_equal(p, p1)
        np.testing.assert_array_equal(s, s1)


def test_control_variate_rejects_one_step(ctx):
    req = cv(geo_requests()["c1_like"], H.FP32)
    req.steps = 1
    with pytest.raises(H.HexoError):
        ctx.price(req)
    with pytest.raises(H.HexoError):
        ctx.geo_asian_price(req)


@pytest.mark.parametrize("case", [
    dict(v0=0.0, theta=1.21e-4, rho=-0.622, kappa=0.367, sigma=0.0823, T=0.0543, steps=60),
    dict(v0=2.2e-4, theta=2.96e-3, rho=0.902, kappa=0.3, sigma=1.8, T=0.0112, steps=330),
    dict(v0=1e-4, theta=1e-4, rho=-0.99, kappa=1.0, sigma=0.05, T=0.02, steps=20),
])
def test_exact_price_near_deterministic(ctx, oracle, case):
    """Nearly deterministic geometric averages (variance near 0, short expiries): wide panels whose
    first one is summed in sub-panels. The GPU's exact prices equal the oracle's to 1e-9 S0 and
    are call prices (non-negative, at least the intrinsic value E[G] - K, decreasing in K)."""
    say("gpu geo exact near-deterministic", case)
    p = [case["v0"], case["theta"], case["rho"], case["kappa"], case["sigma"]]
    t, w = oracle.asian_weights([case["T"]], [1], case["steps"], 0, H.DRIFT)
    W = w.sum()
    EG = W * 100.0 * oracle.geo_cf(p, 0.02, t, w / W, 1.0 + 0j).real
    K = list(np.linspace(0.9, 1.1, 11) * EG)
    req = H.Request(v0=case["v0"], theta=case["theta"], rho=case["rho"], kappa=case["kappa"], sigma=case["sigma"],
                    S0=100.0, expiries=[case["T"]], strikes=[K], steps=case["steps"], n_paths=64, r=0.02,
                    flags=H.DRIFT, payoffs=H.ASIAN_GEO_CV, precision=H.FP64)
    g = ctx.geo_asian_price(req)
    o = oracle.geo_price_weights(p, 100.0, 0.02, t, w, K)
    say(f"  max |gpu - oracle| / S0 {np.abs(g - o).max() / 100.0:.2e}, min price {g.min():.3e}")
    np.testing.assert_allclose(g, o, rtol=0, atol=1e-9 * 100.0)
    tol = 1e-9 * EG
    assert np.all(g >= -tol) and np.all(g >= EG - np.array(K) - tol) and np.all(np.diff(g) <= tol)
