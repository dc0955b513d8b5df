"""GPU SWIFT European pricer through the C ABI (hexo_swift_price), FP64.

Pins: the reference's own golden prices and golden Jacobian (src/UnitTest.cpp:221-261, :278-478,
tolerance sum |diff| < 1e-9), the oracle's SWIFT restatement (oracle/swift_oracle.c) on dynamic
grids, batch invariance, and the Monte-Carlo European prices of the grid workload.
"""
import json
import os

import numpy as np
import pytest

import hestonexotics_amd as H
from hestonexotics_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


def golden_chains():
    g = json.load(open(os.path.join(GOLD, "swift_reference_prices.json")))
    p = g["params"]
    chains = [H.SwiftChain(p["v0"], p["theta"], p["rho"], p["kappa"], p["sigma"], S0=g["S"], T=T, strikes=K,
                           r=g["r"], grid=sp)
              for T, K, sp in zip(g["expiries"], g["strikes"], g["swift_params"])]
    return g, chains


def test_swift_reference_golden_prices_and_jacobian(ctx):
    say("gpu swift golden")
    g, chains = golden_chains()
    res = ctx.swift_price(chains, grads=True)
    prices = np.concatenate([r[0] for r in res])
    grads = np.vstack([r[1] for r in res])
    assert np.abs(prices - np.array(g["prices"])).sum() < 1e-9
    gold = np.array(g["gradients"]).reshape(-1, 5)[:, [4, 1, 3, 0, 2]]
    assert np.abs(grads - gold).sum() < 1e-9
    for (_, _, used), sp in zip(res, g["swift_params"]):
        assert np.allclose(used, sp, rtol=0, atol=0)


def dynamic_chains():
    out = [H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=T, strikes=list(np.linspace(70, 130, 64)), r=0.02)
           for T in W.C2_MATURITIES]
    out.append(H.SwiftChain(0.05, 0.04, -0.6, 1.5, 0.5, S0=100.0, T=0.75, strikes=[80.0, 100.0, 125.0], r=0.01))
    out.append(H.SwiftChain(0.2, 0.1, -0.9, 5.0, 0.9, S0=50.0, T=2.5, strikes=[30.0, 50.0, 70.0], r=0.03))
    return out


def test_swift_matches_oracle_on_dynamic_grids(ctx, oracle):
    say("gpu swift vs oracle, dynamic grids")
    chains = dynamic_chains()
    res = ctx.swift_price(chains, grads=True)
    for ch, (p, g, used) in zip(chains, res):
        op, oused, og = oracle.swift_price([ch.v0, ch.theta, ch.rho, ch.kappa, ch.sigma], ch.S0, ch.r, ch.T,
                                           ch.strikes, trunc=ch.trunc, grads=True)
        # integer grid parameters identical; the range ends agree to rounding (the cumulant
        # expressions are summed in a different order)
        assert [used[i] for i in (0, 1, 5, 6, 7)] == [oused[i] for i in (0, 1, 5, 6, 7)]
        assert np.allclose(used, oused, rtol=1e-13, atol=0)
        np.testing.assert_allclose(p, op, rtol=1e-10, atol=1e-11 * ch.S0)
        np.testing.assert_allclose(g, og, rtol=1e-9, atol=1e-10 * ch.S0)


def test_gpu_grid_choice_equals_host(ctx):
    """The grids hexo_swift_price chooses on the GPU (one wavefront per chain, every scale tried at
    once) equal hexo_swift_grid_for's host loop over the calibration batch's 512 chains and a wide
    parameter box: the integer parameters (m, 2^m, k1, k2, J) exactly; the range ends to 1e-10
    relative. The fourth cumulant's bracket cancels terms ~1e4 times its value at small kappa T,
    so the device's and the host's libm (both within an ulp) leave the ends up to 2e-11 apart
    (measured); the prices move by the same relative amount."""
    say("gpu swift grid choice == host")
    import bench
    rng = np.random.default_rng(17)
    chains = bench.swift_chains(64)
    for _ in range(200):
        v0, th = rng.uniform(0.005, 0.5, 2)
        rho, kappa, sigma = rng.uniform(-0.99, 0.99), rng.uniform(0.05, 20.0), rng.uniform(0.05, 2.0)
        T = float(np.exp(rng.uniform(np.log(0.02), np.log(10.0))))
        K = sorted(rng.uniform(40.0, 250.0, int(rng.integers(1, 40))))
        chains.append(H.SwiftChain(v0, th, rho, kappa, sigma, S0=100.0, T=T, strikes=list(K),
                                   r=float(rng.uniform(0.0, 0.08)), trunc=float(10.0 ** rng.uniform(-12, -6))))
    ok = []
    for ch in chains:
        try:
            ok.append((ch, H.swift_grid(ch)))
        except H.HexoError:
            pass  # grids the host refuses (J beyond HEXO_SWIFT_MAX_J) are refused by the call too
    assert len(ok) > 600
    res = ctx.swift_price([c for c, _ in ok])
    for (ch, host), (_, _, used) in zip(ok, res):
        assert [used[i] for i in (0, 1, 5, 6, 7)] == [host[i] for i in (0, 1, 5, 6, 7)], (ch, used, host)
        assert np.allclose(used, host, rtol=1e-10, atol=0)


def test_swift_batch_equals_single(ctx):
    say("gpu swift batch == single")
    chains = dynamic_chains()
    together = ctx.swift_price(chains, grads=True)
    for ch, (p, g, used) in zip(chains, together):
        p1, g1, u1 = ctx.swift_price([ch], grads=True)[0]
        np.testing.assert_array_equal(p, p1)
        np.testing.assert_array_equal(g, g1)
    plain = ctx.swift_price(chains, grads=False)
    for (p, _, _), (q, g, _) in zip(together, plain):
        np.testing.assert_array_equal(p, q)
        assert g is None


def test_swift_grid_kernel_strikes_beyond_one_wavefront(ctx, oracle):
    """The grid kernel takes each chain's extreme strikes with its 64 lanes striding over the
    strikes: chains of 1, 63, 64, 65 and 300 unsorted strikes get the host's grid, and the oracle's
    prices."""
    say("gpu swift grid kernel, many strikes")
    rng = np.random.default_rng(23)
    chains = [H.SwiftChain(0.05, 0.05, -0.5, 1.5, 0.4, S0=100.0, T=T, r=0.02,
                           strikes=list(rng.permutation(np.linspace(40.0, 220.0, k))))
              for k, T in ((1, 0.3), (63, 0.7), (64, 1.0), (65, 2.0), (300, 4.0))]
    res = ctx.swift_price(chains)
    for ch, (p, _, used) in zip(chains, res):
        host = H.swift_grid(ch)
        assert [used[i] for i in (0, 1, 5, 6, 7)] == [host[i] for i in (0, 1, 5, 6, 7)]
        assert np.allclose(used, host, rtol=1e-10, atol=0)
        op, _ = oracle.swift_price([ch.v0, ch.theta, ch.rho, ch.kappa, ch.sigma], ch.S0, ch.r, ch.T,
                                   ch.strikes, trunc=ch.trunc)
        np.testing.assert_allclose(p, op, rtol=1e-10, atol=1e-11 * ch.S0)


def test_swift_rejects_bad_strikes_in_any_chain(ctx):
    """A non-positive or non-finite strike is refused whether the chain's grid is chosen on the
    GPU (checked by the grid kernel) or given (checked on the host), wherever it sits."""
    ok = H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=list(np.linspace(70, 130, 100)))
    fixed = H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=[90.0, 110.0],
                         grid=list(H.swift_grid(ok)))
    for bad in (0.0, -5.0, float("nan"), float("inf")):
        for pos in (0, 77, 99):
            k = list(ok.strikes)
            k[pos] = bad
            with pytest.raises(H.HexoError):
                ctx.swift_price([ok, H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=k)])
        with pytest.raises(H.HexoError):
            ctx.swift_price([ok, H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=[90.0, bad],
                                              grid=list(fixed.grid))])
    # and the context still prices afterwards
    p, _, _ = ctx.swift_price([ok, fixed])[1]
    assert np.all(np.isfinite(p))


def test_swift_rejects_malformed(ctx):
    bad = H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=[100.0], grid=[5, 32, 5.65, -1, 1, 10, 5, 128])
    with pytest.raises(H.HexoError):
        ctx.swift_price([bad])
    with pytest.raises(H.HexoError):
        ctx.swift_price([H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=[-1.0])])


def test_mc_european_prices_agree_with_swift(ctx):
    """The grid workload's European prices (FP32 Monte-Carlo, 2e5 paths, 52 steps/yr) against
    SWIFT on the same model: within 4 standard errors plus the QE time-discretisation bias."""
    say("gpu mc european vs swift")
    mats = [0.25, 1.0, 3.0]
    reqs = W.c2(n_paths=200_000, maturities=mats, n_strikes=16)
    mc = ctx.price_batch(reqs)
    chains = [H.SwiftChain(r.v0, r.theta, r.rho, r.kappa, r.sigma, S0=r.S0, T=r.expiries[0], strikes=r.strikes[0],
                           r=r.r) for r in reqs]
    sw = ctx.swift_price(chains)
    for r, (p, s), (q, _, _) in zip(reqs, mc, sw):
        euro, se = p[len(r.strikes[0]):], s[len(r.strikes[0]):]
        dev = np.abs(euro - q)
        say("  T", r.expiries[0], "max |mc - swift| / se", (dev / se).max())
        assert np.all(dev < 4 * se + 0.02)


def test_swift_fft_and_direct_sum_paths(ctx, oracle):
    """Chains at both sides of the LDS-FFT size limit (J <= 4096: radix-2 FFTs in one workgroup;
    larger J: direct sums), and a tiny grid, against the oracle's direct DFTs."""
    say("gpu swift fft / direct paths")
    base = H.SwiftChain(0.04, 0.04, -0.7, 2.0, 0.3, S0=100.0, T=1.0, strikes=[80.0, 100.0, 120.0], r=0.02)
    g = list(H.swift_grid(base))
    kw = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, T=1.0, strikes=[80.0, 100.0, 120.0],
              r=0.02)
    chains = [H.SwiftChain(**kw, grid=g[:7] + [J]) for J in (4096, 8192, 16384)]
    chains.append(H.SwiftChain(**kw, grid=[1, 2, 2 ** 0.5, -1.0, 1.0, -2, 2, 2]))
    res = ctx.swift_price(chains, grads=True)
    for ch, (p, gr, used) in zip(chains, res):
        op, _, og = oracle.swift_price([ch.v0, ch.theta, ch.rho, ch.kappa, ch.sigma], ch.S0, ch.r, ch.T, ch.strikes,
                                       swift_params=ch.grid, grads=True)
        say("  J", used[7], "max |gpu - oracle|", np.abs(p - op).max())
        np.testing.assert_allclose(p, op, rtol=1e-10, atol=1e-9)
        np.testing.assert_allclose(gr, og, rtol=1e-9, atol=1e-9)


def test_swift_calibration_batch_within_conditioning(ctx, oracle):
    """The bench's 512-chain calibration batch against the oracle. Chains with a wide truncation
    range carry payoff coefficients of size S e^upper (1.3e7 S at upper = 16.4 for T = 5 and
    sigma = 0.7) that cancel in the transforms, so two FP64 evaluations in different orders (LDS
    FFT here, direct DFTs in the oracle, FFTW in the reference) agree only to about
    eps S e^upper: the tolerance is scaled by that conditioning."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    say("gpu swift calibration batch vs oracle")
    chains = bench.swift_chains(64)
    res = ctx.swift_price(chains, grads=True)
    eps = np.finfo(np.float64).eps
    worst = 0.0
    for ch, (p, g, used) in zip(chains, res):
        op, _, og = oracle.swift_price([ch.v0, ch.theta, ch.rho, ch.kappa, ch.sigma], ch.S0, ch.r, ch.T, ch.strikes,
                                       trunc=ch.trunc, grads=True)
        cond = eps * ch.S0 * max(1.0, np.exp(used[4]))
        worst = max(worst, np.abs(p - op).max() / cond)
        assert np.abs(p - op).max() <= 1e-11 * ch.S0 + 8 * cond
        assert np.abs(g - og).max() <= 1e-10 * ch.S0 + 64 * cond * (1 + np.abs(og).max() / ch.S0)
    say("  max |gpu - oracle| / (eps S e^upper):", worst)
