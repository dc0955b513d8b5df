"""GPU: requests at the limits the C ABI accepts, against the oracle.

(1) Global path indices at and beyond 2^32 (the high word of the Philox counter's path index): a
    request whose paths cross 2^32 and a ragged one far above it, pathwise FP64 against the oracle.
(2) The largest strike count of a request (2^20) and the largest expiry count (4096): FP64 prices
    equal the oracle's; FP32 on the same draws within 1e-4.
(3) The largest step count per expiry (2^22), FP32 with the martingale correction: at K = 0 the
    Asian and European prices are E[A] = sum_j w_j S0 e^{r t_j} and E[X_T] = S0 e^{r T}, met within
    5 standard errors.
(4) hexo_paths refuses windows beyond its limits before it allocates or writes.
"""
import dataclasses

import numpy as np
import pytest

import hestonexotics_amd as H

pytestmark = pytest.mark.gpu

BASE = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0)


def say(*a):
    print(*a, flush=True)


def oracle_run(oracle, req, want_finals=False):
    return oracle.price(req.params, req.S0, req.expiries, req.n_strikes(), req.flat_strikes(), steps=req.steps,
                        n_paths=req.n_paths, seed=req.seed, r=req.r, flags=req.flags, payoffs=req.payoffs,
                        path_offset=req.path_offset, want_finals=want_finals)


@pytest.mark.parametrize("offset,n_paths", [(2 ** 32 - 128, 256), (2 ** 40 + 192, 200)])
def test_path_indices_beyond_2_32(ctx, oracle, offset, n_paths):
    say("gpu path indices beyond 2^32", offset, n_paths)
    req = H.Request(**BASE, expiries=[0.25, 1.0], strikes=[[100.0], [90.0, 110.0]], steps=13, n_paths=n_paths,
                    seed=77, path_offset=offset, precision=H.FP64)
    g = ctx.paths(req, 0, n_paths)
    o = oracle_run(oracle, req, want_finals=True)["finals"]
    assert np.all(np.isfinite(g))
    np.testing.assert_allclose(g, o, rtol=1e-12)
    # the high word takes part: the paths at or above 2^32 differ from those with the same low words
    # below 2^32
    hi = max(0, 2 ** 32 - offset)
    low = dataclasses.replace(req, path_offset=(offset + hi) % (2 ** 32), n_paths=n_paths - hi)
    assert not np.allclose(ctx.paths(low, 0, n_paths - hi), g[hi:])
    p, _ = ctx.price(req)
    np.testing.assert_allclose(p, oracle_run(oracle, req)["prices"], rtol=1e-12, atol=1e-12 * req.S0)


def compare_fp64_fp32(ctx, oracle, req):
    p64, s64 = ctx.price(req)
    o = oracle_run(oracle, req)
    np.testing.assert_allclose(p64, o["prices"], rtol=1e-12, atol=1e-12 * req.S0)
    np.testing.assert_allclose(s64, o["std_errs"], rtol=1e-9, atol=1e-12 * req.S0)
    req.precision = H.FP32
    p32, _ = ctx.price(req)
    big = p64 > 1e-2
    assert np.all(np.isfinite(p32))
    rel = np.abs(p32[big] - p64[big]) / p64[big]
    say(f"  {p64.size} prices, FP32 vs FP64 max rel {rel.max():.2e}")
    assert rel.max() < 1e-4


def test_largest_strike_count(ctx, oracle):
    say("gpu 2^20 strikes")
    req = H.Request(**BASE, expiries=[0.5], strikes=[list(np.linspace(50.0, 150.0, 2 ** 20))], steps=4, n_paths=192,
                    seed=5, r=0.02, flags=H.DRIFT | H.DISCOUNT, precision=H.FP64)
    assert H.validate(req) == 0
    compare_fp64_fp32(ctx, oracle, req)
    req.strikes = [list(np.linspace(50.0, 150.0, 2 ** 20 + 1))]
    assert H.validate(req) == -1


def test_largest_expiry_count(ctx, oracle):
    say("gpu 4096 expiries")
    T = list(np.linspace(0.01, 5.0, 4096))
    req = H.Request(**BASE, expiries=T, strikes=[[100.0]] * len(T), steps=3, n_paths=130, seed=6, precision=H.FP64)
    assert H.validate(req) == 0
    compare_fp64_fp32(ctx, oracle, req)
    req.expiries, req.strikes = T + [6.0], [[100.0]] * (len(T) + 1)
    assert H.validate(req) == -1


def test_largest_step_count(ctx):
    say("gpu 2^22 steps per expiry")
    r = 0.02
    req = H.Request(**BASE, expiries=[1.0], strikes=[[0.0]], steps=2 ** 22, n_paths=64, seed=8, r=r,
                    flags=H.DRIFT | H.MARTINGALE, precision=H.FP32)
    assert H.validate(req) == 0
    p, s = ctx.price(req)
    t, w = H.asian_weights(req, 0)
    expect = [float(np.sum(w * req.S0 * np.exp(r * t))), req.S0 * np.exp(r * 1.0)]
    say(f"  prices {p}, SE {s}, closed forms {expect}")
    assert np.all(np.isfinite(p)) and np.all(s > 0)
    assert np.all(np.abs(p - expect) < 5 * s)
    req.steps = 2 ** 22 + 1
    assert H.validate(req) == -1


def test_paths_window_limits(ctx):
    """hexo_paths refuses windows beyond 2^26 paths or 2^30 finals (paths x expiries x 2) before it
    allocates or writes anything (the output buffer here holds one path's 17 x 2 finals)."""
    import ctypes as C
    from hestonexotics_amd import _lib
    lib = _lib.load()
    buf = np.zeros(17 * 2)
    out = buf.ctypes.data_as(C.POINTER(C.c_double))
    T = [0.1 * (i + 1) for i in range(17)]
    req = H.Request(**BASE, expiries=T, strikes=[[100.0]] * 17, steps=4, n_paths=2 ** 26 + 64, seed=1,
                    precision=H.FP32)
    q = req.to_c()
    assert lib.hexo_paths(ctx.h, C.byref(q), 0, 2 ** 26 + 1, out) == -1      # too many paths
    assert lib.hexo_paths(ctx.h, C.byref(q), 0, 2 ** 25, out) == -1          # 2^25 x 17 x 2 > 2^30 finals
    assert "2^30 finals" in lib.hexo_ctx_last_error(ctx.h).decode()
    assert lib.hexo_paths(ctx.h, C.byref(q), 0, 1, out) == 0 and np.all(np.isfinite(buf))
