"""The C ABI without a GPU: the library loads, exports every entry point include/hexo.h declares,
and its host-only functions (validation, output count, time grid, error strings) behave."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hestonexotics_amd as H
from hestonexotics_amd import _lib, workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hexo.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|uint64_t|const char\*)\s+(hexo_[a-z0-9_]+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_version_and_errors():
    lib = _lib.load()
    assert b"gfx950" in lib.hexo_version()
    for code in range(0, -11, -1):
        assert lib.hexo_strerror(code) != b"unknown status"
    assert lib.hexo_strerror(-99) == b"unknown status"


def good():
    return W.c1(n_paths=1000)


MUTATIONS = {
    "steps": lambda r: setattr(r, "steps", 0),
    "sigma": lambda r: setattr(r, "sigma", 0.0),
    "kappa": lambda r: setattr(r, "kappa", -1.0),
    "theta": lambda r: setattr(r, "theta", 0.0),
    "v0": lambda r: setattr(r, "v0", -0.01),
    "rho": lambda r: setattr(r, "rho", 1.5),
    "S0": lambda r: setattr(r, "S0", 0.0),
    "paths": lambda r: setattr(r, "n_paths", 0),
    "expiry_order": lambda r: (setattr(r, "expiries", [1.0, 0.5]), setattr(r, "strikes", [[100.0], [100.0]])),
    "expiry_nan": lambda r: setattr(r, "expiries", [float("nan")]),
    "strike_nan": lambda r: setattr(r, "strikes", [[float("inf")]]),
    "no_strikes": lambda r: setattr(r, "strikes", [[]]),
    "payoffs": lambda r: setattr(r, "payoffs", 0),
    "payoff_bits": lambda r: setattr(r, "payoffs", 16),
    "flags": lambda r: setattr(r, "flags", 32),
    "precision": lambda r: setattr(r, "precision", 7),
    "path_offset_unaligned": lambda r: setattr(r, "path_offset", 100),
    # the control variate needs an Asian average with weight sum >= 1/4 (one step has 0)
    "control_variate_one_step": lambda r: (setattr(r, "payoffs", H.ASIAN_GEO_CV), setattr(r, "steps", 1)),
    # FP32 below sigma^2 dt = 1e-18 max(v0, theta): the variance step's moment ratio would overflow
    "fp32_sigma_too_small": lambda r: (setattr(r, "sigma", 1e-9), setattr(r, "precision", H.FP32)),
    # the regression coefficient needs the Asian and geometric blocks beside the control variate
    "cv_regression_without_geo": lambda r: (setattr(r, "payoffs", H.ASIAN | H.ASIAN_GEO_CV),
                                            setattr(r, "flags", H.CV_REGRESSION)),
    "cv_regression_without_cv": lambda r: (setattr(r, "payoffs", H.ASIAN | H.GEO_ASIAN),
                                           setattr(r, "flags", H.CV_REGRESSION)),
    # the control variate's exact price does not converge at rho = +-1
    "control_variate_rho_plus_one": lambda r: (setattr(r, "payoffs", H.ASIAN_GEO_CV), setattr(r, "rho", 1.0)),
    "control_variate_rho_minus_one": lambda r: (setattr(r, "payoffs", H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN),
                                                setattr(r, "flags", H.CV_REGRESSION), setattr(r, "rho", -1.0)),
    # tile indices are 32-bit and a batch holds fewer than 2^31 work items: 2^31 - 1 tiles at most
    "paths_2_38": lambda r: setattr(r, "n_paths", 2 ** 38),
    "paths_one_tile_past_limit": lambda r: setattr(r, "n_paths", 2 ** 37 - 63),
}


def test_validation_accepts_good_request():
    assert H.validate(good()) == 0
    r = good()
    r.expiries, r.strikes = [0.3, 0.31, 1.0], [[90, 100], [], [100]]
    assert H.validate(r) == 0
    r.payoffs = H.ASIAN | H.EUROPEAN | H.ASIAN_GEO_CV | H.GEO_ASIAN
    r.flags = H.CV_REGRESSION | H.MARTINGALE | H.DISCOUNT | H.DRIFT | H.TRAPEZOID
    assert H.validate(r) == 0
    # tiny sigma: FP64 takes it; FP32 down to sigma^2 dt = 1e-18 max(v0, theta)
    r.sigma, r.precision = 1e-9, H.FP64
    assert H.validate(r) == 0
    r.sigma, r.precision = 1e-8, H.FP32
    assert H.validate(r) == 0
    # rho = +-1: plain Monte Carlo payoffs (the geometric one included) but no exact price
    r = good()
    r.payoffs, r.rho = H.ASIAN | H.EUROPEAN | H.GEO_ASIAN, 1.0
    assert H.validate(r) == 0
    r.payoffs, r.rho = H.ASIAN_GEO_CV, 0.999999
    assert H.validate(r) == 0
    # the largest path count: 2^31 - 1 tiles of 64 paths
    r = good()
    r.n_paths = 2 ** 37 - 64
    assert H.validate(r) == 0


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_validation_rejects_malformed(name):
    r = good()
    MUTATIONS[name](r)
    assert H.validate(r) == -1


def test_reserved_field_must_be_zero():
    lib = _lib.load()
    q = good().to_c()
    q.reserved = 1
    assert lib.hexo_validate_request(C.byref(q)) == -1


def test_n_outputs():
    lib = _lib.load()
    r = good()
    r.expiries, r.strikes, r.payoffs = [0.5, 1.0], [[90, 100, 110], [100]], H.ASIAN | H.EUROPEAN
    assert lib.hexo_request_n_outputs(C.byref(r.to_c())) == 8
    r.payoffs = H.EUROPEAN
    assert lib.hexo_request_n_outputs(C.byref(r.to_c())) == 4
    r.payoffs = H.ASIAN | H.EUROPEAN | H.ASIAN_GEO_CV
    assert lib.hexo_request_n_outputs(C.byref(r.to_c())) == 12
    r.steps = 0
    assert lib.hexo_request_n_outputs(C.byref(r.to_c())) == 0


def test_time_grid_rejects_bad_input():
    with pytest.raises(H.HexoError):
        H.time_grid([1.0, 0.5], [1, 1], 10)
    with pytest.raises(H.HexoError):
        H.time_grid([1.0], [0], 10)


def test_context_without_gpu_reports_no_device():
    import subprocess, sys
    # run in a child so that a GPU-less HIP initialisation cannot affect this process
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import hestonexotics_amd as H\n"
            "try:\n    H.Context(0); print('ok')\nexcept H.HexoError as e:\n    print(e.code)\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout.strip()
    assert out in ("ok", "-2"), out


def test_workload_generators_are_seeded():
    a = W.c3(n=16, seed=3, path_scale=0.01)
    b = W.c3(n=16, seed=3, path_scale=0.01)
    assert [(r.kappa, r.n_paths, r.seed) for r in a] == [(r.kappa, r.n_paths, r.seed) for r in b]
    reqs, t = W.c4(duration_s=2.0, seed=1)
    assert np.all(np.diff(t) > 0) and 300 < len(reqs) < 500
    assert W.path_steps(W.c2()) == 824_000_000


def test_swift_grid_matches_oracle_and_rejects_bad_chains():
    """The host-side SWIFT grid (SwiftParameters, src/SWIFT.cpp:21-35) equals the oracle's."""
    from oracle import oracle as O
    for p, S, r, T, K in [([0.04, 0.04, -0.7, 2.0, 0.3], 100.0, 0.02, 1.0, list(np.linspace(70, 130, 64))),
                          ([0.05, 0.04, -0.6, 1.5, 0.5], 100.0, 0.01, 0.75, [80.0, 100.0, 125.0])]:
        g = H.swift_grid(H.SwiftChain(*p, S0=S, T=T, strikes=K, r=r, trunc=1e-10))
        _, used = O.swift_price(p, S, r, T, K, trunc=1e-10)
        assert [g[i] for i in (0, 1, 5, 6, 7)] == [used[i] for i in (0, 1, 5, 6, 7)]
        assert np.allclose(g, used, rtol=1e-13, atol=0)
    ok = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, T=1.0, strikes=[100.0])
    for k, v in [("sigma", 0.0), ("rho", 1.0), ("T", 0.0), ("strikes", [0.0]), ("trunc", 0.0)]:
        with pytest.raises(H.HexoError):
            H.swift_grid(H.SwiftChain(**{**ok, k: v}))
    for grid in ([5, 32, 5.65, -1.0, 1.0, 10, 5, 128],      # k1 > k2
                 [5, 32, 5.65, -1.0, 1.0, -300, 300, 128],  # range beyond 4J
                 [5, 32, 5.65, -1.0, 1.0, -10, 10, 100],    # J not a power of two
                 [5, 16, 5.65, -1.0, 1.0, -10, 10, 128]):   # 2^m inconsistent
        with pytest.raises(H.HexoError):
            H.swift_grid(H.SwiftChain(**ok, grid=grid))


@pytest.mark.parametrize("T,strikes,N,flags", [
    ([1.0], [[100.0]], 12, 0),                                     # last step on T
    ([1.0], [[100.0]], 252, 0),                                    # one extra step
    ([0.5, 1.0, 2.0], [[100.0], [], [90.0, 110.0]], 26, 0),        # step-width changes, an empty chain
    ([0.3, 0.31, 0.7, 1.2], [[90, 100], [], [100, 105, 120], [100]], 7, H.TRAPEZOID),
    ([0.1], [[100.0]], 2, H.TRAPEZOID),
])
def test_asian_weights_match_oracle(T, strikes, N, flags):
    """The Asian functional's weights as the path kernels accumulate them (from the run schedule)
    equal the oracle's, taken from the reference's own time loop: the control variate applies
    them to ln X, and its exact price uses them."""
    from oracle import oracle as O
    req = H.Request(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, expiries=T, strikes=strikes,
                    steps=N, n_paths=64, flags=flags)
    for c in range(len(T)):
        t, w = H.asian_weights(req, c)
        to, wo = O.asian_weights(T, [len(k) for k in strikes], N, c, flags)
        np.testing.assert_array_equal(t, to)
        np.testing.assert_allclose(w, wo, rtol=0, atol=1e-15)


def test_asian_weights_refuse_before_allocating():
    """A schedule of 2^22 steps per expiry over 4096 expiries (about 2^34 grid steps) is refused from
    the schedule alone: the weight walks never build per-step arrays the caller has no room for, and
    control-variate validation refuses the oversized expiry."""
    import time
    T = list(np.linspace(0.5, 4096.0, 4096))
    req = H.Request(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, expiries=T,
                    strikes=[[100.0]] * len(T), steps=1 << 22, n_paths=64)
    lib = _lib.load()
    t = np.zeros(8)
    w = np.zeros(8)
    t0 = time.perf_counter()
    q = req.to_c()
    for chain in (0, 4095):
        n = lib.hexo_asian_weights(C.byref(q), chain, t.ctypes.data_as(C.POINTER(C.c_double)),
                                   w.ctypes.data_as(C.POINTER(C.c_double)), 8)
        assert n == -1
    req.payoffs |= H.ASIAN_GEO_CV
    assert H.validate(req) == -1
    assert time.perf_counter() - t0 < 30.0
    # the Python wrapper sizes its buffers from the schedule (one expiry of 2^22 steps: 32 MiB each)
    small = H.Request(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, expiries=[1.0],
                      strikes=[[100.0]], steps=1 << 22, n_paths=64)
    tt, ww = H.asian_weights(small, 0)
    assert len(tt) == len(ww) == (1 << 22) + 1
    # the grid reaches T = 1 exactly (2^22 steps of 2^-22): the reference's final term replaces the
    # last trapezoid, so the weights sum to 1 - 1/N
    assert abs(ww.sum() - (1.0 - 2.0 ** -22)) < 1e-9

