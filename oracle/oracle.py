"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle is a plain-C restatement of the reference's Heston QE Monte-Carlo path
(src/HSimulation.tpp:10-94, src/inc/AsianContract.h:25-47, src/inc/VanillaContract.h:28-39,
src/RNG.cpp:28-43, src/as241.f90:15-119) and of its SWIFT European pricer
(src/SWIFT.cpp:18-118, src/HDistribution.cpp:8-113). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, always as the checker or the timed CPU baseline; the product
in hestonexotics_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import platform

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ASIAN, EUROPEAN, ASIAN_GEO_CV, GEO_ASIAN = 1, 2, 4, 8
DRIFT, DISCOUNT, TRAPEZOID, MARTINGALE = 1, 2, 4, 8

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def _bind(lib: C.CDLL) -> C.CDLL:
    lib.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
    lib.orc_rng_words.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, _u32p]
    lib.orc_uniform_from_word.argtypes = [C.c_uint32]
    lib.orc_uniform_from_word.restype = C.c_double
    for f in (lib.orc_ppnd16_f64, lib.orc_ppnd16_asbuilt):
        f.argtypes = [C.c_double, C.POINTER(C.c_int)]
        f.restype = C.c_double
    lib.orc_normal_f64_from_word.argtypes = [C.c_uint32]
    lib.orc_normal_f64_from_word.restype = C.c_double
    lib.orc_normal_f32_from_word.argtypes = [C.c_uint32]
    lib.orc_normal_f32_from_word.restype = C.c_float
    lib.orc_as241_coefficients.argtypes = [_f64p]
    lib.orc_qe_step.argtypes = [_f64p, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.c_uint32, C.c_uint32]
    lib.orc_qe_step.restype = C.c_int
    lib.orc_qe_step_mean.argtypes = [_f64p, C.c_double, C.c_double, C.c_int, C.c_double, C.c_uint32, C.c_uint32]
    lib.orc_qe_step_mean.restype = C.c_double
    lib.orc_time_grid.argtypes = [_f64p, C.c_uint32, _u32p, C.c_uint32, _u64p, _f64p]
    lib.orc_time_grid.restype = C.c_int64
    lib.orc_price.argtypes = [_f64p, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                              C.c_uint64, C.c_uint32, C.c_uint32, _f64p, _u32p, _f64p, _f64p, _f64p, C.c_void_p,
                              C.c_void_p, C.c_int]
    lib.orc_price.restype = C.c_int
    lib.orc_asian_weights.argtypes = [_f64p, C.c_uint32, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _f64p, _f64p,
                                      C.c_uint64]
    lib.orc_asian_weights.restype = C.c_int64
    lib.orc_geo_cf_out.argtypes = [_f64p, C.c_double, _f64p, _f64p, C.c_int64, C.c_double, C.c_double, _f64p]
    lib.orc_geo_price_weights.argtypes = [_f64p, C.c_double, C.c_double, _f64p, _f64p, C.c_int64, C.c_uint32,
                                          _f64p, _f64p]
    lib.orc_geo_price_weights.restype = C.c_int
    lib.orc_geo_price_weights_tol.argtypes = [_f64p, C.c_double, C.c_double, _f64p, _f64p, C.c_int64, C.c_uint32,
                                              _f64p, C.c_double, _f64p]
    lib.orc_geo_price_weights_tol.restype = C.c_int
    lib.orc_geo_asian_price.argtypes = [_f64p, C.c_double, C.c_double, _f64p, C.c_uint32, _u32p, C.c_uint32,
                                        C.c_uint32, _f64p, _f64p]
    lib.orc_geo_asian_price.restype = C.c_int
    lib.orc_last_seconds.restype = C.c_double
    lib.orc_max_threads.restype = C.c_int
    return lib


_LIBS: dict[str, C.CDLL] = {}


def load(kind: str = "strict") -> C.CDLL:
    """kind: 'strict' (checker build), 'fma' (same source with FMA contraction) or 'fast'
    (reference numerics flags, for timing)."""
    if kind in _LIBS:
        return _LIBS[kind]
    if kind == "strict":
        name = "liboracle.so"
    elif kind == "fma":
        name = "liboracle_fma.so"
    else:
        flags = ""
        try:
            with open("/proc/cpuinfo") as fh:
                for line in fh:
                    if line.startswith("flags"):
                        flags = line
                        break
        except OSError:
            pass
        v4 = " avx512f " in flags and " avx512bw " in flags and platform.machine() == "x86_64"
        name = "liboracle_fast_v4.so" if v4 else "liboracle_fast_v3.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built; run __graft_entry__.build()")
    lib = _bind(C.CDLL(path))
    if kind in ("strict", "fma"):
        lib.orc_swift_price.argtypes = [_f64p, C.c_double, C.c_double, C.c_double, C.c_uint32, _f64p, C.c_void_p,
                                        C.c_double, _f64p, _f64p]
        lib.orc_swift_price.restype = C.c_int
        lib.orc_swift_price_grad.argtypes = [_f64p, C.c_double, C.c_double, C.c_double, C.c_uint32, _f64p, C.c_void_p,
                                             C.c_double, _f64p, C.c_void_p, _f64p]
        lib.orc_swift_price_grad.restype = C.c_int
        lib.orc_heston_chf.argtypes = [_f64p, C.c_double, C.c_double, _f64p]
    lib._hexo_name = name
    _LIBS[kind] = lib
    return lib


def philox(ctr, key) -> np.ndarray:
    out = np.zeros(4, np.uint32)
    load().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def rng_words(seed: int, path0: int, n_paths: int, word0: int, n_words: int, stream: int = 0) -> np.ndarray:
    out = np.zeros((n_paths, n_words), np.uint32)
    load().orc_rng_words(seed, stream, path0, n_paths, word0, n_words, out.reshape(-1))
    return out


def ppnd16_f64(p: float) -> tuple[float, int]:
    f = C.c_int(0)
    z = load().orc_ppnd16_f64(p, C.byref(f))
    return z, f.value


def ppnd16_asbuilt(p: float) -> tuple[float, int]:
    f = C.c_int(0)
    z = load().orc_ppnd16_asbuilt(p, C.byref(f))
    return z, f.value


def normals_f64(words) -> np.ndarray:
    lib = load()
    return np.array([lib.orc_normal_f64_from_word(int(w)) for w in np.asarray(words).reshape(-1)], np.float64)


def normals_f32(words) -> np.ndarray:
    lib = load()
    return np.array([lib.orc_normal_f32_from_word(int(w)) for w in np.asarray(words).reshape(-1)], np.float32)


def as241_coefficients() -> np.ndarray:
    out = np.zeros(45, np.float64)
    load().orc_as241_coefficients(out)
    return out


def qe_step(params, delta: float, V: float, logX: float, w_v: int, w_x: int, r_drift: float = 0.0,
            martingale: bool = False):
    v = C.c_double(V)
    lx = C.c_double(logX)
    quad = load().orc_qe_step(np.asarray(params, np.float64), r_drift, delta, int(martingale), C.byref(v),
                              C.byref(lx), w_v, w_x)
    return v.value, lx.value, quad


def qe_step_mean(params, delta: float, V: float, r_drift: float = 0.0, martingale: bool = False,
                 log2_nv: int = 12, log2_nx: int = 12) -> float:
    """E[X'/X | V] of one QE step by a midpoint rule over its two uniforms."""
    return load().orc_qe_step_mean(np.asarray(params, np.float64), r_drift, delta, int(martingale), V, log2_nv,
                                   log2_nx)


def time_grid(T, n_strikes, steps: int):
    T = np.ascontiguousarray(T, np.float64)
    ns = np.ascontiguousarray(n_strikes, np.uint32)
    fs = np.zeros(len(T), np.uint64)
    fw = np.zeros(len(T), np.float64)
    k = load().orc_time_grid(T, len(T), ns, steps, fs, fw)
    if k < 0:
        raise ValueError("bad time grid input")
    return int(k), fs, fw


def price(params, S0: float, T, n_strikes, strikes, *, steps: int, n_paths: int, seed: int, r: float = 0.0,
          flags: int = 0, payoffs: int = ASIAN | EUROPEAN, path_offset: int = 0, want_finals: bool = False,
          want_counts: bool = False, n_threads: int = 0, kind: str = "strict"):
    """Reference-semantics FP64 Monte-Carlo price. Returns dict with prices, std_errs (payoff-major,
    Asian block first), optional finals [n_paths, n_exp, 2] and branch counts."""
    lib = load(kind)
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    ns = np.ascontiguousarray(np.atleast_1d(n_strikes), np.uint32)
    K = np.ascontiguousarray(np.atleast_1d(strikes), np.float64)
    n_pay = bin(payoffs & 15).count("1")
    n_out = int(ns.sum()) * n_pay
    prices = np.zeros(n_out, np.float64)
    ses = np.zeros(n_out, np.float64)
    finals = np.zeros((n_paths, len(T), 2), np.float64) if want_finals else None
    counts = np.zeros(2, np.uint64) if want_counts else None
    rc = lib.orc_price(np.asarray(params, np.float64), S0, r, flags, payoffs, n_paths, path_offset, seed, steps,
                       len(T), T, ns, K, prices, ses,
                       finals.ctypes.data if finals is not None else None,
                       counts.ctypes.data if counts is not None else None, n_threads)
    if rc != 0:
        raise ValueError(f"orc_price failed: {rc}")
    out = {"prices": prices, "std_errs": ses, "seconds": lib.orc_last_seconds()}
    if finals is not None:
        out["finals"] = finals
    if counts is not None:
        out["counts"] = counts
    return out


def asian_weights(T, n_strikes, steps: int, chain: int, flags: int = 0):
    """(t[0..n], w[0..n]): grid times and weights of chain `chain`'s Asian functional
    A = sum_j w_j X_{t_j} under the reference's time loop."""
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    ns = np.ascontiguousarray(np.atleast_1d(n_strikes), np.uint32)
    cap = 4 * steps * len(T) + 64
    t = np.zeros(cap, np.float64)
    w = np.zeros(cap, np.float64)
    n = load().orc_asian_weights(T, len(T), ns, steps, flags, chain, t, w, cap)
    if n < 0:
        raise ValueError(f"orc_asian_weights failed: {n}")
    return t[:n + 1].copy(), w[:n + 1].copy()


def geo_cf(params, r: float, t, v, s: complex) -> complex:
    """E[exp(s sum_j v_j ln(X_{t_j}/S0))] under Heston (t[0] = 0; v[0] is ignored)."""
    t = np.ascontiguousarray(t, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(2, np.float64)
    load().orc_geo_cf_out(np.asarray(params, np.float64), r, t, v, len(t) - 1, s.real, s.imag, out)
    return complex(out[0], out[1])


def geo_price_weights(params, S0: float, r: float, t, w, strikes, tol: float = 1e-15) -> np.ndarray:
    """E[(W S0 exp(sum_j (w_j/W) ln(X_{t_j}/S0)) - K)^+], W = sum_j w_j, undiscounted; the Fourier
    sum stops at two consecutive panels whose |phi| is below tol u0 (NaN if it never does)."""
    t = np.ascontiguousarray(t, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    K = np.ascontiguousarray(np.atleast_1d(strikes), np.float64)
    out = np.zeros(len(K), np.float64)
    rc = load().orc_geo_price_weights_tol(np.asarray(params, np.float64), S0, r, t, w, len(t) - 1, len(K), K,
                                          float(tol), out)
    if rc not in (0, -4):
        raise ValueError(f"orc_geo_price_weights failed: {rc}")
    return out


def geo_asian_price(params, S0: float, T, n_strikes, strikes, *, steps: int, r: float = 0.0,
                    flags: int = 0) -> np.ndarray:
    """Exact price of the geometric counterpart of every option's Asian functional (undiscounted;
    r is the drift, as with the DRIFT flag)."""
    T = np.ascontiguousarray(np.atleast_1d(T), np.float64)
    ns = np.ascontiguousarray(np.atleast_1d(n_strikes), np.uint32)
    K = np.ascontiguousarray(np.atleast_1d(strikes), np.float64)
    out = np.zeros(len(K), np.float64)
    rc = load().orc_geo_asian_price(np.asarray(params, np.float64), S0, r, T, len(T), ns, steps, flags, K, out)
    if rc != 0:
        raise ValueError(f"orc_geo_asian_price failed: {rc}")
    return out


def swift_price(params, S: float, r: float, T: float, strikes, swift_params=None, trunc: float = 1e-7,
                grads: bool = False):
    """SWIFT European call prices (and, with grads=True, their gradients, shape (n, 5) in the
    order v0, theta, rho, kappa, sigma); returns (prices, used SWIFT parameters[, grads])."""
    lib = load()
    K = np.ascontiguousarray(np.atleast_1d(strikes), np.float64)
    prices = np.zeros(len(K), np.float64)
    used = np.zeros(8, np.float64)
    g = np.zeros((len(K), 5), np.float64) if grads else None
    sp = None
    if swift_params is not None:
        sp_arr = np.ascontiguousarray(swift_params, np.float64)
        sp = sp_arr.ctypes.data
    rc = lib.orc_swift_price_grad(np.asarray(params, np.float64), S, r, T, len(K), K, sp, trunc, prices,
                                  g.ctypes.data if grads else None, used)
    if rc != 0:
        raise ValueError(f"orc_swift_price_grad failed: {rc}")
    return (prices, used, g) if grads else (prices, used)


def heston_chf(params, tau: float, u: float) -> complex:
    out = np.zeros(2, np.float64)
    load().orc_heston_chf(np.asarray(params, np.float64), tau, u, out)
    return complex(out[0], out[1])


def qe_moments(params, delta: float, V: float, n: int, seed: int = 1) -> dict:
    lib = load()
    if not hasattr(lib, "_qe_bound"):
        lib.orc_qe_moments.argtypes = [_f64p, C.c_double, C.c_double, C.c_uint64, C.c_uint64, _f64p]
        lib._qe_bound = True
    out = np.zeros(6, np.float64)
    lib.orc_qe_moments(np.asarray(params, np.float64), delta, V, n, seed, out)
    return dict(zip(["mean", "var", "m4", "zero_frac", "quad_frac", "mean_exp_dlogx"], out.tolist()))
