"""Python mirror of the pricing interface, over the C ABI (include/hexo.h).

`Request` carries what one call of the reference's
`HSimulation::price<HQEAnderson<ffloat, Policy>>(p, S, chains, n_simulations, n_opts, steps)`
(src/inc/HSimulation.h:61-63) needs: Heston parameters in HParams order, spot, the expiry chains
with their strikes, path count, steps per expiry, plus the RNG seed, the payoffs to price from the
same paths, the precision and the documented extensions (drift, discount, corrected trapezoid).
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import (ASIAN, ASIAN_GEO_CV, CALIB_INITIAL_GIVEN, CV_REGRESSION, DISCOUNT, DRIFT, EUROPEAN, FP32, FP32_COMPENSATED,
                   FP64, GEO_ASIAN, MARTINGALE, SWIFT_FIXED_GRID, TRAPEZOID, HexoError, hexo_batch_stats, hexo_calib_chain,
                   hexo_calib_config, hexo_calib_result, hexo_request, hexo_server_config, hexo_server_stats,
                   hexo_swift_grid, hexo_swift_request, hexo_ticket_info, hexo_tuning)

__all__ = ["Request", "Context", "Batch", "ASIAN", "EUROPEAN", "ASIAN_GEO_CV", "GEO_ASIAN", "DRIFT", "DISCOUNT",
           "TRAPEZOID", "MARTINGALE", "CV_REGRESSION", "FP64", "FP32", "FP32_COMPENSATED", "HexoError", "validate", "time_grid", "asian_weights",
           "SwiftChain", "SwiftCall", "swift_grid", "cv_optimal"]

GRID_FIELDS = ("m", "exp2_m", "sqrt_exp2_m", "lower", "upper", "k1", "k2", "J")


@dataclass
class SwiftChain:
    """One expiry chain for the SWIFT European pricer (hexo_swift_request). grid: None (chosen from
    the model and the strikes, src/SWIFT.cpp:21-35) or the 8 SwiftParameters values
    (m, 2^m, sqrt(2^m), lower, upper, k1, k2, J) to use as given."""
    v0: float
    theta: float
    rho: float
    kappa: float
    sigma: float
    S0: float
    T: float
    strikes: Sequence[float]
    r: float = 0.0
    trunc: float = 1e-10
    grid: Sequence[float] | None = None

    def to_c(self):
        q = hexo_swift_request()
        for f in ("v0", "theta", "rho", "kappa", "sigma", "S0", "r", "T", "trunc"):
            setattr(q, f, float(getattr(self, f)))
        self._k = np.ascontiguousarray(self.strikes, np.float64)  # kept alive with the struct's owner
        q.n_strikes = len(self._k)
        q.strikes = self._k.ctypes.data_as(C.POINTER(C.c_double))
        if self.grid is not None:
            q.flags = SWIFT_FIXED_GRID
            for f, v in zip(GRID_FIELDS, self.grid):
                setattr(q.grid, f, type(getattr(q.grid, f))(v))
        return q


def _grid_tuple(g: hexo_swift_grid):
    return tuple(getattr(g, f) for f in GRID_FIELDS)


def swift_grid(chain: SwiftChain):
    """The SWIFT grid the pricer chooses for a chain (host only)."""
    g = hexo_swift_grid()
    q = chain.to_c()
    rc = _lib.load().hexo_swift_grid_for(C.byref(q), C.byref(g))
    if rc != 0:
        raise HexoError(rc, "hexo_swift_grid_for")
    return _grid_tuple(g)


@dataclass
class Request:
    v0: float
    theta: float
    rho: float
    kappa: float
    sigma: float
    S0: float
    expiries: Sequence[float]
    strikes: Sequence[Sequence[float]]  # one list of strikes per expiry
    steps: int
    n_paths: int
    seed: int = 1
    r: float = 0.0
    payoffs: int = ASIAN | EUROPEAN
    flags: int = 0
    precision: int = FP32
    path_offset: int = 0
    _keep: list = field(default_factory=list, repr=False, compare=False)

    @property
    def params(self) -> list[float]:
        return [self.v0, self.theta, self.rho, self.kappa, self.sigma]

    @property
    def n_opts(self) -> int:
        return int(sum(len(k) for k in self.strikes))

    @property
    def n_payoffs(self) -> int:
        return bin(self.payoffs & (ASIAN | EUROPEAN | ASIAN_GEO_CV | GEO_ASIAN)).count("1")

    @property
    def n_outputs(self) -> int:
        return self.n_opts * self.n_payoffs

    def flat_strikes(self) -> np.ndarray:
        return np.ascontiguousarray(np.concatenate([np.asarray(k, np.float64) for k in self.strikes])
                                    if self.strikes else np.zeros(0), np.float64)

    def n_strikes(self) -> np.ndarray:
        return np.ascontiguousarray([len(k) for k in self.strikes], np.uint32)

    def to_c(self) -> hexo_request:
        T = np.ascontiguousarray(self.expiries, np.float64)
        ns = self.n_strikes()
        K = self.flat_strikes()
        if K.size == 0:
            K = np.zeros(1, np.float64)
        self._keep = [T, ns, K]
        q = hexo_request()
        q.v0, q.theta, q.rho, q.kappa, q.sigma = self.params
        q.S0, q.r = self.S0, self.r
        q.n_paths, q.seed, q.path_offset = int(self.n_paths), int(self.seed) & (2**64 - 1), int(self.path_offset)
        q.steps, q.n_expiries = int(self.steps), len(T)
        q.expiries = T.ctypes.data_as(C.POINTER(C.c_double))
        q.n_strikes = ns.ctypes.data_as(C.POINTER(C.c_uint32))
        q.strikes = K.ctypes.data_as(C.POINTER(C.c_double))
        q.payoffs, q.flags, q.precision, q.reserved = self.payoffs, self.flags, self.precision, 0
        return q

    def split(self, prices: np.ndarray) -> dict:
        """Outputs are payoff-major: the Asian, European, control-variate Asian and geometric
        blocks, those priced, in that order."""
        out, o = {}, 0
        n = self.n_opts
        for bit, name in ((ASIAN, "asian"), (EUROPEAN, "european"), (ASIAN_GEO_CV, "asian_geo_cv"),
                          (GEO_ASIAN, "geo_asian")):
            if self.payoffs & bit:
                out[name] = prices[o:o + n]
                o += n
        return out


def _c_array(reqs: Sequence[Request]):
    arr = (hexo_request * len(reqs))()
    for i, r in enumerate(reqs):
        arr[i] = r.to_c()
    return arr


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def validate(req: Request) -> int:
    lib = _lib.load()
    return lib.hexo_validate_request(C.byref(req.to_c()))


def cv_optimal(asian, geo, cv, geo_exact):
    """The control variate with the regression coefficient beta = Cov(A, G) / Var(G) instead of 1.

    asian, geo, cv: (prices, std_errs) of the ASIAN, GEO_ASIAN and ASIAN_GEO_CV blocks of one
    request (same paths, same discounting); geo_exact: its exact geometric prices
    (Context.geo_asian_price). Var(A - G) = Var A + Var G - 2 Cov(A, G) gives the covariance; the
    estimator A - beta (G - E[G]) has variance Var A (1 - corr^2), never above the plain one.
    Returns (prices, std_errs, beta)."""
    (pa, sa), (pg, sg), (_, sd) = [(np.asarray(p, np.float64), np.asarray(s, np.float64))
                                   for p, s in (asian, geo, cv)]
    cov = 0.5 * (sa * sa + sg * sg - sd * sd)
    beta = np.where(sg > 0.0, cov / np.maximum(sg * sg, 1e-300), 0.0)
    price = pa - beta * (pg - np.asarray(geo_exact, np.float64))
    se = np.sqrt(np.maximum(sa * sa - beta * cov, 0.0))
    return price, se, beta


def asian_weights(req: Request, chain: int):
    """(t, w): grid times and weights of expiry `chain`'s Asian average A = sum_j w_j X(t_j), as
    the path kernels accumulate it (host only)."""
    lib = _lib.load()
    # sized from the run schedule: expiry `chain` is finalised after fin_step[chain] grid steps
    _, fin_step, _ = time_grid(req.expiries, req.n_strikes(), req.steps)
    if not 0 <= chain < len(fin_step) or fin_step[chain] == 0:
        raise HexoError(-1, f"asian weights: expiry {chain} is not finalised")
    cap = int(fin_step[chain]) + 1
    t = np.zeros(cap, np.float64)
    w = np.zeros(cap, np.float64)
    n = lib.hexo_asian_weights(C.byref(req.to_c()), chain, _dp(t), _dp(w), cap)
    if n < 0:
        raise HexoError(int(n), "asian weights")
    return t[:n + 1].copy(), w[:n + 1].copy()


def time_grid(expiries, n_strikes, steps: int):
    lib = _lib.load()
    T = np.ascontiguousarray(expiries, np.float64)
    ns = np.ascontiguousarray(n_strikes, np.uint32)
    fs = np.zeros(len(T), np.uint64)
    fw = np.zeros(len(T), np.float64)
    tot = C.c_uint64(0)
    rc = lib.hexo_time_grid(_dp(T), len(T), ns.ctypes.data_as(C.POINTER(C.c_uint32)), steps,
                            fs.ctypes.data_as(C.POINTER(C.c_uint64)), _dp(fw), C.byref(tot))
    if rc != 0:
        raise HexoError(rc, "time grid")
    return int(tot.value), fs, fw


class Batch:
    """A set of requests resident on the GPU; its launch sequence is a captured HIP graph."""

    def __init__(self, ctx: "Context", reqs: Sequence[Request]):
        self.ctx = ctx
        self.reqs = list(reqs)
        self._arr = _c_array(self.reqs)
        self.n_out = sum(r.n_outputs for r in self.reqs)
        self._h = C.c_void_p()
        ctx._check(ctx.lib.hexo_batch_create(ctx.h, self._arr, len(self.reqs), C.byref(self._h)))

    def run(self):
        self.ctx._check(self.ctx.lib.hexo_batch_run(self._h))

    def run_eager(self):
        self.ctx._check(self.ctx.lib.hexo_batch_run_eager(self._h))

    def sync(self):
        self.ctx._check(self.ctx.lib.hexo_batch_sync(self._h))

    def results(self):
        p = np.zeros(self.n_out, np.float64)
        s = np.zeros(self.n_out, np.float64)
        self.ctx._check(self.ctx.lib.hexo_batch_results(self._h, _dp(p), _dp(s)))
        return p, s

    def per_request(self, arr: np.ndarray) -> list[np.ndarray]:
        out, o = [], 0
        for r in self.reqs:
            out.append(arr[o:o + r.n_outputs])
            o += r.n_outputs
        return out

    def stats(self) -> dict:
        st = hexo_batch_stats()
        self.ctx._check(self.ctx.lib.hexo_batch_stats_get(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def branch_counts(self) -> np.ndarray:
        c = np.zeros(4, np.uint64)
        self.ctx._check(self.ctx.lib.hexo_batch_branch_counts(self._h, c.ctypes.data_as(C.POINTER(C.c_uint64))))
        return c

    def close(self):
        if self._h:
            self.ctx.lib.hexo_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One GPU. Wraps hexo_ctx_create / hexo_ctx_destroy."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.h = C.c_void_p()
        self._batches = weakref.WeakSet()
        rc = self.lib.hexo_ctx_create(device, C.byref(self.h))
        if rc != 0:
            raise HexoError(rc, "hexo_ctx_create")

    def _check(self, rc: int):
        if rc != 0:
            raise HexoError(rc, self.lib.hexo_ctx_last_error(self.h).decode())

    def close(self):
        if self.h:
            for b in list(self._batches):
                b.close()
            self.lib.hexo_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def device_info(self) -> dict:
        name = C.create_string_buffer(128)
        cu, clk = C.c_int(0), C.c_int(0)
        self._check(self.lib.hexo_device_info(self.h, name, 128, C.byref(cu), C.byref(clk)))
        return {"arch": name.value.decode(), "n_cu": cu.value, "clock_khz": clk.value}

    # ---- pricing ----
    def price(self, req: Request):
        p = np.zeros(req.n_outputs, np.float64)
        s = np.zeros(req.n_outputs, np.float64)
        self._check(self.lib.hexo_price(self.h, C.byref(req.to_c()), _dp(p), _dp(s)))
        return p, s

    def price_batch(self, reqs: Sequence[Request]):
        arr = _c_array(reqs)
        n = sum(r.n_outputs for r in reqs)
        p = np.zeros(n, np.float64)
        s = np.zeros(n, np.float64)
        self._check(self.lib.hexo_price_batch(self.h, arr, len(reqs), _dp(p), _dp(s)))
        out, o = [], 0
        for r in reqs:
            out.append((p[o:o + r.n_outputs], s[o:o + r.n_outputs]))
            o += r.n_outputs
        return out

    def batch(self, reqs: Sequence[Request]) -> Batch:
        b = Batch(self, reqs)
        self._batches.add(b)
        return b

    def paths(self, req: Request, first: int, n: int) -> np.ndarray:
        out = np.zeros((n, len(req.expiries), 2), np.float64)
        self._check(self.lib.hexo_paths(self.h, C.byref(req.to_c()), first, n, _dp(out.reshape(-1))))
        return out

    def geo_asian_price(self, req: Request) -> np.ndarray:
        """Exact Heston price of each option's geometric control variate (chain-major)."""
        out = np.zeros(req.n_opts, np.float64)
        self._check(self.lib.hexo_geo_asian_price(self.h, C.byref(req.to_c()), _dp(out)))
        return out

    # ---- components ----
    def rng_fill(self, seed: int, path0: int, n_paths: int, word0: int, n_words: int, stream: int = 0):
        out = np.zeros((n_paths, n_words), np.uint32)
        self._check(self.lib.hexo_rng_fill(self.h, seed, stream, path0, n_paths, word0, n_words,
                                           out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def philox(self, ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
        ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
        key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
        out = np.zeros_like(ctr)
        u = C.POINTER(C.c_uint32)
        self._check(self.lib.hexo_philox4x32_10(self.h, ctr.ctypes.data_as(u), key.ctypes.data_as(u), len(ctr),
                                                out.ctypes.data_as(u)))
        return out

    def normals(self, words: np.ndarray, precision: int) -> np.ndarray:
        w = np.ascontiguousarray(words, np.uint32).reshape(-1)
        out = np.zeros(w.size, np.float64)
        self._check(self.lib.hexo_normals(self.h, w.ctypes.data_as(C.POINTER(C.c_uint32)), w.size, precision,
                                          _dp(out)))
        return out

    # ---- SWIFT European pricer ----
    def swift_price(self, chains: Sequence[SwiftChain], grads: bool = False):
        """European call prices of each chain (and gradients, shape (n_strikes, 5) in the order
        v0, theta, rho, kappa, sigma). Returns a list of (prices, grads or None, grid) per chain."""
        return SwiftCall(self, chains, grads).run().split()


    # ---- calibration ----
    def calibrate(self, S0: float, chains, initial=None, **config) -> dict:
        """Heston calibration to market prices (hexo_calibrate). chains: sequence of
        (T, strikes, prices[, volumes]). initial: None (the reference's starting point from the
        average implied volatility) or (v0, theta, rho, kappa, sigma). config: max_iter,
        mu_scale, eps_grad, eps_step, eps_err, r, trunc (defaults: the reference's)."""
        arr = (hexo_calib_chain * len(chains))()
        keep = []
        for i, ch in enumerate(chains):
            T, K, P = ch[0], ch[1], ch[2]
            V = ch[3] if len(ch) > 3 else None
            k = np.ascontiguousarray(K, np.float64)
            p = np.ascontiguousarray(P, np.float64)
            keep += [k, p]
            arr[i].T = float(T)
            arr[i].n_options = len(k)
            arr[i].strikes = _dp(k)
            arr[i].prices = _dp(p)
            if V is not None:
                v = np.ascontiguousarray(V, np.float64)
                keep.append(v)
                arr[i].volumes = _dp(v)
        cfg = hexo_calib_config()
        self.lib.hexo_calib_defaults(C.byref(cfg))
        for key, val in config.items():
            setattr(cfg, key, val)
        if initial is not None:
            cfg.flags |= CALIB_INITIAL_GIVEN
            for j in range(5):
                cfg.initial[j] = float(initial[j])
        res = hexo_calib_result()
        self._check(self.lib.hexo_calibrate(self.h, float(S0), arr, len(chains), C.byref(cfg), C.byref(res)))
        return {"params": np.array(res.params[:]), "initial": np.array(res.initial[:]), "err0": res.err0,
                "err": res.err, "iterations": res.iterations, "reason": res.reason,
                "nfev": res.n_price_evals, "njev": res.n_jacobian_evals}

    # ---- tuning / measurement ----
    def get_tuning(self) -> dict:
        t = hexo_tuning()
        self._check(self.lib.hexo_get_tuning(self.h, C.byref(t)))
        return {"paths_per_thread": t.paths_per_thread, "block_threads": t.block_threads,
                "waves_per_simd": t.waves_per_simd}

    def set_tuning(self, paths_per_thread: int, block_threads: int, waves_per_simd: int = 0):
        t = hexo_tuning(paths_per_thread, block_threads, waves_per_simd, 0)
        self._check(self.lib.hexo_set_tuning(self.h, C.byref(t)))

    def autotune(self, reqs: Sequence[Request], reps: int = 2) -> dict:
        arr = _c_array(reqs)
        t = hexo_tuning()
        self._check(self.lib.hexo_autotune(self.h, arr, len(reqs), reps, C.byref(t)))
        return {"paths_per_thread": t.paths_per_thread, "block_threads": t.block_threads,
                "waves_per_simd": t.waves_per_simd}

    def copy_probe(self, nbytes: int, reps: int = 10) -> float:
        ms = C.c_double(0)
        self._check(self.lib.hexo_copy_probe(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    # ---- serving ----
    def server_start(self, memory_budget_bytes: int = 0, streams: int = 2, max_batch: int = 0):
        cfg = hexo_server_config(memory_budget_bytes, streams, max_batch, 0.0)
        self._check(self.lib.hexo_server_start(self.h, C.byref(cfg)))

    def server_stop(self):
        self._check(self.lib.hexo_server_stop(self.h))

    def submit(self, req: Request) -> int:
        t = C.c_uint64(0)
        rc = self.lib.hexo_submit(self.h, C.byref(req.to_c()), C.byref(t))
        if rc != 0:
            raise HexoError(rc, self.lib.hexo_ctx_last_error(self.h).decode())
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int(0)
        self._check(self.lib.hexo_poll(self.h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int, n_out: int, timeout_s: float = -1.0):
        p = np.zeros(n_out, np.float64)
        s = np.zeros(n_out, np.float64)
        info = hexo_ticket_info()
        rc = self.lib.hexo_wait(self.h, ticket, timeout_s, _dp(p), _dp(s), C.byref(info))
        if rc != 0:
            # refusals of the call itself carry a message; a ticket's own failure status does not
            msg = self.lib.hexo_ctx_last_error(self.h).decode() if rc in (-1, -9) else "wait"
            raise HexoError(rc, msg)
        return p, s, {k: getattr(info, k) for k, _ in info._fields_}

    def cancel(self, ticket: int):
        self._check(self.lib.hexo_cancel(self.h, ticket))

    def server_stats(self) -> dict:
        st = hexo_server_stats()
        self._check(self.lib.hexo_server_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def inject_fault(self, kind: int, count: int):
        self._check(self.lib.hexo_inject_fault(self.h, kind, count))

    def run_trace(self, reqs: Sequence[Request], arrival_s: Sequence[float]):
        arr = _c_array(reqs)
        n = sum(r.n_outputs for r in reqs)
        p = np.zeros(max(n, 1), np.float64)
        s = np.zeros(max(n, 1), np.float64)
        info = (hexo_ticket_info * max(len(reqs), 1))()
        at = np.ascontiguousarray(arrival_s, np.float64)
        rc = self.lib.hexo_server_run_trace(self.h, arr, _dp(at), len(reqs), _dp(p), _dp(s), info)
        infos = [{k: getattr(info[i], k) for k, _ in info[i]._fields_} for i in range(len(reqs))]
        return rc, p[:n], s[:n], infos


class SwiftCall:
    """One hexo_swift_price call with its C request array and output buffers built once, so that
    repeated calls (a calibration loop, a benchmark) go straight to the C ABI."""

    def __init__(self, ctx: Context, chains: Sequence[SwiftChain], grads: bool = False):
        self.ctx, self.chains, self.grads = ctx, list(chains), grads
        self.arr = (hexo_swift_request * len(self.chains))()
        for i, ch in enumerate(self.chains):
            self.arr[i] = ch.to_c()
        n = sum(len(ch.strikes) for ch in self.chains)
        self.prices = np.zeros(n, np.float64)
        self.grad_buf = np.zeros(5 * n, np.float64) if grads else None
        self.used = (hexo_swift_grid * len(self.chains))()
        self._pp = _dp(self.prices)
        self._gp = _dp(self.grad_buf) if grads else None

    def run(self) -> "SwiftCall":
        c = self.ctx
        c._check(c.lib.hexo_swift_price(c.h, self.arr, len(self.chains), self._pp, self._gp, self.used))
        return self

    def split(self):
        out, o = [], 0
        for i, ch in enumerate(self.chains):
            k = len(ch.strikes)
            g = self.grad_buf[5 * o:5 * (o + k)].reshape(k, 5).copy() if self.grads else None
            out.append((self.prices[o:o + k].copy(), g, _grid_tuple(self.used[i])))
            o += k
        return out
