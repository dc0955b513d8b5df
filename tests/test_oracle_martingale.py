"""CPU oracle: Andersen's martingale correction of the QE log-price step (HEXO_MARTINGALE).

Pins (none of them restates the correction's formula):
 * one step: E[X'/X | V] = exp(r delta) exactly, computed by a product midpoint rule over the
   step's two uniforms through the oracle's own step function; the uncorrected step misses it by
   up to 8% on the same cases, so a dropped term, a wrong sign or the wrong A fails the check;
 * whole paths: the Monte-Carlo mean of X_T (European call at K = 0) equals S0 exp(r T) within
   the estimator's standard error with the correction, and is tens of standard errors away
   without it.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

C5 = [0.04, 0.04, -0.9, 0.5, 0.9]       # Feller badly violated, exponential branch dominates
C1 = [0.04, 0.04, -0.7, 2.0, 0.3]       # quadratic branch dominates
POS_RHO = [0.04, 0.04, 0.5, 2.0, 0.3]   # A = K2 + K4/2 > 0: the existence condition is active

STEP_CASES = [(p, d, V) for p in (C5, C1, POS_RHO) for d in (1.0, 0.25) for V in (0.001, 0.04, 0.2)]


def branch(params, delta, V):
    """1 quadratic, 0 exponential (the branch depends on V only, not on the draws)."""
    return O.qe_step(params, delta, V, 0.0, 1 << 31, 1 << 31)[2]


def test_step_cases_cover_both_branches():
    kinds = {branch(p, d, V) for p, d, V in STEP_CASES}
    assert kinds == {0, 1}


def test_one_step_conditional_mean_is_exact():
    r = 0.03
    worst_corr, worst_plain = 0.0, 0.0
    for p, d, V in STEP_CASES:
        target = math.exp(r * d)
        corr = O.qe_step_mean(p, d, V, r_drift=r, martingale=True, log2_nv=13, log2_nx=12) / target - 1.0
        plain = O.qe_step_mean(p, d, V, r_drift=r, martingale=False, log2_nv=13, log2_nx=12) / target - 1.0
        # the midpoint rule's own error at this resolution is a few 1e-6 (it halves when the
        # grid doubles); the positive-rho quadratic cases converge slowest
        assert abs(corr) < 3e-5, (p, d, V, corr)
        worst_corr = max(worst_corr, abs(corr))
        worst_plain = max(worst_plain, abs(plain))
    assert worst_plain > 1e-2 > 300 * worst_corr, (worst_plain, worst_corr)


def test_quadrature_error_shrinks_with_the_grid():
    p, d, V = C5, 1.0, 0.2
    e1 = abs(O.qe_step_mean(p, d, V, martingale=True, log2_nv=11, log2_nx=10) - 1.0)
    e2 = abs(O.qe_step_mean(p, d, V, martingale=True, log2_nv=13, log2_nx=12) - 1.0)
    assert e2 < 0.5 * e1


def test_correction_vanishes_without_correlation():
    """rho = 0: A = K2 + K4/2 = 0 and K1 + K3/2 = 0, so K0* = K0 and the corrected step is the
    plain one (which is then conditionally lognormal with the exact mean)."""
    p = [0.04, 0.04, 0.0, 2.0, 0.5]
    for wv, wx in [(123456789, 987654321), (1 << 31, 17), (4000000000, 3)]:
        a = O.qe_step(p, 1 / 52, 0.04, 0.0, wv, wx, r_drift=0.02)
        b = O.qe_step(p, 1 / 52, 0.04, 0.0, wv, wx, r_drift=0.02, martingale=True)
        assert a[0] == b[0] and abs(a[1] - b[1]) < 1e-12


@pytest.mark.parametrize("params,steps,T", [
    ([0.2] + C5[1:], 1, 1.0),
    ([0.2] + C5[1:], 4, 4.0),
    ([0.2] + POS_RHO[1:], 2, 2.0),
    ([0.01] + POS_RHO[1:], 3, 3.0),
])
def test_terminal_mean_is_forward(params, steps, T):
    S0, r, n = 100.0, 0.02, 2_000_000
    fwd = S0 * math.exp(r * T)
    out = {}
    for mart in (False, True):
        flags = O.DRIFT | (O.MARTINGALE if mart else 0)
        res = O.price(params, S0, [T], [1], [0.0], steps=steps, n_paths=n, seed=5, r=r, flags=flags,
                      payoffs=O.EUROPEAN)
        out[mart] = (res["prices"][0] - fwd) / res["std_errs"][0]
    assert abs(out[True]) < 4.0, out
    assert abs(out[False]) > 20.0, out
