"""CPU oracle of the calibration (oracle/calib.py): the reference's own Levenberg-Marquardt test
(src/UnitTest.cpp:498-524: synthetic SWIFT prices at p, start p + 0.2, recover p to a summed
absolute error below 1e-2) and the Black-Scholes implied-volatility bisection against known
volatilities."""
import json
import os

import numpy as np

from oracle import calib as CB

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def reference_levmar_problem():
    g = json.load(open(os.path.join(GOLD, "swift_reference_prices.json")))
    p_true = np.array([0.09, 0.09, 0.04, 1.0, 1.0])  # src/UnitTest.cpp:182-188
    chains = []
    for T, K in zip(g["expiries"], g["strikes"]):
        prices, _ = CB.O.swift_price(list(p_true), 1.0, 0.02, T, K, trunc=1e-10)
        chains.append((T, K, prices, [1.0] * len(K)))
    return p_true, chains


def test_reference_levmar_test():
    p_true, chains = reference_levmar_problem()
    res = CB.levenberg_marquardt(1.0, chains, p_true + 0.2, r=0.02)
    assert res["reason"] in (2, 6)
    assert np.abs(res["params"] - p_true).sum() < 1e-2  # LEVMAR_TOLERANCE, src/UnitTest.cpp:497
    assert res["err"] < 1e-10 < res["err0"]


def test_implied_volatility_bisection():
    for S, K, T, sig in [(100.0, 90.0, 0.5, 0.25), (100.0, 120.0, 2.0, 0.6), (1.0, 1.0427, 0.119, 0.3)]:
        price = CB.bs_call(S, K, sig, T)
        assert abs(CB.imp_vol(S, K, price, T) - sig) < 2e-6
    assert CB.imp_vol(100.0, 50.0, 49.0, 1.0) == -1.0  # below the arbitrage bound
    assert CB.imp_vol(100.0, 50.0, 100.0, 1.0) == -2.0
