"""CPU oracle: the SWIFT European pricer (FP64 European oracle), pinned by the reference's own
golden vectors (src/UnitTest.cpp:221-261, tolerance sum |diff| < 1e-9 at :218,270) and by an
independent closed-form Heston integral (tests/golden/heston_closed_form.json)."""
import json
import os

import numpy as np

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_swift_reference_golden_prices():
    g = json.load(open(os.path.join(GOLD, "swift_reference_prices.json")))
    p = g["params"]
    params = [p["v0"], p["theta"], p["rho"], p["kappa"], p["sigma"]]
    got = []
    for T, K, sp in zip(g["expiries"], g["strikes"], g["swift_params"]):
        prices, used = O.swift_price(params, g["S"], g["r"], T, K, swift_params=sp)
        assert list(used) == [float(x) for x in sp]
        got.extend(prices)
    diff = np.abs(np.array(got) - np.array(g["prices"]))
    assert diff.sum() < 1e-9, diff.sum()


def test_swift_golden_case_against_closed_form():
    """The golden SWIFT prices (truncated at m = 5) agree with the Heston integral to ~2e-7."""
    g = json.load(open(os.path.join(GOLD, "swift_reference_prices.json")))
    cf = json.load(open(os.path.join(GOLD, "heston_closed_form.json")))
    for case in cf["cases"][:2]:
        i = int(np.argmin(np.abs(np.array(g["expiries"]) - case["T"])))
        ref = np.array(g["prices"][5 * i:5 * i + 5])
        assert np.abs(ref - np.array(case["price"])).max() < 5e-7


def test_swift_dynamic_parameters_against_closed_form():
    """Dynamically chosen SWIFT parameters (src/SWIFT.cpp:21-35) on the grid workload's model."""
    cf = json.load(open(os.path.join(GOLD, "heston_closed_form.json")))
    for case in cf["cases"][2:]:
        params = [case["v0"], case["theta"], case["rho"], case["kappa"], case["sigma"]]
        prices, used = O.swift_price(params, case["S"], case["r"], case["T"], case["K"], trunc=1e-10)
        ref = np.array(case["price"])
        assert np.abs(prices - ref).max() < 1e-5 * case["S"], (case["T"], prices - ref, used)


def test_chf_is_characteristic_function():
    p = [0.04, 0.04, -0.7, 2.0, 0.3]
    assert abs(O.heston_chf(p, 1.0, 0.0) - 1.0) < 1e-14
    # |phi(u)| <= 1 for real u
    for u in [0.3, 1.0, 5.0, 40.0]:
        assert abs(O.heston_chf(p, 1.0, u)) <= 1.0 + 1e-14


def test_swift_reference_golden_gradients():
    """Price gradients (SWIFT::price_opts_grad) against the reference's golden Jacobian
    (src/UnitTest.cpp:278-478, tolerance sum |diff| < 1e-9 at :275,493)."""
    g = json.load(open(os.path.join(GOLD, "swift_reference_prices.json")))
    p = g["params"]
    params = [p["v0"], p["theta"], p["rho"], p["kappa"], p["sigma"]]
    gold = np.array(g["gradients"]).reshape(-1, 5)[:, [4, 1, 3, 0, 2]]
    ours = np.vstack([O.swift_price(params, g["S"], g["r"], T, K, swift_params=sp, grads=True)[2]
                      for T, K, sp in zip(g["expiries"], g["strikes"], g["swift_params"])])
    assert np.abs(ours - gold).sum() < 1e-9


def test_swift_gradient_against_finite_differences():
    """With the SWIFT grid held fixed, the gradient equals central differences of the price."""
    params = np.array([0.05, 0.04, -0.6, 1.5, 0.5])
    K = [80.0, 100.0, 125.0]
    _, used = O.swift_price(params, 100.0, 0.01, 0.75, K, trunc=1e-10)
    _, _, g = O.swift_price(params, 100.0, 0.01, 0.75, K, swift_params=used, grads=True)
    for k in range(5):
        h = 1e-5 * max(abs(params[k]), 0.1)
        up, dn = params.copy(), params.copy()
        up[k] += h
        dn[k] -= h
        fd = (O.swift_price(up, 100.0, 0.01, 0.75, K, swift_params=used)[0] -
              O.swift_price(dn, 100.0, 0.01, 0.75, K, swift_params=used)[0]) / (2 * h)
        assert np.abs(g[:, k] - fd).max() < 1e-6 * (1 + np.abs(fd).max()), (k, g[:, k], fd)
