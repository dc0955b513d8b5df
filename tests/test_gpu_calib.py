"""GPU calibration (hexo_calibrate) against the oracle's restatement of the reference's
calibration (oracle/calib.py) on the reference's own Levenberg-Marquardt test and from the
reference's implied-volatility starting point."""
import numpy as np
import pytest

import hestonexotics_amd as H

from test_oracle_calib import reference_levmar_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_reference_levmar_test_on_gpu():
    say("gpu calibration: reference levmar test")
    from oracle import calib as CB
    p_true, chains = reference_levmar_problem()
    with H.Context(0) as ctx:
        res = ctx.calibrate(1.0, chains, initial=p_true + 0.2, r=0.02)
    ref = CB.levenberg_marquardt(1.0, chains, p_true + 0.2, r=0.02)
    say("  gpu", res["params"], res["reason"], res["iterations"], "oracle", ref["params"], ref["reason"],
        ref["iterations"])
    assert res["reason"] in (2, 6)
    assert np.abs(res["params"] - p_true).sum() < 1e-2
    assert res["iterations"] == ref["iterations"] and res["reason"] == ref["reason"]
    np.testing.assert_allclose(res["params"], ref["params"], rtol=1e-6, atol=1e-9)


def test_calibration_from_implied_volatility_start():
    say("gpu calibration: reference starting point")
    from oracle import calib as CB
    p_true = np.array([0.04, 0.05, -0.6, 1.8, 0.5])
    S = 100.0
    chains = []
    for T in (0.25, 0.5, 1.0, 2.0):
        K = list(np.linspace(80.0, 120.0, 9))
        prices, _ = CB.O.swift_price(list(p_true), S, CB.YEARLY_RISK_FREE, T, K, trunc=1e-10)
        chains.append((T, K, prices, [1.0] * len(K)))
    with H.Context(0) as ctx:
        res = ctx.calibrate(S, chains)
        with pytest.raises(H.HexoError):
            ctx.calibrate(S, [(1.0, [100.0], [np.nan])])
        with pytest.raises(H.HexoError):
            ctx.calibrate(-1.0, chains)
    ref = CB.levenberg_marquardt(S, chains, CB.initial_point(S, chains))
    say("  gpu", res["params"], res["reason"], res["iterations"], "oracle", ref["params"], ref["reason"])
    np.testing.assert_allclose(res["initial"], CB.initial_point(S, chains), rtol=1e-12)
    assert res["reason"] in (2, 6)
    np.testing.assert_allclose(res["params"], ref["params"], rtol=1e-5, atol=1e-8)
    assert np.abs(res["params"] - p_true).max() < 1e-3
