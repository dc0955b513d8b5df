"""The drop-in boundary on the GPU: the `hexo -p asian all` CLI (reference grammar and output
columns, src/Main.cpp:75-96) and the header-only HSimulation::price<HQEAnderson<...>> wrapper it
is built on give the C ABI's prices for the same request."""
import os
import re
import subprocess

import numpy as np
import pytest

import hestonexotics_amd as H

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "hestonexotics_amd", "csrc", "hexo")
GOLD = os.path.join(ROOT, "tests", "golden")


def run_cli(*args, env_extra=None):
    env = dict(os.environ, HEXO_DATA_DIR=GOLD, **(env_extra or {}))
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120, env=env)


def test_cli_prices_equal_c_abi(ctx):
    print("gpu cli -p asian all DEMO", flush=True)
    out = run_cli("-p", "asian", "all", "DEMO", env_extra={"HEXO_PATHS": "20000", "HEXO_STEPS": "200"})
    assert out.returncode == 0, out.stderr
    lines = [l for l in out.stdout.splitlines() if l.startswith("S:")]
    got = [float(re.search(r"asian-option-price: ([-0-9.]+)", l).group(1)) for l in lines]
    # the chain file: expiries of 20 (dropped, T <= 0.1), 65 and 261 trading days
    req = H.Request(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, expiries=[65 / 261, 1.0],
                    strikes=[[90, 100, 110], [100, 120]], steps=200, n_paths=20000, seed=1, payoffs=H.ASIAN,
                    precision=H.FP64)
    p, _ = ctx.price(req)
    assert len(got) == 5
    np.testing.assert_allclose(got, np.round(p, 2), atol=0.0051)
    # volume filter: open interest >= 300 keeps 3 calls
    out = run_cli("-p", "asian", "all", "DEMO", "open_interest", "300",
                  env_extra={"HEXO_PATHS": "5000", "HEXO_STEPS": "50"})
    assert out.returncode == 0 and len([l for l in out.stdout.splitlines() if l.startswith("S:")]) == 3


@pytest.mark.parametrize("cv,mart,prec", [("geo", "0", "fp64"), ("regression", "1", "fp32"), ("geo-unit", "0", "fp64"),
                                         ("none", "1", "fp64")])
def test_cli_variance_reduction_equals_c_abi(ctx, cv, mart, prec):
    """The wrapper's variance-reduction policies and martingale-corrected scheme, selected in the CLI
    with HEXO_CONTROL_VARIATE and HEXO_MARTINGALE, give the C ABI's prices of the same request: the
    control-variate block (the regression coefficient for "geo" and "regression", coefficient 1 for
    "geo-unit") or the corrected Asian."""
    print("gpu cli variance reduction", cv, mart, prec, flush=True)
    out = run_cli("-p", "asian", "all", "DEMO",
                  env_extra={"HEXO_PATHS": "20000", "HEXO_STEPS": "200", "HEXO_CONTROL_VARIATE": cv,
                             "HEXO_MARTINGALE": mart, "HEXO_PRECISION": prec})
    assert out.returncode == 0, out.stderr
    got = [float(re.search(r"asian-option-price: ([-0-9.]+)", l).group(1))
           for l in out.stdout.splitlines() if l.startswith("S:")]
    payoffs, flags, block = {"geo": (H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN, H.CV_REGRESSION, 1),
                             "regression": (H.ASIAN | H.ASIAN_GEO_CV | H.GEO_ASIAN, H.CV_REGRESSION, 1),
                             "geo-unit": (H.ASIAN_GEO_CV, 0, 0),
                             "none": (H.ASIAN, 0, 0)}[cv]
    req = H.Request(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3, S0=100.0, expiries=[65 / 261, 1.0],
                    strikes=[[90, 100, 110], [100, 120]], steps=200, n_paths=20000, seed=1, payoffs=payoffs,
                    flags=flags | (H.MARTINGALE if mart == "1" else 0),
                    precision=H.FP32 if prec == "fp32" else H.FP64)
    p, _ = ctx.price(req)
    n = req.n_opts
    assert len(got) == n
    np.testing.assert_allclose(got, np.round(p[block * n:(block + 1) * n], 2), atol=0.0051)
    bad = run_cli("-p", "asian", "all", "DEMO", env_extra={"HEXO_CONTROL_VARIATE": "antithetic"})
    assert bad.returncode == 2


def test_cli_self_tests():
    """`hexo -t rng` is the reference's rng_test at its sample size (2^28 uniforms and normals);
    `hexo -t simulation` prices one Asian call with both kernels."""
    print("gpu cli -t rng -t simulation", flush=True)
    out = run_cli("-t", "rng", "-t", "simulation")
    print(out.stdout, flush=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "RNG test passed" in out.stdout and "Simulation test passed" in out.stdout


def test_cli_rejects_bad_grammar():
    assert run_cli("-p", "european", "all", "DEMO").returncode == 2
    assert run_cli("-q").returncode == 2


def test_cli_calibrates_and_stores_parameters(tmp_path):
    """`hexo -c <SYM>` (src/Main.cpp:64-72) on a chain file whose asks are SWIFT prices of known
    parameters: the stored parameters recover them; `hexo -p asian all` then uses them."""
    print("gpu cli -c SYNTH", flush=True)
    from oracle import oracle as O
    p_true = [0.04, 0.05, -0.6, 1.8, 0.5]
    lines = ["quote 100.0"]
    for days in (65, 130, 261, 522):
        T = days / 261.0
        K = [80.0, 85.0, 90.0, 95.0, 100.0, 105.0, 110.0, 115.0, 120.0]
        prices, _ = O.swift_price(p_true, 100.0, 0.0005, T, K, trunc=1e-10)
        lines.append(f"expiry {days}")
        lines += [f"call {k} {p:.12f} {p:.12f} 100 100" for k, p in zip(K, prices)]
    (tmp_path / "SYNTH.chains").write_text("\n".join(lines) + "\n")
    env = {"HEXO_DATA_DIR": str(tmp_path)}
    out = subprocess.run([CLI, "-c", "SYNTH"], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, **env))
    assert out.returncode == 0, out.stdout + out.stderr
    if (tmp_path / "HParams.db").exists():  # the reference's table (src/inc/DB.h:31), via libsqlite3
        import sqlite3
        con = sqlite3.connect(str(tmp_path / "HParams.db"))
        row = con.execute("SELECT V0, Vm, Rho, Kappa, Sigma FROM HestonParameters WHERE Name='SYNTH'").fetchone()
        con.close()
        stored = list(row)
    else:
        stored = [float(v) for v in (tmp_path / "SYNTH.params").read_text().split()]
    assert np.abs(np.array(stored) - np.array(p_true)).max() < 1e-3, stored
    out = subprocess.run([CLI, "-p", "asian", "all", "SYNTH"], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, HEXO_PATHS="2000", HEXO_STEPS="50", **env))
    assert out.returncode == 0, out.stderr
    assert len([l for l in out.stdout.splitlines() if l.startswith("S:")]) == 36
    assert "setup completed" not in out.stdout  # the stored parameters were used, no recalibration
