#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden from independent sources (run in the build container).

philox_rocrand.json   Philox4x32-10 blocks computed by rocRAND's philox4x32_10_engine (ROCm's
                      open-source generator, compiled here for the host) on random counters/keys.
ndtri.json            scipy.special.ndtri on a grid of probabilities from 1e-300 to 1 - 1e-16
                      (a library inverse normal CDF, not AS241).
heston_closed_form.json  Heston European call prices from Lewis' single-integral formula with the
                      "little trap" characteristic function, integrated with scipy.integrate.quad
                      (textbook closed form, independent of SWIFT and of the Monte-Carlo path).
geometric_asian_gbm.json  Black-Scholes closed forms used by the sigma -> 0 limit tests.

The reference's own golden vectors (the 40 SWIFT prices of src/UnitTest.cpp:221-261 with their
inputs, src/UnitTest.cpp:155-188, and the 200 SWIFT price gradients of :278-477 with their index
map :478) are data copied into swift_reference_prices.json.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
from scipy import integrate, special

HERE = os.path.dirname(os.path.abspath(__file__))

ROCRAND_SRC = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdlib>
struct Exposed : rocrand_device::philox4x32_10_engine {
    using rocrand_device::philox4x32_10_engine::ten_rounds;
};
int main(int argc, char** argv) {
    Exposed e;
    unsigned v[6];
    while (scanf("%x %x %x %x %x %x", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]) == 6) {
        uint4 c = {v[0], v[1], v[2], v[3]};
        uint2 k = {v[4], v[5]};
        uint4 r = e.ten_rounds(c, k);
        printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
"""


def philox_rocrand():
    rng = np.random.default_rng(123)
    inputs = rng.integers(0, 2**32, size=(64, 6), dtype=np.uint64).astype(np.uint32)
    inputs[0] = 0
    inputs[1] = 0xFFFFFFFF
    inputs[2, :4] = [7, 1, 0, 0]  # a step-pair counter of path 1, the layout the pricer uses
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "philox.cpp")
        exe = os.path.join(d, "philox")
        open(src, "w").write(ROCRAND_SRC)
        subprocess.run(["hipcc", "-x", "hip", "--offload-arch=gfx950", "-O1", src, "-o", exe], check=True,
                       stderr=subprocess.DEVNULL)
        stdin = "\n".join(" ".join(f"{x:08x}" for x in row) for row in inputs)
        out = subprocess.run([exe], input=stdin, capture_output=True, text=True, check=True).stdout.split("\n")
    res = [[int(x, 16) for x in line.split()] for line in out if line.strip()]
    return {"source": "rocRAND philox4x32_10_engine::ten_rounds (ROCm 7.2, host build)",
            "ctr": inputs[:, :4].tolist(), "key": inputs[:, 4:].tolist(), "out": res}


def ndtri():
    p = np.concatenate([10.0 ** -np.arange(300, 0, -7.0), np.linspace(0.001, 0.999, 401),
                        1 - 10.0 ** -np.arange(1, 17, 0.5), [0.025, 0.5, 0.975, 2.0 ** -33, 1 - 2.0 ** -33]])
    p = np.unique(p[(p > 0) & (p < 1)])
    return {"source": "scipy.special.ndtri", "p": p.tolist(), "z": special.ndtri(p).tolist()}


def heston_call_lewis(S, K, T, r, v0, theta, rho, kappa, sigma):
    """C = S - sqrt(S K) e^{-rT/2}/pi * int_0^inf Re[e^{i u k} phi(u - i/2)] / (u^2 + 1/4) du,
    k = log(S/K) + rT, phi the characteristic function of log(S_T/S_0) - rT (little trap form)."""
    def phi(u):
        xi = kappa - sigma * rho * 1j * u
        d = np.sqrt(xi * xi + sigma * sigma * (u * u + 1j * u))
        g = (xi - d) / (xi + d)
        e = np.exp(-d * T)
        C = kappa * theta / sigma ** 2 * ((xi - d) * T - 2.0 * np.log((1 - g * e) / (1 - g)))
        D = (xi - d) / sigma ** 2 * (1 - e) / (1 - g * e)
        return np.exp(C + D * v0)

    k = np.log(S / K) + r * T

    def f(u):
        return (np.exp(1j * u * k) * phi(u - 0.5j)).real / (u * u + 0.25)

    val, _ = integrate.quad(f, 0.0, np.inf, limit=2000, epsabs=1e-14, epsrel=1e-13)
    return S - np.sqrt(S * K) * np.exp(-r * T / 2) / np.pi * val


def heston_closed_form():
    cases = []
    # the reference pricing test's model (src/UnitTest.cpp:182-188): S = 1, r = 0.02
    p = dict(v0=0.09, theta=0.09, rho=0.04, kappa=1.0, sigma=1.0)
    for T, Ks in [(0.119047619047619, [0.9371, 0.9956, 1.0427, 1.2287, 1.3939]),
                  (1.42857142857143, [0.6137, 0.9025, 1.0766, 1.3046, 1.5328])]:
        cases.append({"S": 1.0, "r": 0.02, "T": T, "K": Ks, **p,
                      "price": [heston_call_lewis(1.0, K, T, 0.02, **p) for K in Ks]})
    # the strike/maturity grid's model: S0 = 100, r = 0.02
    p = dict(v0=0.04, theta=0.04, rho=-0.7, kappa=2.0, sigma=0.3)
    Ks = list(np.linspace(70.0, 130.0, 9))
    for T in [0.1, 0.25, 1.0, 5.0]:
        cases.append({"S": 100.0, "r": 0.02, "T": T, "K": Ks, **p,
                      "price": [heston_call_lewis(100.0, K, T, 0.02, **p) for K in Ks]})
    return {"source": "Lewis (2000) single integral, little-trap Heston chf, scipy.integrate.quad", "cases": cases}


def main():
    for name, fn in [("philox_rocrand.json", philox_rocrand), ("ndtri.json", ndtri),
                     ("heston_closed_form.json", heston_closed_form)]:
        data = fn()
        with open(os.path.join(HERE, name), "w") as fh:
            json.dump(data, fh, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
