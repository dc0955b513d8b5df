"""CPU oracle: the FP32 kernel's inverse normal (oracle/hexo_oracle.c orc_normal_f32_from_word, the
restatement of device.hpp normal_word_f32), pinned by what the mathematics fixes:

- the FP64 PPND16 of the same word (itself pinned against scipy's ndtri and AS241's hash sums in
  test_oracle_rng_as241.py): within 8e-7 absolute where the central polynomial applies and 3.5e-7
  relative in the tail, on random words, both ends of the word range and both sides of the
  central / tail switch;
- exact antisymmetry z(~k) = -z(k) (u and 1 - u are mirrored words), which makes the mean of the
  normals over all words exactly zero;
- monotonicity in k, up to the 2e-6 step allowed where the two polynomials meet.

The same check rejects CPU copies with a planted mistake (a dropped top coefficient, a wrong sign
of one coefficient, the fit's shifted variable l + 6.5 used with the expanded coefficients, the
tail switch at the wrong place).
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "hestonexotics_amd", "csrc"))
import normal_f32_fit as F  # noqa: E402

L_SWITCH_J = 2 ** 21  # l = log2(s (1 - s)) crosses -11 at s ~ 2^-11, i.e. j = min(k, ~k) ~ 2^21


def words():
    rng = np.random.default_rng(7)
    k = np.concatenate([rng.integers(0, 2 ** 32, 60000), np.arange(0, 3000), np.arange(2 ** 32 - 3000, 2 ** 32),
                        np.arange(2 ** 31 - 3000, 2 ** 31 + 3000), np.arange(L_SWITCH_J - 3000, L_SWITCH_J + 3000),
                        2 ** 32 - 1 - np.arange(L_SWITCH_J - 3000, L_SWITCH_J + 3000)])
    return np.unique(k).astype(np.uint64)


def check_normal_f32(fn):
    """(ok, message) for a words -> FP32 normals function."""
    k = words()
    z = np.asarray(fn(k), np.float64)
    ref = O.normals_f64(k.astype(np.uint32))
    t = np.minimum(k, 2 ** 32 - 1 - k) < L_SWITCH_J  # tail words (l < -11), up to the switch's rounding
    err = np.abs(z - ref)
    central_abs = err[~t].max()
    tail_rel = (err[t] / np.abs(ref[t])).max()
    if central_abs > 8e-7:
        return False, f"central error {central_abs:.3e}"
    if tail_rel > 3.5e-7:
        return False, f"tail relative error {tail_rel:.3e}"
    mirror = np.asarray(fn(2 ** 32 - 1 - k), np.float64)
    if not np.array_equal(mirror, -z):
        return False, "not antisymmetric"
    for lo in (0, L_SWITCH_J - 2000, 2 ** 31 - 2000, 2 ** 32 - 2000):
        seg = np.asarray(fn(np.arange(lo, lo + 2000, dtype=np.uint64)), np.float64)
        if np.diff(seg).min() < -2e-6:
            return False, f"not monotone near {lo}"
    return True, ""


def test_oracle_normal_f32_pinned():
    ok, msg = check_normal_f32(lambda k: O.normals_f32(k.astype(np.uint32)))
    assert ok, msg


def test_oracle_normal_f32_special_words():
    z = O.normals_f32(np.array([0, 1, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 2, 2 ** 32 - 1], np.uint32)).astype(np.float64)
    ref = O.normals_f64(np.array([0, 1, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 2, 2 ** 32 - 1], np.uint32))
    # the extreme words (u = 2^-33: z = -6.3380) to FP32 precision; the two middle words are within
    # the conversion's rounding of 0 (|z| = 3e-10 exactly)
    assert np.abs(z[[0, 1, 4, 5]] - ref[[0, 1, 4, 5]]).max() < 2e-6
    assert abs(z[2]) < 1e-7 and abs(z[3]) < 1e-7 and z[0] < z[1] < 0 < z[4] < z[5]


def test_fit_script_reproduces_the_coefficients():
    """normal_f32_fit.py's coefficients, evaluated in float32 steps, are the oracle's (the same
    values to the last bit: the mirror and the oracle agree to libm's log2f rounding)."""
    central, tail, yc = F.fit()
    k = words()[::7]
    a = F.normal_f32(k, central, tail, yc).astype(np.float64)
    b = O.normals_f32(k.astype(np.uint32)).astype(np.float64)
    assert np.abs(a - b).max() < 3e-7
    assert yc == 4.5305938720703125


def _planted(kind):
    central, tail, yc = F.fit()
    central, tail = central.copy(), tail.copy()
    shift = 0.0
    switch = F.L_C
    if kind == "drop_top":
        central[-1] = 0.0
    elif kind == "coef_sign":
        central[3] = -central[3]
    elif kind == "shifted":
        shift = F.X_C
    elif kind == "switch":
        switch = -9.0

    def fn(k):
        k = np.asarray(k, np.uint64)
        j = np.where(k >= 2 ** 31, 2 ** 32 - 1 - k, k)
        s = F.f32_fma(j.astype(np.float32), np.float32(2.0 ** -32), np.float32(2.0 ** -33))
        t = F.f32_fma(-s, s, s)
        l = np.log2(t.astype(np.float64)).astype(np.float32)
        hc = F.horner(central, (l.astype(np.float64) + shift).astype(np.float32))
        y = (np.sqrt(-l.astype(np.float64)).astype(np.float32).astype(np.float64) - yc).astype(np.float32)
        h = np.where(l < np.float32(switch), F.horner(tail, y), hc)
        pr = (h.astype(np.float64) * F.f32_fma(np.full_like(s, -2.0), s, 1.0).astype(np.float64)).astype(np.float32)
        return np.where(k >= 2 ** 31, pr, -pr)
    return fn


def test_check_accepts_the_unplanted_copy():
    ok, msg = check_normal_f32(_planted("none"))
    assert ok, msg


@pytest.mark.parametrize("kind", ["drop_top", "coef_sign", "shifted", "switch"])
def test_check_rejects_planted_mistakes(kind):
    ok, _ = check_normal_f32(_planted(kind))
    assert not ok, kind
