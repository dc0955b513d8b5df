"""CPU oracle: Philox4x32-10 and AS241, pinned by independent sources.

Philox: the Random123 known-answer vectors and rocRAND's philox4x32_10 engine (tests/golden).
AS241: scipy.special.ndtri (a library inverse normal), the coefficient hash sums printed in the
algorithm's published listing (src/as241.f90:45,64,83), symmetry and the IFAULT cases.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_random123_kat(ctr, key, out):
    assert O.philox(ctr, key).tolist() == out


def test_philox_matches_rocrand():
    g = json.load(open(os.path.join(GOLD, "philox_rocrand.json")))
    for ctr, key, out in zip(g["ctr"], g["key"], g["out"]):
        assert O.philox(ctr, key).tolist() == out


def test_stream_layout_and_jump_ahead():
    """Word w of path p is element w%4 of Philox({w/4, lo(p), hi(p), stream}, seed): any window of
    any path is reachable directly (jump-ahead), and windows agree with the full stream."""
    seed, path = 0x0123456789ABCDEF, (5 << 32) + 17
    full = O.rng_words(seed, path, 1, 0, 64)[0]
    for w in range(0, 64, 4):
        blk = O.philox([w // 4, path & 0xFFFFFFFF, path >> 32, 0], [seed & 0xFFFFFFFF, seed >> 32])
        assert full[w:w + 4].tolist() == blk.tolist()
    win = O.rng_words(seed, path, 1, 37, 11)[0]
    assert win.tolist() == full[37:48].tolist()
    other_stream = O.rng_words(seed, path, 1, 0, 8, stream=1)[0]
    assert other_stream.tolist() != full[:8].tolist()


def test_uniform_moments():
    """Moments of u = (k + 1/2) 2^-32 over 2^20 stream words, in the spirit of src/UnitTest.cpp:528-564."""
    w = O.rng_words(42, 0, 1024, 0, 1024).reshape(-1)
    u = (w.astype(np.float64) + 0.5) * 2.0 ** -32
    n = u.size
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
    assert abs(u.var() - 1 / 12) < 5 * np.sqrt(1 / 180 / n)
    # bits are balanced and adjacent words uncorrelated
    bits = np.unpackbits(w.view(np.uint8)).mean()
    assert abs(bits - 0.5) < 1e-3
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 5 / np.sqrt(n)


def test_uniform_symmetry_and_range():
    for k in [0, 1, 2**31 - 1, 2**31, 2**32 - 1]:
        u = O.load().orc_uniform_from_word(k)
        assert 0.0 < u < 1.0
        assert u + O.load().orc_uniform_from_word((~k) & 0xFFFFFFFF) == 1.0


def test_as241_coefficient_hash_sums():
    """Sums of the decimal mantissas of the coefficients, as printed with the published algorithm
    (src/as241.f90:45,64,83): a transcription error in any coefficient changes them."""
    c = O.as241_coefficients()
    ab = sum(mant(x) for x in c[0:15])
    cd = sum(mant(x) for x in c[15:30])
    ef = sum(mant(x) for x in c[30:45])
    assert abs(ab - 55.8831928806149014439) < 1e-12
    assert abs(cd - 49.3320650330161028903) < 1e-12  # 49.33206 50330 16102 89036
    assert abs(ef - 47.5258331754928967163) < 1e-12  # 47.52583 31754 92896 71629


def mant(x):
    return x / 10.0 ** np.floor(np.log10(abs(x)))


def test_as241_f64_against_ndtri():
    g = json.load(open(os.path.join(GOLD, "ndtri.json")))
    p = np.array(g["p"])
    z = np.array(g["z"])
    ours = np.array([O.ppnd16_f64(x)[0] for x in p])
    rel = np.abs(ours - z) / np.maximum(np.abs(z), 1e-300)
    central = np.abs(p - 0.5) <= 0.425
    assert rel[central & (np.abs(z) > 1e-3)].max() < 5e-15
    assert rel[~central].max() < 5e-15  # AS241 claims ~1e-16; both tail rationals
    assert np.abs(ours[np.abs(z) <= 1e-3] - z[np.abs(z) <= 1e-3]).max() < 1e-17


def test_as241_special_cases():
    assert O.ppnd16_f64(0.5) == (0.0, 0)
    for p in [0.0, 1.0, -0.1, 1.5]:
        assert O.ppnd16_f64(p) == (0.0, 1)
    # dyadic p: 1 - p is exact, so q and the tail argument are exactly mirrored
    for p in [2.0 ** -40, 2.0 ** -7, 0.3125, 0.421875, 0.4375]:
        assert O.ppnd16_f64(p)[0] == -O.ppnd16_f64(1.0 - p)[0]


def test_as241_asbuilt_fp32_error_is_single_precision():
    """The reference evaluates AS241 in REAL*4 (src/as241.f90:20-25): ~1e-7 relative, not 1e-16."""
    g = json.load(open(os.path.join(GOLD, "ndtri.json")))
    p = np.array(g["p"])
    z = np.array(g["z"])
    sel = (p > 1e-30) & (np.abs(z) > 1e-2)
    ours = np.array([O.ppnd16_asbuilt(x)[0] for x in p[sel]])
    rel = np.abs(ours - z[sel]) / np.abs(z[sel])
    assert rel.max() < 2e-6
    assert rel.max() > 1e-9  # it really is single precision


def test_fp32_normal_from_word_against_ndtri():
    """The FP32 kernel's normal (its oracle restatement), driven by the 32-bit word, against scipy's
    ndtri: within 8e-7 absolute, and full relative precision in both tails."""
    rng = np.random.default_rng(3)
    w = rng.integers(0, 2**32, 4000, dtype=np.uint64).astype(np.uint32)
    w[:4] = [0, 1, 2**32 - 2, 2**32 - 1]
    u = (w.astype(np.float64) + 0.5) * 2.0 ** -32
    from scipy.special import ndtri
    z = ndtri(u)
    ours = O.normals_f32(w).astype(np.float64)
    assert np.abs(ours - z).max() < 8e-7
    sel = np.abs(z) > 1.0
    assert (np.abs(ours[sel] - z[sel]) / np.abs(z[sel])).max() < 3.5e-7
    # the extreme words reach |z| ~ 6.34 without losing precision (no 1 - u cancellation)
    assert abs(ours[0] - z[0]) / abs(z[0]) < 3e-7 and abs(ours[3] - z[3]) / abs(z[3]) < 3e-7


def test_normal_moments():
    w = O.rng_words(7, 0, 256, 0, 1024).reshape(-1)
    z = O.normals_f64(w)
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs((z ** 4).mean() - 3) < 5 * np.sqrt(96 / n)
