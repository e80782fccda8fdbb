"""Deterministic math and random-number layer shared by kernels and oracle.

Pinned against: mpmath at 50 digits (elementary functions, ulp bounds) and
the Philox4x32-10 known-answer values published with the Random123 release
of Salmon et al., SC'11."""
import math
import struct

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")
mpmath.mp.dps = 50


def ulp_err(got, exact):
    e = float(exact)
    if e == 0.0:
        return 0.0 if got == 0.0 else float("inf")
    return abs(mpmath.mpf(got) - exact) / math.ulp(e)


CASES = {
    "trk_det_exp": (lambda x: mpmath.exp(x), lambda r: r.uniform(-700, 700, 400)),
    "trk_det_log": (lambda x: mpmath.log(x), lambda r: np.exp(r.uniform(-690, 690, 400))),
    "trk_det_sin": (lambda x: mpmath.sin(x), lambda r: r.uniform(-60, 60, 400)),
    "trk_det_cos": (lambda x: mpmath.cos(x), lambda r: r.uniform(-60, 60, 400)),
    "trk_det_acos": (lambda x: mpmath.acos(x), lambda r: r.uniform(-1, 1, 400)),
    "trk_det_atan": (lambda x: mpmath.atan(x), lambda r: np.sinh(r.uniform(-20, 20, 400))),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_detmath_ulp(lib, name):
    ref, gen = CASES[name]
    xs = gen(np.random.default_rng(abs(hash(name)) % 2**32))
    f = getattr(lib, name)
    worst = 0.0
    for x in xs:
        x = float(x)
        ex = ref(mpmath.mpf(x))
        if name in ("trk_det_sin", "trk_det_cos") and abs(ex) < 1e-3:
            continue          # relative error near zeros is not the contract
        worst = max(worst, ulp_err(f(x), ex))
    assert worst <= 16, worst


def test_detmath_pow(lib):
    r = np.random.default_rng(7)
    worst = 0.0
    for x, y in zip(np.exp(r.uniform(-20, 20, 300)), r.uniform(-3, 3, 300)):
        ex = mpmath.power(mpmath.mpf(float(x)), mpmath.mpf(float(y)))
        worst = max(worst, ulp_err(lib.trk_det_pow(float(x), float(y)), ex))
    assert worst <= 64, worst


def test_detmath_special_values(lib):
    assert lib.trk_det_log(1.0) == 0.0
    assert lib.trk_det_exp(0.0) == 1.0
    assert lib.trk_det_acos(1.0) == 0.0
    assert abs(lib.trk_det_acos(-1.0) - math.pi) <= math.ulp(math.pi)
    assert lib.trk_det_sin(0.0) == 0.0
    assert lib.trk_det_cos(0.0) == 1.0
    assert lib.trk_det_pow(2.0, 10.0) == 1024.0
    assert lib.trk_det_pow(8.0, 1.0 / 3.0) == pytest.approx(2.0, rel=1e-15)
    assert lib.trk_det_exp(-800.0) == 0.0


# --------------------------------------------------------------------- Philox
def philox4x32_10(ctr, key):
    """Philox4x32-10 written from the SC'11 paper (round function, Table 2
    multipliers, Weyl key schedule)."""
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = list(ctr)
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        c = [(p1 >> 32) ^ c[1] ^ k0, p1 & 0xFFFFFFFF, (p0 >> 32) ^ c[3] ^ k1, p0 & 0xFFFFFFFF]
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return c


KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(ctr, key, out):
    assert tuple(philox4x32_10(ctr, key)) == out


def uniform_from_words(a, b):
    u = ((a << 32) | b) >> 11
    return (u + 0.5) * 2.0 ** -53


def test_rng_uniform_kat(lib):
    # draw 0 of particle 0, iteration 0, seed 0 = words 0,1 of the all-zero KAT block
    w = KAT[0][2]
    assert lib.trk_rng_uniform(0, 0, 0, 0) == uniform_from_words(w[0], w[1])
    assert lib.trk_rng_uniform(0, 0, 0, 1) == uniform_from_words(w[2], w[3])


def test_rng_uniform_stream_layout(lib):
    r = np.random.default_rng(3)
    for _ in range(200):
        seed = int(r.integers(0, 2**64, dtype=np.uint64))
        pid = int(r.integers(0, 2**64, dtype=np.uint64))
        it = int(r.integers(0, 2**32))
        n = int(r.integers(0, 2**32))
        w = philox4x32_10((n >> 1, pid & 0xFFFFFFFF, pid >> 32, it), (seed & 0xFFFFFFFF, seed >> 32))
        exp = uniform_from_words(w[0], w[1]) if n % 2 == 0 else uniform_from_words(w[2], w[3])
        got = lib.trk_rng_uniform(seed, pid, it, n)
        assert got == exp
        assert 0.0 < got < 1.0


def test_rng_uniform_moments(lib):
    xs = np.array([lib.trk_rng_uniform(0x5EED0002, 12345, 7, n) for n in range(20000)])
    assert abs(xs.mean() - 0.5) < 0.01
    assert abs(xs.var() - 1.0 / 12.0) < 0.003
