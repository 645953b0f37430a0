"""Known-answer checks of the oracle's random number generators."""
import ctypes

import numpy as np
import pytest

from oracle import pyoracle as po

# Random123 kat_vectors, philox4x32 10 rounds: (ctr, key) -> out
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert po.philox4x32_10(ctr, key) == out


def test_mt19937_standard_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937
    assert int(po.mt19937(5489, 10000)[-1]) == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 1234, 2**32 - 1])
def test_mt19937_matches_numpy_legacy_seeding(seed):
    g = np.random.MT19937()
    g._legacy_seeding(seed)
    assert np.array_equal(g.random_raw(2000), po.mt19937(seed, 2000))


@pytest.mark.parametrize("seed", [0, 1, 2, 42, 123456789, 2**31 - 1])
def test_glibc_rand_restatement_matches_libc(seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    want = np.array([libc.rand() for _ in range(3000)], dtype=np.int64)
    assert np.array_equal(po.glibc_rand(seed, 3000), want)


def test_poisson_uniform_range_and_mean():
    us = np.array([po.poisson_u(7, i, s) for i in range(50) for s in range(200)])
    assert us.min() > 0.0 and us.max() <= 1.0
    assert abs(us.mean() - 0.5) < 0.01
    # 24-bit grid: every value is a multiple of 2^-24
    assert np.all(np.round(us * 2**24) == us * 2**24)


def test_poisson_stream_uses_philox_word_of_step():
    # u(seed, id, step) = ((x >> 8) + 1) * 2^-24, x = word (step & 3) of Philox(ctr={id, step>>2, 0, 'POIS'})
    seed, nid = 0x1234_5678_9ABC, 17
    for step in (0, 1, 2, 3, 4, 9, 1001):
        out = po.philox4x32_10([nid, step >> 2, 0, 0x504F4953], [seed & 0xFFFFFFFF, seed >> 32])
        x = out[step & 3]
        assert po.poisson_u(seed, nid, step) == np.float32(((x >> 8) + 1) * 2.0**-24)
