"""Pins of the oracle's xorshift128+ (src/mmc_rand_xorshift128p.c:55-83) and
scattering-length draw (src/mmc_rand_common.h:50-51)."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O

MASK = (1 << 64) - 1


def step(state):
    st = (C.c_uint64 * 2)(*state)
    u = O.lib().orc_rand_next(st)
    return u, (st[0], st[1])


def test_known_answer_single_bit_state():
    # state (1, 0): s1 = 1 ^ (1 << 23); t1 = s1 ^ 0 ^ (s1 >> 18) ^ (0 >> 5) = 1 + 2^5 + 2^23
    # output low word 8388641 >> 9 = 16384 -> u = 16384 * 2^-23 = 2^-9
    u, st = step((1, 0))
    assert st == (0, 1 + (1 << 5) + (1 << 23))
    assert u == np.float32(2.0 ** -9)


def test_known_answer_second_word():
    # state (0, 1): s1 = 0, s0 = 1; t1 = 0 ^ 1 ^ 0 ^ (1 >> 5) = 1; sum t1 + s0 = 2 -> u = 0
    u, st = step((0, 1))
    assert st == (1, 1)
    assert u == 0.0
    # state (0, 2^40): s0 >> 5 = 2^35; t1 = 2^40 ^ 2^35; sum = 2^41 + 2^35 -> low word 0 -> u = 0
    u, st = step((0, 1 << 40))
    assert st == (1 << 40, (1 << 40) ^ (1 << 35))
    assert u == 0.0


def _next_state(a):
    """Transition only (no output) via the oracle."""
    return step(a)[1]


def test_state_transition_is_linear_over_gf2():
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = tuple(int(x) for x in rng.integers(0, 2**63, size=2, dtype=np.int64))
        b = tuple(int(x) for x in rng.integers(0, 2**63, size=2, dtype=np.int64))
        ab = (a[0] ^ b[0], a[1] ^ b[1])
        na, nb, nab = _next_state(a), _next_state(b), _next_state(ab)
        assert nab == (na[0] ^ nb[0], na[1] ^ nb[1])


def test_output_is_sum_of_new_state_and_old_second_word():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = tuple(int(x) for x in rng.integers(0, 2**63, size=2, dtype=np.int64))
        u, st = step(a)
        low = ((st[1] + a[1]) & MASK) & 0xFFFFFFFF
        assert u == np.float32((low >> 9) * 2.0 ** -23)


def test_seed_words():
    seed = (C.c_uint32 * 4)(0x11111111, 0x22222222, 0x33333333, 0x44444444)
    st = (C.c_uint64 * 2)()
    O.lib().orc_rand_seed(seed, st)
    assert st[0] == 0x1111111122222222 and st[1] == 0x3333333344444444


def test_uniform_and_exponential_draws():
    st = (C.c_uint64 * 2)(0x0123456789ABCDEF, 0xFEDCBA9876543210)
    L = O.lib()
    u = np.array([L.orc_rand_next(st) for _ in range(200000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert np.all(u * 2**23 == np.floor(u * 2**23))
    counts, _ = np.histogram(u, bins=50, range=(0, 1))
    chi2 = ((counts - len(u) / 50) ** 2 / (len(u) / 50)).sum()
    assert chi2 < 100  # 49 dof, p ~ 1e-5
    s = np.array([L.orc_rand_scatlen(st) for _ in range(200000)])
    assert abs(s.mean() - 1.0) < 0.01
    # Kolmogorov-Smirnov against Exp(1) (the +1e-6 shift is far below the KS resolution)
    xs = np.sort(s)
    cdf = 1 - np.exp(-xs)
    d = np.max(np.abs(cdf - np.arange(1, len(xs) + 1) / len(xs)))
    assert d < 1.63 / np.sqrt(len(xs))


@pytest.mark.parametrize("math", ["glibc", "portable"])
def test_both_oracle_builds_share_the_generator(math):
    a = (C.c_uint64 * 2)(7, 9)
    b = (C.c_uint64 * 2)(7, 9)
    for _ in range(100):
        assert O.lib("glibc").orc_rand_next(a) == O.lib(math).orc_rand_next(b)
