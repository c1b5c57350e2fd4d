"""Host twin of the device start-band generator (systems.counter_bands).

Pinned against the published SplitMix64 output for state 0
(0xE220A8397B1DCDAF, the generator's first value from seed 0), and against
exact Python-integer arithmetic for large seeds, band offsets and indices,
which checks numpy's wrapping uint64 products and shifts.
"""
import numpy as np

from rsdft_develop_amd import systems

M = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def _value(seed, n, b, j):
    z = (seed + (b * n + j + 1) * GOLDEN) & M
    return (_mix(z) >> 11) * 2.0 ** -53


def test_first_value_is_splitmix64_of_state_zero():
    u = systems.counter_bands(5, 1, seed=0)
    assert u[0, 0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53


def test_matches_exact_integer_arithmetic():
    n = 1000
    for seed, n0 in ((0, 0), (2 ** 63 + 12345, 3), (M, 7)):
        u = systems.counter_bands(n, 4, seed=seed, n0=n0)
        for b in range(4):
            for j in (0, 1, 499, 999):
                assert u[j, b] == _value(seed, n, n0 + b, j), (seed, n0, b, j)
        assert u.flags["F_CONTIGUOUS"] and u.min() >= 0.0 and u.max() < 1.0


def test_uniform_moments():
    u = systems.counter_bands(20000, 3, seed=9)
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1.0 / 12.0) < 0.005
