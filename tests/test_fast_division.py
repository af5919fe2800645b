"""The multiplier division the device assembly uses for index digits
(quac_amd/csrc/qh_assemble.hip: fdiv_make / fdiv): with l = ceil(log2 d) and
m = floor(2^32 (2^l - d) / d) + 1, n / d = (mulhi(n, m) + n) >> l for every
32-bit n (Granlund & Montgomery 1994), the sum taken in 64 bits.  Restated here
in integer Python and checked against exact division."""
import random

import pytest

M32 = (1 << 32) - 1


def fdiv_make(d):
    sh = 0
    while sh < 32 and (1 << sh) < d:
        sh += 1
    m = ((1 << 32) * ((1 << sh) - d)) // d + 1
    assert 0 <= m <= M32
    return m & M32, sh


def fdiv(n, m, sh):
    return (((n * m) >> 32) + n) >> sh


def numerators(d, rng):
    out = {0, 1, d - 1, d, d + 1, M32, M32 - 1, M32 - d, (1 << 31), (1 << 31) - 1}
    q = M32 // d
    for k in (1, 2, 3, q - 1, q):
        for e in (-1, 0, 1):
            out.add(k * d + e)
    out.update(rng.randrange(0, 1 << 32) for _ in range(64))
    return [n for n in out if 0 <= n <= M32]


@pytest.mark.parametrize("lo,hi", [(1, 1500), (1500, 5000)])
def test_small_divisors(lo, hi):
    rng = random.Random(lo)
    for d in range(lo, hi):
        m, sh = fdiv_make(d)
        for n in numerators(d, rng):
            assert fdiv(n, m, sh) == n // d, (d, n)


def test_large_and_power_of_two_divisors():
    rng = random.Random(7)
    ds = [1 << k for k in range(32)] + [(1 << k) + 1 for k in range(1, 32)] + [(1 << k) - 1 for k in range(2, 33)]
    ds += [rng.randrange(5000, 1 << 32) for _ in range(400)]
    for d in ds:
        m, sh = fdiv_make(d)
        for n in numerators(d, rng):
            assert fdiv(n, m, sh) == n // d, (d, n)
