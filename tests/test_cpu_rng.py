"""RNG pins: the MWC recurrence, the draw conversion, the safe-prime table and the photon
stream seeding (src/octmps_kernel.cuh:10-29, 52-172)."""
import random

import numpy as np
import pytest

import octmps as O
import oracle_py as OR


def is_prime(n):
    # Miller-Rabin with the first 13 prime bases: deterministic below 3.3e24.
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def qualifies(a):
    return is_prime(a * 2 ** 32 - 1) and is_prime(a * 2 ** 31 - 1)


def test_safeprimes_property_order_and_completeness():
    m = [int(v) for v in O.safeprimes(24)]
    assert all(qualifies(a) for a in m)
    assert m == sorted(m, reverse=True) and m[0] < 2 ** 32
    # No qualifying multiplier is skipped: brute force every a from 2^32-1 down to m[5].
    for a in range(2 ** 32 - 1, m[5] - 1, -1):
        assert qualifies(a) == (a in m[:6]), a


def test_safeprimes_start_of_table():
    # Largest multipliers below 2^32 with a*2^32-1 and a*2^31-1 prime, found by brute force.
    m = O.safeprimes(3)
    expect, a = [], 2 ** 32 - 1
    while len(expect) < 3:
        if qualifies(a):
            expect.append(a)
        a -= 1
    assert list(map(int, m)) == expect


def rz24(u):
    # __uint2float_rz: keep the 24 most significant bits.
    b = u.bit_length()
    return u if b <= 24 else (u >> (b - 24)) << (b - 24)


def test_mwc_draws_against_big_integers():
    a = int(O.safeprimes(4)[3])
    rng = random.Random(5)
    for _ in range(20):
        x = (rng.randrange(a - 1) << 32) | rng.randrange(2 ** 32 - 1)
        vals, xe = OR.rand_co(x, a, 50)
        xs = x
        for v in vals:
            xs = (xs & 0xFFFFFFFF) * a + (xs >> 32)
            u = xs & 0xFFFFFFFF
            assert v == rz24(u) / 2.0 ** 32
            assert 0.0 <= v <= 1.0 - 2.0 ** -24
        assert xe == xs


def test_mwc_is_the_lcg_with_inverse_base():
    # b * X' = X (mod a*b - 1): the MWC step is multiplication by b^-1 modulo p = a*2^32-1,
    # which gives the generator its period (p-1)/2 when p is a safe prime.
    a = int(O.safeprimes(1)[0])
    p = a * 2 ** 32 - 1
    binv = pow(2 ** 32, -1, p)
    rng = random.Random(9)
    for _ in range(200):
        x = (rng.randrange(a - 1) << 32) | rng.randrange(2 ** 32 - 1)
        _, xn = OR.rand_co(x, a, 1)
        assert xn == x * binv % p


def test_photon_streams_valid_and_identical_in_product_and_oracle(table):
    for photon in list(range(64)) + [2 ** 33 + 7, 10 ** 12]:
        x1, a1 = O.photon_streams(1, 3, photon)
        x2, a2 = OR.photon_streams(1, 3, photon, table)
        assert np.array_equal(x1, x2) and np.array_equal(a1, a2)
        k = photon % table.k
        assert list(map(int, a1)) == [int(v) for v in table.mult[3 + 3 * k:6 + 3 * k]]
        for s in range(3):
            X, A = int(x1[s]), int(a1[s])
            assert X != 0 and (X >> 32) < A - 1 and (X & 0xFFFFFFFF) < 2 ** 32 - 1


def test_streams_differ_across_photons_salts_and_seeds():
    seen = set()
    for seed in (1, 2):
        for salt in (0, 1):
            for p in range(200):
                x, _ = O.photon_streams(seed, salt, p)
                seen.add(tuple(int(v) for v in x))
    assert len(seen) == 2 * 2 * 200


@pytest.mark.parametrize("a,b", [(1.0, 3.0), (-1.0, 3.0), (2.0, -7.0), (0.1, 6.0), (1e-3, 1e-9)])
def test_div_rd_is_round_toward_minus_infinity(a, b):
    from fractions import Fraction
    q = OR.div_rd(a, b)
    exact = Fraction(a) / Fraction(b)
    assert Fraction(q) <= exact < Fraction(np.nextafter(q, np.inf))


def test_div_rd_random():
    from fractions import Fraction
    rng = np.random.default_rng(1)
    for a, b in rng.normal(size=(2000, 2)) * 10 ** rng.uniform(-3, 3, size=(2000, 2)):
        q = OR.div_rd(float(a), float(b))
        exact = Fraction(float(a)) / Fraction(float(b))
        assert Fraction(q) <= exact < Fraction(np.nextafter(q, np.inf))
