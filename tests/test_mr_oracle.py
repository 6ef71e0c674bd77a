"""CPU checks that pin the GMP Miller-Rabin oracle: N and rev(N) against
Python big integers and decimal strings, per-base verdicts against a
textbook strong-probable-prime test and known (pseudo)primes."""
import random
import sys

import pytest

from _support import (gmp_build, gmp_powm, gmp_probab_prime, gmp_sprp, gmp_version, inv_pow2_mod,
                      sprp_python)

MERSENNE_EXP = [13, 17, 19, 31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281]


def test_gmp_version_recorded():
    assert gmp_version().startswith("6.")


@pytest.mark.parametrize("a,b,c,k,j", [(30, 15, 7, 12, 34), (50, 20, 3, 4096, 7), (100, 99, 1, 1, 1),
                                       (31, 27, 5, 4096, 7), (32, 8, 5, 4096, 4096), (30, 25, 5, 4096, 7),
                                       (60, 40, 10, 100, 100), (12, 6, 3, 9, 99)])
def test_build_matches_python(a, b, c, k, j):
    n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
    assert gmp_build(a, b, c, k, j, 0) == n
    assert gmp_build(a, b, c, k, j, 1) == int(str(n)[::-1])
    lk, lj = len(str(k)), len(str(j))
    if a - b - lk < 0 or a - c - lj < 0:
        with pytest.raises(ValueError):      # formula undefined: negative exponent
            gmp_build(a, b, c, k, j, 2)
        return
    formula = 10 ** a - int(str(k)[::-1]) * 10 ** (a - b - lk) - int(str(j)[::-1]) * 10 ** (a - c - lj) - 1
    assert gmp_build(a, b, c, k, j, 2) == formula


def test_rev_formula_differs_only_when_blocks_overlap():
    # F4: the closed formula is the true reversal iff the k and j digit
    # blocks do not overlap and N keeps a digits.
    rng = random.Random(4)
    for _ in range(200):
        a = rng.randint(20, 60)
        b = rng.randint(2, a - 1)
        c = rng.randint(1, b - 1)
        k, j = rng.randint(1, 4096), rng.randint(1, 4096)
        lk, lj = len(str(k)), len(str(j))
        if a - b - lk < 0 or a - c - lj < 0:
            continue
        n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
        same = gmp_build(a, b, c, k, j, 1) == gmp_build(a, b, c, k, j, 2)
        clean = b >= c + lj and b + lk <= a and len(str(n)) == a
        if clean:
            assert same, (a, b, c, k, j)


def test_sprp_matches_textbook():
    rng = random.Random(9)
    for bits in (40, 200, 1000, 3000):
        for _ in range(6):
            n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
            for base in (2, 3, 5, 7, 53):
                v, r, s = gmp_sprp(n, base)
                tv, tr, ts = sprp_python(n, base)
                assert (bool(v), r, s) == (tv, tr, ts)


def test_known_primes_and_strong_pseudoprimes():
    for p in MERSENNE_EXP:
        n = 2 ** p - 1
        for base in (2, 3, 5, 7, 11, 13):
            assert gmp_sprp(n, base)[0] == 1
        assert gmp_probab_prime(n) >= 1
    # 2047 = 23 * 89 is a strong pseudoprime to base 2 only; 3215031751 to
    # bases 2, 3, 5 and 7 (OEIS A014233) -- per-base MR passes, BPSW does not.
    assert [gmp_sprp(2047, b)[0] for b in (2, 3)] == [1, 0]
    assert [gmp_sprp(3215031751, b)[0] for b in (2, 3, 5, 7, 11)] == [1, 1, 1, 1, 0]
    assert gmp_probab_prime(3215031751) == 0
    # the reference's phase-2 test: N = 9999 - 100k - 10j against trial division
    for k in range(1, 6):
        for j in range(1, 6):
            n = 10000 - 100 * k - 10 * j - 1
            td = all(n % d for d in range(2, int(n ** 0.5) + 1))
            assert (gmp_probab_prime(n) >= 1) == td


def test_powm_matches_python():
    """mr_oracle_powm (GMP mpz_powm with a full exponent; the checker of the
    checkpoint residues and the Miller-Rabin CPU baseline) against Python's
    pow, including exponents longer and shorter than the modulus, base 0
    and base >= N."""
    rng = random.Random(41)
    for bits in (17, 64, 300, 2000):
        for _ in range(6):
            n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
            b = rng.randrange(3 * n)
            e = rng.getrandbits(rng.choice([1, bits // 2, bits, 2 * bits]))
            assert gmp_powm(b, e, n) == pow(b, e, n)
        assert gmp_powm(0, 5, n) == 0
        assert gmp_powm(7, 0, n) == 1 % n


def test_inverse_of_the_montgomery_radix():
    """inv_pow2_mod (R^-1 mod n by Newton, used to check Montgomery
    products) against pow(R, -1, n)."""
    rng = random.Random(42)
    for bits in (24, 100, 5000):
        n = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
        R = 1 << (24 * (bits // 24 + 1))
        assert inv_pow2_mod(n, R) == pow(R, -1, n)
