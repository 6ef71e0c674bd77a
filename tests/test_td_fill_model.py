"""CPU model of the row kernel's word-pattern fills (the small-prime Phase 0
of the reference, sieve_table.c:93-243 / sieve.c:4-31, done per row).

For the fill primes 3, 7, 11, ..., 31 the row kernel does not mark hits one
by one: in row k, prime p eliminates exactly the j of one class mod p (the
N condition) and the j whose padded reversal R(j) lies in one class mod p
(the merged rev condition), so the row's bits are ORs of precomputed class
masks of 32-bit words, funnel-shifted to the row's bit offset (rows start
mid-word when 32 does not divide J).  The model below builds the class-mask
table with the kernel's layout, derives each row's classes from Python
modular arithmetic, assembles the bitmap word by word with the same shifts
and compares it with the oracle's restatement of trial_div_cpu.c over the
fill primes alone; a planted off-by-one class fails it."""
import random

import numpy as np
import pytest

from _support import oracle_sieve

FILL_P = [3, 7, 11, 13, 17, 19, 23, 29, 31]
FILL_OFF = [0, 3, 10, 21, 34, 51, 70, 93, 122]
CLASSES = 153


def ndig(x):
    return len(str(x))


def padded_rev(j, dj):
    return int(str(j)[::-1]) * 10 ** (dj - ndig(j))


def fill_table(J, dj):
    """Class masks, word (kind * CLASSES + off[f] + r) * qw + q."""
    qw = (J + 31) // 32
    tab = np.zeros(2 * CLASSES * qw, dtype=np.uint64)
    for f, p in enumerate(FILL_P):
        for j in range(1, J + 1):
            q, i = divmod(j - 1, 32)
            tab[(FILL_OFF[f] + j % p) * qw + q] |= np.uint64(1 << i)
            tab[(CLASSES + FILL_OFF[f] + padded_rev(j, dj) % p) * qw + q] |= np.uint64(1 << i)
    return tab.astype(np.uint32), qw


def row_classes(p, a, b, c, k, dj, skew=0):
    """(N class, rev class or None) of row k: the j with p | N are those of
    class (10^a - 1 - k 10^b) / 10^c mod p, the j with p | rev(N) those whose
    R(j) is of class (10^a - 1 - rev(k) 10^(a-b-lk)) / 10^(a-c-dj)."""
    inv = lambda x: pow(x, p - 2, p)
    cn = (pow(10, a, p) - 1 - k * pow(10, b, p)) * inv(pow(10, c, p)) % p
    lk = ndig(k)
    cr = None
    if a - b - lk >= 0:
        rk = int(str(k)[::-1])
        cr = (pow(10, a, p) - 1 - rk * pow(10, a - b - lk, p)) * inv(pow(10, a - c - dj, p)) % p
    return (cn + skew) % p, cr


def model_fill(a, b, c, K, J, primes, skew=0):
    dj = ndig(J)
    assert a - c >= dj, "fills serve the merged rev path only"
    tab, qw = fill_table(J, dj)
    nwords = (K * J + 31) // 32
    bits = np.zeros(nwords + 1, dtype=np.uint64)
    for k in range(1, K + 1):
        for q in range(qw):
            m = 0
            for p in primes:
                f = FILL_P.index(p)
                cn, cr = row_classes(p, a, b, c, k, dj, skew)
                m |= int(tab[(FILL_OFF[f] + cn) * qw + q])
                if cr is not None:
                    m |= int(tab[(CLASSES + FILL_OFF[f] + cr) * qw + q])
            if not m:
                continue
            b0 = (k - 1) * J + 32 * q
            sh, w = b0 % 32, b0 // 32
            bits[w] |= np.uint64((m << sh) & 0xFFFFFFFF)
            if sh:
                bits[w + 1] |= np.uint64(m >> (32 - sh))
    assert bits[nwords] == 0, "no mark beyond the grid"
    return bits[:nwords].astype(np.uint32)


CASES = [(50000, 25000, 100, 20, 4096), (1000, 600, 100, 37, 100), (300, 150, 7, 25, 37),
         (40, 12, 3, 9, 9), (60, 58, 2, 33, 65), (500, 101, 100, 12, 999)]


@pytest.mark.parametrize("a,b,c,K,J", CASES)
def test_fill_model_matches_oracle(a, b, c, K, J):
    primes = np.array(FILL_P, dtype=np.uint32)
    got = model_fill(a, b, c, K, J, FILL_P)
    ref = oracle_sieve(primes, a, b, c, K, J, threads=1)
    assert (got == ref).all()


def test_fill_model_random_subsets():
    rng = random.Random(2026)
    for _ in range(6):
        J = rng.choice([1, 5, 31, 32, 33, 64, 100, 257, 1000, 1500])
        dj = ndig(J)
        c = rng.randint(0, 50)
        a = c + dj + rng.randint(0, 200)
        b = rng.randint(0, a)
        K = rng.randint(1, 40)
        sub = sorted(rng.sample(FILL_P, rng.randint(1, 9)))
        got = model_fill(a, b, c, K, J, sub)
        ref = oracle_sieve(np.array(sub, dtype=np.uint32), a, b, c, K, J, threads=1)
        assert (got == ref).all(), (a, b, c, K, J, sub)


def test_planted_class_slip_fails():
    a, b, c, K, J = 1000, 600, 100, 37, 100
    ref = oracle_sieve(np.array(FILL_P, dtype=np.uint32), a, b, c, K, J, threads=1)
    bad = model_fill(a, b, c, K, J, FILL_P, skew=1)
    assert not (bad == ref).all()
