"""CPU checks that pin the Phase-1 oracle before it is trusted as the GPU's
checker: against the reference's own trial_div_cpu.c (built from its
sources), against the committed golden bitmaps, against direct per-pair
evaluation, and against exact big-integer arithmetic."""
import json
import random

import numpy as np
import pytest

from _support import (GOLDEN, TEST_GPU_PRIMES, bit_get, golden_cases, have_ref_td,
                      numpy_primes, oracle_brute, oracle_check_pairs, oracle_sieve, popcount,
                      ref_trial_div_cpu, survivors)

def test_golden_fixture_matches_oracle():
    n = 0
    for case, primes, words in golden_cases():
        got = oracle_sieve(primes, case["a"], case["b"], case["c"], threads=1)
        assert (got == words).all(), case
        n += 1
    assert n == 12


@pytest.mark.skipif(not have_ref_td(), reason="oracle/_ref not built (reference sources absent)")
def test_oracle_equals_reference_cpu_random_triples():
    rng = random.Random(0x5EED)
    primes = numpy_primes(30000)
    for _ in range(40):
        a = rng.randint(5, 100000)
        b = rng.randint(0, a)
        c = rng.randint(0, b)
        ref = ref_trial_div_cpu(primes, a, b, c)
        got = oracle_sieve(primes, a, b, c, threads=1)
        assert (ref == got).all(), (a, b, c)


@pytest.mark.skipif(not have_ref_td(), reason="oracle/_ref not built (reference sources absent)")
def test_oracle_equals_reference_cpu_edge_exponents():
    # a-b-l and a-c-l crossing zero for l = 1..3; 2 and 5 in the list.
    primes = numpy_primes(5000)
    for d in range(0, 6):
        for e in range(0, 6):
            a = 200
            b = a - d
            c = max(0, a - e) if a - e < b else b - 1
            ref = ref_trial_div_cpu(primes, a, b, c)
            got = oracle_sieve(primes, a, b, c, threads=1)
            assert (ref == got).all(), (a, b, c)


@pytest.mark.parametrize("K,J", [(100, 100), (150, 130), (1200, 37), (37, 1200), (9, 11), (1, 1)])
def test_sieve_equals_brute(K, J):
    primes = numpy_primes(400)
    rng = random.Random(K * 10000 + J)
    for _ in range(3):
        a = rng.randint(10, 5000)
        b = rng.randint(1, a - 1)
        c = rng.randint(0, b - 1)
        s = oracle_sieve(primes, a, b, c, K, J, threads=1)
        t = oracle_brute(primes, a, b, c, K, J)
        assert (s == t).all(), (K, J, a, b, c)


def test_brute_equals_bigint_divisibility():
    """Tiny a: build N and the rev formula as Python integers and test
    divisibility directly; pins the residue arithmetic of both routines."""
    primes = [int(p) for p in numpy_primes(120)]
    for (a, b, c, K, J) in [(12, 7, 3, 60, 45), (9, 5, 1, 30, 120), (14, 10, 2, 110, 20)]:
        bits = oracle_brute(np.array(primes, np.uint32), a, b, c, K, J)
        for k in range(1, K + 1):
            lk = len(str(k))
            rk = int(str(k)[::-1])
            for j in range(1, J + 1):
                lj = len(str(j))
                rj = int(str(j)[::-1])
                n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
                hit = False
                for p in primes:
                    if pow(10, c, p) != 0 and n % p == 0:
                        hit = True
                        break
                    eb, ec = a - b - lk, a - c - lj
                    if eb >= 0 and ec >= 0 and pow(10, eb, p) and pow(10, ec, p):
                        r = 10 ** a - rk * 10 ** eb - rj * 10 ** ec - 1
                        if r % p == 0:
                            hit = True
                            break
                assert bit_get(bits, J, k, j) == int(hit), (a, b, c, k, j)


def test_check_pairs_agrees_with_sieve():
    primes = numpy_primes(3000)
    a, b, c, K, J = 777, 500, 101, 300, 280
    bits = oracle_sieve(primes, a, b, c, K, J)
    rng = np.random.default_rng(7)
    ks = rng.integers(1, K + 1, 500)
    js = rng.integers(1, J + 1, 500)
    flags = oracle_check_pairs(primes, a, b, c, ks, js)
    for k, j, f in zip(ks, js, flags):
        assert bit_get(bits, J, int(k), int(j)) == int(f)


def test_mt_equals_single_thread():
    primes = numpy_primes(50000)
    a, b, c = 50000, 25000, 100
    s1 = oracle_sieve(primes, a, b, c, 256, 256, threads=1)
    s8 = oracle_sieve(primes, a, b, c, 256, 256, threads=8)
    assert (s1 == s8).all()


def test_survivor_order_and_k100_gate():
    # The reference's all-eliminated gate (sieve.c:24-31) reads 313 words,
    # the last with 16 valid bits; survivors are lexicographic (worker.c:45-54).
    primes = numpy_primes(20000)
    bits = oracle_sieve(primes, 50001, 33334, 2749)
    surv = survivors(bits, 100, 100)
    assert surv == sorted(surv)
    assert len(surv) == 10000 - popcount(bits)
    assert (bits[312] >> 16) == 0


def test_pipeline_fixture_survivors_match_oracle():
    """tests/golden/pipeline_config1.json's survivors are the oracle's for
    workload 1 (the fixture the GPU pipeline test compares against)."""
    import hashlib
    import json
    from _support import ROOT, numpy_primes, oracle_sieve, survivors
    fx = json.loads((ROOT / "tests" / "golden" / "pipeline_config1.json").read_text())
    bits = oracle_sieve(numpy_primes(fx["prime_limit"]), fx["a"], fx["b"], fx["c"], fx["K"], fx["J"])
    surv = survivors(bits, fx["K"], fx["J"])
    assert len(surv) == fx["survivor_count"]
    packed = np.array([(k << 16) | j for k, j in surv], dtype="<u4")
    assert hashlib.sha256(packed.tobytes()).hexdigest() == fx["survivors_sha256"]


def test_vanishing_power_of_ten_groups_follow_trial_div_cpu():
    """p = 2 and 5 divide every power of ten above 10^0.  When a k's rev
    exponent a-b-lk is 0 and a j digit group's exponent a-c-lj is positive,
    rev(N) mod p does not depend on j inside that group, so exact
    divisibility eliminates the whole group when it eliminates one pair.
    The reference's GPU kernel marks those pairs (its flat scan,
    trial_div.cu:97-115); its CPU sieve trial_div_cpu.c, the contract for
    Phase-1 parity, solves each group through the inverse of the group's
    power of ten and skips a group whose power vanishes mod p
    (trial_div_cpu.c:62-63,73,82), so it marks none of them.  This build
    follows trial_div_cpu.c; the test pins the divergence: at (a, b, c) =
    (4, 3, 1), p = 2, exact divisibility of the rev formula eliminates 499
    pairs and the CPU sieve 4 (the j = 100 group, whose exponent is 0).
    No workload reaches such triples (c = 100, b > c)."""
    a, b, c = 4, 3, 1
    exact = set()
    for k in range(1, 101):
        for j in range(1, 101):
            lk, lj = len(str(k)), len(str(j))
            if a - b - lk < 0 or a - c - lj < 0:
                continue
            rev = 10 ** a - int(str(k)[::-1]) * 10 ** (a - b - lk) - \
                int(str(j)[::-1]) * 10 ** (a - c - lj) - 1
            n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
            assert n % 2 == 1
            if rev % 2 == 0:
                exact.add((k, j))
    assert len(exact) == 499
    got = oracle_sieve(np.array([2], np.uint32), a, b, c, threads=1)
    marked = {(k, j) for k in range(1, 101) for j in range(1, 101) if bit_get(got, 100, k, j)}
    assert marked == {(k, 100) for k in (2, 4, 6, 8)}
    assert marked < exact
    if have_ref_td():
        assert (ref_trial_div_cpu(np.array([2], np.uint32), a, b, c) == got).all()
