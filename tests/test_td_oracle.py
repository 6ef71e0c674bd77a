"""CPU checks that pin the Phase-1 oracle before it is trusted as the GPU's
checker: against the reference's own trial_div_cpu.c (built from its
sources), against the committed golden bitmaps, against direct per-pair
evaluation, and against exact big-integer arithmetic."""
import json
import random

import numpy as np
import pytest

from _support import (GOLDEN, bit_get, numpy_primes, oracle_brute, oracle_check_pairs,
                      oracle_sieve, popcount, ref_td, ref_trial_div_cpu, survivors)

TEST_GPU_PRIMES = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
                   73, 79, 83, 89, 97]


def golden_cases():
    data = json.loads((GOLDEN / "td_ref_k100.json").read_text())
    p20k = numpy_primes(20000)
    for case in data["cases"]:
        primes = np.array(TEST_GPU_PRIMES, np.uint32) if case["primes"] == "test_gpu" else p20k
        words = np.array([int(case["bits"][i:i + 8], 16) for i in range(0, len(case["bits"]), 8)],
                         np.uint32)
        yield case, primes, words


def test_golden_fixture_matches_oracle():
    n = 0
    for case, primes, words in golden_cases():
        got = oracle_sieve(primes, case["a"], case["b"], case["c"], threads=1)
        assert (got == words).all(), case
        n += 1
    assert n == 12


@pytest.mark.skipif(ref_td() is None, reason="oracle/_ref not built (reference sources absent)")
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


@pytest.mark.skipif(ref_td() is None, reason="oracle/_ref not built (reference sources absent)")
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
