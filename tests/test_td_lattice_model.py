"""CPU model of the trial-division lattice kernel.

The kernel's per-prime routine (td_lattice_prime in
prime-finder_amd/csrc/td_lattice.h: the N condition and each digit class of
the rev condition as "first t with (c0 + t sig) mod p < T", found by a
continued-fraction descent and repeated from the next t) is compiled for the
host with g++ (tests/model/td_lattice_model.cpp) and compared bit for bit
with the oracle's restatement of trial_div_cpu.c on primes above 10^4:
random triples with exponent edges, grids up to 9999, primes next to 2^16,
2^21, 2^24 and 2^31.  A composite list entry must end (no hang)."""
import ctypes as C
import random
import subprocess

import numpy as np
import pytest

from _support import ROOT, nwords, numpy_primes, oracle_sieve

u32p = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    out = tmp_path_factory.mktemp("latmodel") / "td_lattice_model.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall",
                    f"-I{ROOT / 'prime-finder_amd' / 'csrc'}",
                    str(ROOT / "tests" / "model" / "td_lattice_model.cpp"), "-o", str(out)],
                   check=True)
    L = C.CDLL(str(out))
    L.lattice_sieve.argtypes = [u32p, C.c_uint32, C.c_longlong, C.c_longlong, C.c_longlong,
                                C.c_uint32, C.c_uint32, u32p]
    L.lattice_sieve.restype = C.c_int
    L.gaps_rule_mismatches.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.gaps_rule_mismatches.restype = C.c_int
    L.first_hit_mismatches.argtypes = [C.c_uint32] * 6
    L.first_hit_mismatches.restype = C.c_int
    L.top_window_mismatch.restype = C.c_uint32
    L.lazy_pow_mismatches.argtypes = [C.c_uint32] * 3
    L.lazy_pow_mismatches.restype = C.c_int
    return L


def run_model(L, primes, a, b, c, K, J):
    p = np.ascontiguousarray(primes, dtype=np.uint32)
    bits = np.zeros(nwords(K, J), np.uint32)
    assert L.lattice_sieve(p.ctypes.data_as(u32p), len(p), a, b, c, K, J,
                           bits.ctypes.data_as(u32p)) == 0
    return bits


def _isprime(n):
    if n < 2 or n % 2 == 0:
        return n == 2
    d = 3
    while d * d <= n:
        if n % d == 0:
            return False
        d += 2
    return True


def _primes_near(x, count, rng):
    out, v = [], x - rng.randrange(0, 5000)
    while len(out) < count:
        if _isprime(v) and 10 ** 4 < v < 2 ** 31:
            out.append(v)
        v += 1
    return out


@pytest.fixture(scope="module")
def big_primes():
    rng = random.Random(11)
    allp = numpy_primes(3_000_000)
    ps = list(allp[allp > 10 ** 4][::997])
    for x in (1 << 16, 1 << 21, 1 << 24, 10 ** 8, (1 << 31) - 200):
        ps += _primes_near(x, 6, rng)
    ps += [2147483647, 2147483629, 99999989, 16777259, 2097143, 2097169]
    return np.array(sorted(set(ps)), np.uint32)


@pytest.mark.parametrize("K,J", [(4096, 4096), (100, 100), (9999, 37), (37, 9999), (1000, 999),
                                 (150, 130), (1, 1), (9, 10)])
def test_model_equals_oracle(model, big_primes, K, J):
    rng = random.Random(K * 10007 + J)
    dj = len(str(J))
    for _ in range(4):
        a = rng.choice([rng.randrange(dj + 1, 200), rng.randrange(50000, 100001)])
        c = rng.randrange(0, a - dj + 1)
        b = rng.choice([rng.randrange(0, a + 4), c + 1, a - 1, a - 4, a])
        got = run_model(model, big_primes, a, b, c, K, J)
        want = oracle_sieve(big_primes, a, b, c, K, J)
        assert (got == want).all(), (a, b, c, K, J)


def test_model_workload2_triple_sample(model):
    """A workload triple at K = J = 4096 over primes of every size class of
    the lane kernels (the lattice path takes p >= 2^21 on the device)."""
    import json
    from _support import GOLDEN
    cfg = json.loads((GOLDEN / "configs.json").read_text())["2"]
    rng = np.random.default_rng(3)
    allp = numpy_primes(10 ** 8)
    sample = np.sort(rng.choice(allp[allp > 10 ** 4], 600, replace=False)).astype(np.uint32)
    for (a, b, c) in cfg["triples"][:3]:
        got = run_model(model, sample, a, b, c, 4096, 4096)
        assert (got == oracle_sieve(sample, a, b, c, 4096, 4096)).all(), (a, b, c)


def test_composite_entries_terminate(model):
    comps = np.array([3 * 5 * 7 * 11 * 13 * 17, 1000001, 2147483647 * 0 + 2147483645,
                      10 ** 8 + 1, 65535 * 3], np.uint32)
    run_model(model, comps, 60000, 30000, 100, 4096, 4096)


def test_return_time_rule_brute_force(model):
    """The three-gap rule the kernel steps by after a hit (td_gaps): from
    every v of the window [0, T), the first return of v + n sig (mod p) and
    where it lands, against brute force, on random (p, sig, T) with small p,
    windows from 2 to p - 1, steps next to 1 and p - 1."""
    rng = random.Random(3)
    primes = [p for p in range(11, 4000) if _isprime(p)]
    cases = 0
    for _ in range(300):
        p = rng.choice(primes)
        sig = rng.choice([1, 2, p - 1, p - 2, rng.randrange(1, p)])
        T = rng.choice([2, 3, p - 1, rng.randrange(2, p)])
        assert model.gaps_rule_mismatches(p, sig, T) == 0, (p, sig, T)
        cases += 1
    assert cases == 300


def test_windowed_first_hit_brute_force(model):
    """The first-hit query with its window (td_first_hit: a query that only
    needs t < W stops once the descent shows the first hit is W or more):
    exact first hit and residue whenever it lies below W, and never a t
    below W otherwise, against brute force on random (p, sig, T, W) with
    small p, steps next to 1 and p - 1, windows from 1 to beyond p."""
    rng = random.Random(17)
    primes = [p for p in range(10007, 40000) if _isprime(p)]
    for _ in range(300):
        p = rng.choice(primes)
        sig = rng.choice([1, 2, p - 1, p - 2, rng.randrange(1, p)])
        T = rng.choice([1, 2, p - 1, rng.randrange(1, 100), rng.randrange(1, p)])
        W = rng.choice([1, 2, 10, 100, 1000, p, rng.randrange(1, p)])
        c0 = rng.randrange(0, p)
        assert model.first_hit_mismatches(p, sig, T, W, c0, 40) == 0, (p, sig, T, W, c0)


def test_top_class_window_brute_force(model):
    """The top rev class searches t = rev(k) only up to the largest reversal
    of a k <= K (td_top_window: K = 100 leaves only k = 100, t = 1): checked
    against brute force for every K from 1 to 9999."""
    assert model.top_window_mismatch() == 0


def test_lazy_powers_equal_reduced_powers(model):
    """The setup's powers with lazily reduced products (no final
    subtraction, operands below 2p, for p < 2^30) equal the fully reduced
    chain and plain 64-bit square-and-multiply, on odd moduli up to just
    below 2^30 (primes and composites: the bound needs only p odd and
    4 p^2 <= p 2^32) and exponents up to 2^20, zero included."""
    for p in (3, 10007, 65537, 999999937, (1 << 30) - 35, (1 << 30) - 1, 1000001):
        assert model.lazy_pow_mismatches(p, p & 0xFFFF, 400) == 0, p
