"""GPU parity of the batched Miller-Rabin stack (HIP library through the C
ABI): exact products and Montgomery products against Python integers,
strong-probable-prime residues and verdicts against GMP and textbook
arithmetic, and known primes / strong pseudoprimes."""
import random
import sys

import pytest

from _support import gmp_sprp, sprp_python

pytestmark = pytest.mark.gpu

if hasattr(sys, "set_int_max_str_digits"):
    sys.set_int_max_str_digits(0)      # decimal reversal of 100 000-digit numbers


def say(*a):
    print("[mr-gpu]", *a, flush=True)


def rand_odd(rng, bits):
    return rng.getrandbits(bits) | 1 | (1 << (bits - 1))


@pytest.mark.parametrize("bits", [100, 2000, 20000, 100000, 160000])
def test_product_exact(gpu, bits):
    rng = random.Random(bits)
    mods = [rand_odd(rng, bits) for _ in range(3)]
    ctx = gpu.MrContext(mods)
    say("product", bits, "bits, L", ctx.limbs, "logM", ctx.log_len)
    a = [rng.getrandbits(bits) for _ in mods]
    b = [rng.getrandbits(bits) for _ in mods]
    a[0] = (1 << bits) - 1          # all-ones operands: longest carry chains
    b[0] = (1 << bits) - 1
    got = ctx.debug_product(a, b)
    for x, y, g in zip(a, b, got):
        assert g == x * y
    ctx.close()


@pytest.mark.parametrize("bits", [64, 1500, 30000, 166097])
def test_montmul_and_mulsmall(gpu, bits):
    rng = random.Random(bits + 1)
    mods = [rand_odd(rng, bits) for _ in range(4)]
    ctx = gpu.MrContext(mods)
    R = ctx.R
    a = [rng.randrange(n) for n in mods]
    b = [rng.randrange(n) for n in mods]
    a[1], b[1] = mods[1] - 1, mods[1] - 1
    got = ctx.debug_montmul(a, b)
    for x, y, n, g in zip(a, b, mods, got):
        assert g == x * y * pow(R, -1, n) % n
    for s in (2, 3, 7, 53, (1 << 24) - 1):
        got = ctx.debug_mulsmall(a, s)
        for x, n, g in zip(a, mods, got):
            assert g == x * s % n
    ctx.close()


def test_known_primes_and_pseudoprimes(gpu):
    """Mersenne primes (s = 1), strong pseudoprimes that separate per-base
    MR from BPSW, composites, and primes with s = 2, 3, 5."""
    from sympy import nextprime
    cases = [2 ** p - 1 for p in (61, 89, 107, 127, 521, 607, 1279)]
    cases += [2047, 3215031751, 3215031751 * 3 + 2, 25326001, 341550071728321]
    for k in (2, 3, 5):              # primes with N - 1 = 2^s d, s = k
        x = nextprime(2 ** 300)
        while (x - 1) % (1 << k) != 0 or ((x - 1) >> k) % 2 == 0:
            x = nextprime(x)
        cases.append(int(x))
    cases += [rand_odd(random.Random(5), 700) for _ in range(3)]
    bases = [2, 3, 5, 7, 11, 13]
    ctx = gpu.MrContext(cases)
    alive, res = ctx.run(bases)
    for i, n in enumerate(cases):
        exp_alive = True
        for bi, base in enumerate(bases):
            ok, r, s = sprp_python(n, base)
            assert ctx.s[i] == s
            if exp_alive:
                assert res[i][bi] == r, (n, base)
            else:
                assert res[i][bi] == 0
            exp_alive = exp_alive and ok
        assert alive[i] == exp_alive, n
    # strong pseudoprime 3215031751 passes 2, 3, 5, 7 and fails 11
    i = cases.index(3215031751)
    alive4, _ = ctx.run([2, 3, 5, 7], residues=False)
    assert alive4[i] and not alive[i]
    ctx.close()


def test_large_mersenne_closed_form(gpu):
    """M = 2^p - 1 prime, N - 1 = 2 d: a^d = (a / M) (Legendre symbol), so
    2^d = 1 (M = 7 mod 8) and 3^d = M - 1 (quadratic reciprocity)."""
    ps = [9689, 21701, 44497, 86243, 110503, 132049]
    mods = [2 ** p - 1 for p in ps]
    ctx = gpu.MrContext(mods)
    say("mersenne", ps, "L", ctx.limbs, "logM", ctx.log_len)
    alive, res = ctx.run([2, 3])
    for i, m in enumerate(mods):
        assert alive[i]
        assert res[i][0] == 1
        assert res[i][1] == m - 1
    ctx.close()


def test_residues_vs_gmp_medium(gpu):
    """Candidate-form numbers of 1 000 - 12 000 digits, N and rev(N): every
    residue equals GMP's mpz_powm."""
    rng = random.Random(77)
    mods = []
    for digits in (1000, 3000, 6000, 12000):
        a = digits
        b = rng.randint(100, a - 10)
        c = rng.randint(20, b - 5)
        k, j = rng.randint(1, 4096), rng.randint(1, 4096)
        n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
        mods += [n, int(str(n)[::-1])]
    ctx = gpu.MrContext(mods)
    bases = [2, 3, 5, 7]
    alive, res = ctx.run(bases)
    for i, n in enumerate(mods):
        exp_alive = True
        for bi, base in enumerate(bases):
            v, r, s = gmp_sprp(n, base)
            if exp_alive:
                assert res[i][bi] == r
            exp_alive = exp_alive and v == 1
        assert alive[i] == exp_alive
    ctx.close()


def test_multiplicative_consistency_50k_digits(gpu):
    """Size-independent checks at the 50 000-digit working size: residues
    are multiplicative in the base, (ab)^d = a^d b^d mod N, and
    4^d = (2^d)^2 mod N."""
    rng = random.Random(50000)
    mods = []
    for _ in range(2):
        b = rng.randint(1000, 49000)
        c = rng.randint(100, b - 1)
        n = 10 ** 50000 - rng.randint(1, 4096) * 10 ** b - rng.randint(1, 4096) * 10 ** c - 1
        mods += [n, int(str(n)[::-1])]
    ctx = gpu.MrContext(mods)
    say("50k-digit consistency: L", ctx.limbs, "logM", ctx.log_len)
    bases = [2, 3, 6, 4]
    # run all four bases on every candidate regardless of earlier failures
    rows = []
    for base in bases:
        _, r = ctx.run([base])
        rows.append([x[0] for x in r])
    for i, n in enumerate(mods):
        r2, r3, r6, r4 = rows[0][i], rows[1][i], rows[2][i], rows[3][i]
        assert r6 == r2 * r3 % n
        assert r4 == r2 * r2 % n
    ctx.close()
