"""GPU parity of the batched Miller-Rabin stack (HIP library through the C
ABI): exact products and Montgomery products against Python integers,
strong-probable-prime residues and verdicts against GMP and textbook
arithmetic, and known primes / strong pseudoprimes."""
import random
import sys

import pytest

from _support import gmp_sprp, inv_pow2_mod, sprp_python

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


@pytest.mark.parametrize("bits", [64, 96, 1500, 30000, 166097, 332193, 384000])
def test_montmul_and_mulsmall(gpu, bits):
    """Montgomery products (a zero operand, N - 1 squared) and small-factor
    products against Python, up to the largest supported modulus (2^15
    transforms, where m N is a wrapped product).  96 and 384 000 bits fill
    whole 32-bit words and whole 24-bit limbs at once."""
    rng = random.Random(bits + 1)
    mods = [rand_odd(rng, bits) for _ in range(4)]
    ctx = gpu.MrContext(mods)
    say("montmul", bits, "bits, L", ctx.limbs, "logM", ctx.log_len)
    R = ctx.R
    a = [rng.randrange(n) for n in mods]
    b = [rng.randrange(n) for n in mods]
    a[1], b[1] = mods[1] - 1, mods[1] - 1
    a[2] = 0
    got = ctx.debug_montmul(a, b)
    for x, y, n, g in zip(a, b, mods, got):
        assert g == x * y * inv_pow2_mod(n, R) % n
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


def test_multiplicative_consistency_20k_digits(gpu):
    """Size-independent checks on candidate-form numbers of 20 000 digits:
    residues are multiplicative in the base, (ab)^d = a^d b^d mod N, and
    4^d = (2^d)^2 mod N."""
    rng = random.Random(20000)
    mods = []
    for _ in range(2):
        b = rng.randint(1000, 19000)
        c = rng.randint(100, b - 1)
        n = 10 ** 20000 - rng.randint(1, 4096) * 10 ** b - rng.randint(1, 4096) * 10 ** c - 1
        mods += [n, int(str(n)[::-1])]
    ctx = gpu.MrContext(mods)
    rows = []
    for base in (2, 3, 6, 4):      # every base on every candidate
        _, r = ctx.run([base])
        rows.append([x[0] for x in r])
        say("20k-digit consistency: base", base, "done")
    for i, n in enumerate(mods):
        r2, r3, r6, r4 = rows[0][i], rows[1][i], rows[2][i], rows[3][i]
        assert r6 == r2 * r3 % n
        assert r4 == r2 * r2 % n
    ctx.close()


def test_device_builder_matches_gmp_and_python(gpu):
    """N, its decimal reversal and the reference's rev formula, built on the
    device from (a, b, c, k, j), equal GMP's mpz construction (mpz_ui_pow_ui
    / mul_ui / sub, and mpz_set_str of the reversed string)."""
    from _support import gmp_build
    import json
    from _support import GOLDEN
    cfg = json.loads((GOLDEN / "configs.json").read_text())
    cases = [(5, 2, 1, 3, 5), (30, 15, 7, 12, 34), (31, 27, 5, 4096, 7), (60, 59, 1, 9, 1),
             (40, 3, 1, 4096, 4096), (25, 24, 10, 9, 9), (2000, 1000, 100, 4095, 77)]
    cases += [tuple(x) for x in cfg["3"]["candidates"][:6]]
    cands = []
    for (a, b, c, k, j) in cases:
        for which in (0, 1, 2):
            lk, lj = len(str(k)), len(str(j))
            if which == 2 and (a - b - lk < 0 or a - c - lj < 0):
                continue
            cands.append((a, b, c, k, j, which))
    ctx = gpu.MrContext(candidates=cands)
    got = ctx.moduli()
    for cd, g in zip(cands, got):
        a, b, c, k, j, which = cd
        assert g == gmp_build(a, b, c, k, j, which), cd
        if a <= 2000:
            n = 10 ** a - k * 10 ** b - j * 10 ** c - 1
            want = n if which == 0 else int(str(n)[::-1]) if which == 1 else g
            assert g == want
    ctx.close()


def test_even_and_tiny_moduli_are_decided(gpu):
    ctx = gpu.MrContext([2, 3, 1, 4, 10 ** 30 + 8, 7, 2 ** 89 - 1])
    alive, res = ctx.run([2, 3])
    assert alive == [True, True, False, False, False, True, True]
    ctx.close()


def test_bases_that_are_multiples_of_n(gpu):
    """N = 5 and 7 with bases 2, 3, 5, 7: the base equal to N tests nothing
    and the round passes (GMP mpz_probab_prime_p calls both prime); 25 and
    49 stay composite."""
    from _support import gmp_probab_prime, sprp_python
    nums = [5, 7, 11, 13, 25, 49, 35]
    ctx = gpu.MrContext(nums)
    alive, res = ctx.run([2, 3, 5, 7])
    for n, a in zip(nums, alive):
        want = all(sprp_python(n, b)[0] for b in (2, 3, 5, 7))
        assert a == want, n
        if n in (5, 7, 11, 13):
            assert a and gmp_probab_prime(n) >= 1, n
    ctx.close()
    v, _ = gpu.mr_test_batch([(1, 0, 0, 1, 1, 0)], [2, 3, 5, 7])   # N = 10 - 1 - 1 - 1 = 7
    assert v == [1]
    code, _ = gpu.miller_rabin_test(1, 0, 0, 1, 1)                # N = rev(N) = 7
    assert code == 2


@pytest.mark.parametrize("logm", [12, 13, 15])
def test_every_transform_length(gpu, logm):
    """The same moduli through forced transform lengths 2^12 .. 2^15: exact
    products, Montgomery products and full rounds against Python/GMP."""
    rng = random.Random(logm)
    mods = [rand_odd(rng, 1000), rand_odd(rng, 2500), 2 ** 1279 - 1, 3215031751]
    ctx = gpu.MrContext(mods, min_log_len=logm)
    assert ctx.log_len == logm
    a = [rng.randrange(n) for n in mods]
    b = [rng.randrange(n) for n in mods]
    assert ctx.debug_product(a, b) == [x * y for x, y in zip(a, b)]
    R = ctx.R
    assert ctx.debug_montmul(a, b) == [x * y * inv_pow2_mod(n, R) % n for x, y, n in zip(a, b, mods)]
    alive, res = ctx.run([2, 3, 5, 7])
    for i, n in enumerate(mods):
        exp_alive = True
        for bi, base in enumerate((2, 3, 5, 7)):
            ok, r, _ = sprp_python(n, base)
            if exp_alive:
                assert res[i][bi] == r
            exp_alive = exp_alive and ok
        assert alive[i] == exp_alive
    ctx.close()


@pytest.mark.parametrize("bits", [49120, 98272, 196576])
def test_largest_modulus_per_length(gpu, bits):
    """The largest modulus of each unsplit transform length (L = M/2 - 1
    limbs: 2^12, 2^13, 2^14; the bit lengths are whole 32-bit words).  There
    the high layout's top limbs sit in the last thread's run, which once hid
    the final subtraction's borrow, and the wrapped m N product of M/2
    points has the least room.  Montgomery and small-factor products
    against Python (the reductions of the exponentiation loop are the same
    code); at 2^12 also full base-2 rounds against GMP (GMP takes minutes
    per round at the larger sizes)."""
    rng = random.Random(bits)
    mods = [rand_odd(rng, bits) for _ in range(3)]
    mods[2] = (1 << bits) - 1 - 2 * rng.getrandbits(8)      # long runs of ones
    ctx = gpu.MrContext(mods)
    say("largest modulus", bits, "bits, L", ctx.limbs, "logM", ctx.log_len)
    assert 2 * ctx.limbs + 2 == 1 << ctx.log_len
    a = [rng.randrange(n) for n in mods]
    b = [rng.randrange(n) for n in mods]
    a[0], b[0] = mods[0] - 1, mods[0] - 1
    R = ctx.R
    assert ctx.debug_montmul(a, b) == [x * y * inv_pow2_mod(n, R) % n for x, y, n in zip(a, b, mods)]
    for s in (3, (1 << 24) - 1):
        assert ctx.debug_mulsmall(a, s) == [x * s % n for x, n in zip(a, mods)]
    if ctx.log_len == 12:
        alive, res = ctx.run([2])
        for i, n in enumerate(mods[:2]):
            ok2, r2, _ = gmp_sprp(n, 2)
            assert res[i][0] == r2 and alive[i] == ok2
    ctx.close()


def test_100k_digit_arithmetic(gpu):
    """At 100 000 digits (2^15 transforms): device-built candidates equal
    GMP, products and Montgomery products are exact."""
    from _support import gmp_build
    rng = random.Random(100000)
    cands = [(100000, 61234, 789, 4095, 17, 0), (100000, 61234, 789, 4095, 17, 1),
             (99990, 12345, 6789, 123, 4096, 0)]
    ctx = gpu.MrContext(candidates=cands)
    say("100k digits: L", ctx.limbs, "logM", ctx.log_len)
    assert ctx.log_len == 15
    mods = ctx.moduli()
    for cd, n in zip(cands, mods):
        assert n == gmp_build(*cd)
    a = [rng.randrange(n) for n in mods]
    b = [rng.randrange(n) for n in mods]
    a[0] = mods[0] - 1
    R = ctx.R
    assert ctx.debug_product(a, b) == [x * y for x, y in zip(a, b)]
    assert ctx.debug_montmul(a, b) == [x * y * inv_pow2_mod(n, R) % n for x, y, n in zip(a, b, mods)]
    ctx.close()


def test_mersenne_216091_closed_form(gpu):
    """M216091 (65 050 digits, 2^15 transforms): M - 1 = 2 d with d odd and
    3 a quadratic non-residue, so the base-3 residue 3^d is M - 1."""
    m = 2 ** 216091 - 1
    ctx = gpu.MrContext([m])
    say("M216091: L", ctx.limbs, "logM", ctx.log_len, "- one base-3 round")
    assert ctx.log_len == 15
    alive, r3 = ctx.run([3])
    assert r3[0][0] == m - 1 and alive[0]
    ctx.close()
