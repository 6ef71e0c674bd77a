"""miller_rabin_test (include/miller_rabin.h) on the GPU against the
reference's own Phase-2 checks (tests/test_phase2.c:46-71: N prime by trial
division) and against GMP's mpz_probab_prime_p on N and the reference rev
formula (the oracle's restatement of miller_rabin.c:32-118)."""
import pytest

from _support import gmp_build, gmp_probab_prime

pytestmark = pytest.mark.gpu


def is_prime_td(n: int) -> bool:
    if n < 2:
        return False
    d = 2
    while d * d <= n:
        if n % d == 0:
            return False
        d += 1
    return True


def reference_code(a, b, c, k, j) -> int:
    """miller_rabin.c:76-118 with GMP: 0 / 1 / 2."""
    if gmp_probab_prime(gmp_build(a, b, c, k, j, 0)) == 0:
        return 0
    return 2 if gmp_probab_prime(gmp_build(a, b, c, k, j, 2)) > 0 else 1


def test_phase2_small_vs_trial_division(gpu):
    # test_phase2.c:51-69: N = 10^4 - 100k - 10j - 1, k, j in 1..5
    for k in range(1, 6):
        for j in range(1, 6):
            n = 10000 - 100 * k - 10 * j - 1
            code, _ = gpu.miller_rabin_test(4, 2, 1, k, j)
            assert code in (0, 1, 2)
            assert (code >= 1) == is_prime_td(n), (k, j, n, code)


def test_codes_match_reference_semantics(gpu):
    # test_phase2.c:37-44 grid (a=6, b=3, c=2) and (5, 2, 1, 1, 1), (4, 2, 1, 3, 5)
    cases = [(6, 3, 2, k, j) for k in range(1, 11) for j in range(1, 11)]
    cases += [(5, 2, 1, 1, 1), (4, 2, 1, 3, 5), (9, 4, 1, 7, 3), (12, 7, 3, 11, 9)]
    seen = set()
    for cs in cases:
        code, t = gpu.miller_rabin_test(*cs)
        assert code == reference_code(*cs), cs
        seen.add(code)
        if code == 0:
            assert t.build_revn == 0.0 and t.mr_revn == 0.0
    assert seen == {0, 1, 2}


def test_codes_at_1000_digits(gpu):
    # (k, j) of the triple (1000, 500, 100) where GMP finds N prime (code 1),
    # mixed with composites
    cases = [(32, 5), (47, 57), (54, 26), (69, 14), (1, 7), (2, 8), (3, 9)]
    codes = []
    for k, j in cases:
        cs = (1000, 500, 100, k, j)
        code, _ = gpu.miller_rabin_test(*cs)
        assert code == reference_code(*cs), cs
        codes.append(code)
    assert codes == [1, 1, 1, 1, 0, 0, 0]


def test_invalid_candidates_fail_loudly(gpu):
    assert gpu.miller_rabin_test(4, 2, 1, 0, 1)[0] == -1        # k < 1
    assert gpu.miller_rabin_test(3, 5, 1, 1, 1)[0] == -1        # N < 0


def _limbs16(n, n_limbs):
    return [(n >> (16 * i)) & 0xFFFF for i in range(n_limbs)]


def test_cpp_runner_surface(gpu):
    """batch_mod_ctx.hpp driven like bench_mr_gpu.cu:60-64 (16-bit limbs in
    uint64, DEFAULT_WITNESSES): verdicts equal GMP's per-witness rounds."""
    import subprocess
    from _support import PKG, gmp_sprp
    exe = PKG / "lib" / "mr_cpp_check"
    assert exe.exists(), "mr_cpp_check not built"
    nums = [2 ** 521 - 1, 2 ** 607 - 1, 2047 * (2 ** 400 + 1), 3215031751,
            gmp_build(1000, 500, 100, 32, 5, 0), gmp_build(1000, 500, 100, 1, 7, 0),
            gmp_build(1000, 500, 100, 54, 26, 2)]
    n_limbs = max((n.bit_length() + 15) // 16 for n in nums)
    text = f"{n_limbs} {len(nums)}\n" + " ".join(str(x) for n in nums for x in _limbs16(n, n_limbs))
    r = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = [int(l.split()[1]) for l in r.stdout.splitlines() if l and l[0].isdigit()]
    witnesses = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53]
    exp = [int(all(gmp_sprp(n, w)[0] for w in witnesses)) for n in nums]
    assert got == exp
    assert exp[3] == 0          # 3215031751 fails at 11


def _cpp(mode, text):
    import subprocess
    from _support import PKG
    exe = PKG / "lib" / "mr_cpp_check"
    assert exe.exists(), "mr_cpp_check not built"
    r = subprocess.run([str(exe), mode], input=text, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cpp_runner_arithmetic(gpu):
    """BatchModCtx's device-pointer arithmetic (batch_mod_ctx.cuh:141-155) from a
    compiled C++ caller, as correctness_tests.cuh:217-218,318,353,387 use it:
    to_residue_batch, modmul_batch, modsq_batch, check_passed and
    from_residue_batch give x y mod N and x^2 mod N (Python integers), and
    check_passed accepts the working forms of 1 and N - 1 only."""
    import random
    rng = random.Random(44)
    nums = [2 ** 521 - 1, 2 ** 1279 - 1, gmp_build(1000, 500, 100, 32, 5, 0),
            gmp_build(3000, 1200, 100, 77, 41, 1), gmp_build(6000, 2500, 300, 4001, 17, 0)]
    nums = [n | 1 for n in nums]
    n_limbs = max((n.bit_length() + 15) // 16 for n in nums)
    xs = [rng.randrange(2, n) for n in nums]
    xs[1] = nums[1] - 1                       # N - 1 passes check_passed
    xs[2] = 1                                 # and so does 1
    ys = [rng.randrange(1, n) for n in nums]
    text = f"{n_limbs} {len(nums)}\n" + " ".join(
        str(v) for group in (nums, xs, ys) for n in group for v in _limbs16(n, n_limbs))
    out = _cpp("arith", text)
    rows = [l.split() for l in out.splitlines() if l and l[0].isdigit()]
    assert len(rows) == len(nums)
    for (i, xy, sq, p1, px), n, x, y in zip(rows, nums, xs, ys):
        assert int(xy, 16) == x * y % n, i
        assert int(sq, 16) == x * x % n, i
        assert p1 == "1"
        assert px == ("1" if x in (1, n - 1) else "0"), i


def test_cpp_runner_rejects_wrong_s_and_d(gpu):
    """gpu_miller_rabin with an s or an exp_all that are not those of N throws
    (the reference would silently test another exponent)."""
    nums = [2 ** 521 - 1, gmp_build(1000, 500, 100, 32, 5, 0)]
    n_limbs = max((n.bit_length() + 15) // 16 for n in nums)
    text = f"{n_limbs} {len(nums)}\n" + " ".join(str(x) for n in nums for x in _limbs16(n, n_limbs))
    assert _cpp("badargs", text).strip() == "ok"
