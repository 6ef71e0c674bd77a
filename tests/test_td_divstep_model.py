"""CPU model of the remainder step of the lattice kernel's first-hit descent.

Each continued-fraction level needs L mod A and M mod A (td_mod_est in
prime-finder_amd/csrc/td_lattice.h).  The kernel estimates the quotient
from the binary32 reciprocal of A scaled by 1 - 2^-20 and truncates: the
rounding of x, of A, of the reciprocal (v_rcp_f32 is accurate to 1 ulp,
the host's division to half an ulp) and of the product stays below
5 * 2^-24 relative, so the biased estimate never exceeds floor(x / A) and,
while it is below 2^18, is at most one less; one conditional subtraction
then gives x mod A exactly, and one conditional increment the quotient
(the return-time computation's runs, td_div_est).  Larger estimates take
the exact division.

Replayed here in numpy binary32 arithmetic with the reciprocal rounded to
nearest and perturbed by one ulp either way, on random and edge operands
(x at and next to multiples of A, A next to powers of two, x next to
2^31).  An unbiased estimate breaks the bound, which shows the bias is
what makes the step exact."""
import numpy as np

BIAS = np.float32(1.0 - 2.0 ** -20)


def operands(rng, n):
    A = rng.integers(1, 2 ** 31, n, dtype=np.int64)
    small = rng.integers(1, 1 << 12, n // 4, dtype=np.int64)
    p2 = (np.int64(1) << rng.integers(0, 31, n // 4)) + rng.integers(-2, 3, n // 4)
    A = np.concatenate([A, small, np.clip(p2, 1, 2 ** 31 - 1)])
    m = len(A)
    x = rng.integers(0, 2 ** 31, m, dtype=np.int64)
    # x at and next to multiples of A, and next to 2^31
    k = rng.integers(0, 2 ** 20, m, dtype=np.int64)
    xk = k * A + rng.integers(-1, 2, m)
    top = (2 ** 31 - 1) - rng.integers(0, 3, m)
    x = np.concatenate([x, np.clip(xk, 0, 2 ** 31 - 1), top])
    A = np.concatenate([A, A, A])
    return x.astype(np.uint64), A.astype(np.uint64)


def estimate(x, A, ulp_shift, bias=True):
    fA = A.astype(np.float32)
    r = (np.float32(1.0) / fA).astype(np.float32)
    if ulp_shift < 0:
        r = np.nextafter(r, np.float32(0)).astype(np.float32)
    elif ulp_shift > 0:
        r = np.nextafter(r, np.float32(np.inf)).astype(np.float32)
    rA = (r * BIAS).astype(np.float32) if bias else r
    prod = (x.astype(np.float32) * rA).astype(np.float32)
    return np.trunc(prod).astype(np.uint64)


def test_biased_estimate_bounds_and_exact_remainder():
    rng = np.random.default_rng(20261203)
    x, A = operands(rng, 400000)
    q = x // A
    for shift in (-1, 0, 1):
        est = estimate(x, A, shift)
        assert (est <= q).all(), "estimate above the quotient"
        fast = est < (1 << 18)
        assert (est[fast] + 1 >= q[fast]).all(), "fast-path estimate more than one below"
        assert (q[fast] < (1 << 19)).all()
        # the kernel's remainder: r = x - est A (in [0, 2A)), then min(r, r - A) in uint32
        r = (x - est * A) & 0xFFFFFFFF
        d = (r - A) & 0xFFFFFFFF
        rem = np.minimum(r, d)
        assert (rem[fast] == (x % A)[fast]).all()


def test_quotient_form_of_the_return_time_steps():
    """td_div_est (the return-time computation's runs): the estimate plus
    one when x - est Y >= Y is floor(x / Y) whenever est < 2^18."""
    rng = np.random.default_rng(99)
    x, A = operands(rng, 400000)
    q = x // A
    for shift in (-1, 0, 1):
        est = estimate(x, A, shift)
        fast = est < (1 << 18)
        r = (x - est * A) & 0xFFFFFFFF
        got = est + (r >= A)
        assert (got[fast] == q[fast]).all()


def test_unbiased_estimate_breaks_the_bound():
    rng = np.random.default_rng(7)
    x, A = operands(rng, 400000)
    est = estimate(x, A, 1, bias=False)
    assert (est > x // A).any()
