"""CPU model of the Montgomery step's wrapped tail (limb level): with
W = m N mod (x^H - 1) evaluated at x = 2^24 (H = M/2 > L), folding
V = T + W modulo B^H - 1 and moving the wrapped limbs back up gives
t = (T + m N) / R exactly, for every limb count up to M/2 - 1.  This is
the arithmetic the device's wrap_tail implements (fold, one end-around
carry, unwrap into the high layout).  The end-around carry needs the low
half of V to be all ones above position 2L - H, which natural operands
reach with probability about 2^-48 per step, so it is driven here with
values V of the right residue class built for it; the device's products
are checked by the GPU parity tests."""
import random

B = 1 << 24


def limbs(x, n):
    return [(x >> (24 * i)) & (B - 1) for i in range(n)]


def value(v):
    return sum(x << (24 * i) for i, x in enumerate(v))


def cyclic(a, b, H):
    c = [0] * H
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                c[(i + j) % H] += x * y
    return c


def fold_unwrap(V, L, M, end_around=True):
    """The device's sequence on V = T + W: S = V_lo + V_hi with one
    end-around carry; t's limbs at [L, H) stay, those at [H, 2L] come from
    [0, 2L + 1 - H)."""
    H = M // 2
    assert V < B ** M
    v = limbs(V, M)
    S = value(v[:H]) + value(v[H:])
    if S >= B ** H:
        S = S - B ** H + 1 if end_around else S - B ** H
    s = limbs(S, H)
    r = [0] * M
    for pos in range(M):
        r[pos] = 0 if pos < L else (s[pos] if pos < H else (s[pos - H] if pos <= 2 * L else 0))
    return value(r) >> (24 * L)


def tail(T, W, L, M):
    V = T + W
    H = M // 2
    v = limbs(V, M)
    S = value(v[:H]) + value(v[H:])
    S = S - B ** H + 1 if S >= B ** H else S
    assert S < B ** H - 1 or S == 0          # never the all-ones form of zero
    return fold_unwrap(V, L, M)


def test_wrapped_tail_every_limb_count():
    rng = random.Random(1)
    for M in (8, 16, 32, 64):
        H = M // 2
        for L in range(1, H):
            R = B ** L
            for trial in range(6):
                N = rng.randrange(B ** (L - 1) if L > 1 else 3, R // 2) | 1
                if trial == 0:
                    N = (R // 2 - 1) | 1
                Np = (-pow(N, -1, R)) % R
                a, b = rng.randrange(N), rng.randrange(N)
                if trial == 1:
                    a = 0
                if trial == 2:
                    a = b = N - 1
                T = a * b
                m = (T % R) * Np % R
                W = value(cyclic(limbs(m, L), limbs(N, L), H))
                assert (T + m * N) % R == 0
                assert tail(T, W, L, M) == (T + m * N) // R, (M, L, trial)


def crafted(rng, M):
    """(V, L, t) with V == t B^L (mod B^H - 1) chosen so that V_lo + V_hi
    overflows B^H: the end-around carry, which real operands reach with
    probability about B^-2 per step at the largest moduli."""
    H = M // 2
    L = rng.randrange(1, H)
    t = rng.randrange(1, B ** (L + 1) // 2)
    X = t * B ** L % (B ** H - 1)
    q = rng.randrange(X + 1, B ** H)             # q > X forces the carry
    return X + q * (B ** H - 1), L, t


def test_end_around_carry_cases():
    rng = random.Random(3)
    for M in (8, 16, 32, 64):
        H = M // 2
        for _ in range(50):
            V, L, t = crafted(rng, M)
            v = limbs(V, M)
            assert value(v[:H]) + value(v[H:]) >= B ** H
            assert fold_unwrap(V, L, M) == t


def test_model_catches_a_dropped_end_around_carry():
    """Planted mistake: without the end-around carry the crafted cases give
    a wrong t (so the test above constrains that step)."""
    rng = random.Random(4)
    wrong = 0
    for _ in range(50):
        V, L, t = crafted(rng, 16)
        wrong += fold_unwrap(V, L, 16, end_around=False) != t
    assert wrong == 50
