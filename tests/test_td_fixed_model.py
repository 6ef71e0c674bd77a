"""CPU model of the lane kernel's fixed-point residue images (primes from
2^24): a residue x mod p is carried as round(x 2^32 / p) in a 32-bit word
and advanced by subtracting the image of the step modulo 2^32, which wraps
exactly where x - s wraps mod p.  Checked against exact residues with the
device's rounding (double precision, round to nearest): the accumulated
error stays within (K + 1) / 2 units after K steps, far inside the bound
eoff = 2K + 32 the kernel uses, so every true hit (x < thr) falls in the
widened window [0, thr fp + 2 eoff) of the shifted image; and without the
widening some hits would be missed (the widening is needed).  The device's
bitmaps are checked bit for bit by the GPU parity tests."""
import random

import numpy as np

M32 = 1 << 32


def fix(x, fp):
    return int(round(x * fp)) % M32          # __double2ull_rn((double)x * fp)


def run_stream(p, x0, step, K, thr, widen=True):
    """Images of x_k = x0 - k step (mod p), k = 1..K, as the kernel advances
    them; returns (max |error| in units, hits missed by the window)."""
    fp = 4294967296.0 / p
    eoff = 2 * K + 32 if widen else 0
    img = (fix(x0, fp) + eoff) % M32
    s_img = fix(step, fp)
    win = (int(np.ceil(thr * fp)) + 2 * eoff) if widen else int(np.ceil(thr * fp))
    x, worst, missed = x0, 0.0, 0
    for _ in range(K):
        img = (img - s_img) % M32
        x = (x - step) % p
        err = ((img - eoff - x * M32 / p + M32 / 2) % M32) - M32 / 2
        worst = max(worst, abs(err))
        if x < thr and not img < win:
            missed += 1
    return worst, missed


def test_error_bound_and_no_missed_hits():
    rng = random.Random(24)
    K = 4096
    for _ in range(40):
        p = rng.choice([16777259, 16777289, 50331653, 99999989, 2147483647, 33554467])
        x0, step = rng.randrange(p), rng.randrange(1, p)
        thr = rng.choice([4096, 9999])
        worst, missed = run_stream(p, x0, step, K, thr)
        assert worst <= (K + 1) / 2 + 1e-3, (p, worst)
        assert missed == 0


def test_hits_near_the_window_edge():
    """Streams that land on x = thr - 1 (the last hit) and x = 0: both must
    be inside the widened window."""
    rng = random.Random(7)
    K = 9999
    for p in (16777259, 99999989, 2147483647):
        for target in (0, 4095, 9998):
            step = rng.randrange(1, p)
            k_hit = rng.randrange(1, K + 1)
            x0 = (target + k_hit * step) % p        # x_{k_hit} = target
            worst, missed = run_stream(p, x0, step, K, 9999 if target == 9998 else 4096)
            assert missed == 0 and worst <= (K + 1) / 2 + 1e-3


def test_unwidened_window_misses_hits():
    """Planted mistake: with eoff = 0 the rounding drift pushes some true
    hits out of the window (so the widening is what the test above pins)."""
    rng = random.Random(3)
    missed = 0
    for _ in range(200):
        p = 16777259
        step = rng.randrange(1, p)
        x0 = (0 + rng.randrange(1, 4097) * step) % p   # a hit at x = 0 somewhere
        missed += run_stream(p, x0, step, 4096, 4096, widen=False)[1]
    assert missed > 0
