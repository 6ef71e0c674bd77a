"""Continuous batching (mr_sched_* / mr_test_batch): mixed sizes across
transform buckets, slot reuse, early retirement and serving-style
submission, checked against GMP strong-probable-prime rounds base by base
(the oracle's mpz_powm harness)."""
import random
import sys
import time

import pytest

from _support import gmp_build, gmp_sprp

pytestmark = pytest.mark.gpu

BASES = [2, 3, 5, 7]


def say(*a):
    print("[mr-sched]", *a, flush=True)


def expected(cd, bases=BASES):
    """(verdict, rounds) of per-base Miller-Rabin with retirement."""
    n = gmp_build(*cd)
    rounds = 0
    for b in bases:
        rounds += 1
        ok, _, _ = gmp_sprp(n, b)
        if not ok:
            return 0, rounds
    return 1, rounds


def mixed_candidates(rng, digits, count, which=(0, 1, 2)):
    out = []
    for _ in range(count):
        a = digits
        b = rng.randrange(a // 3, a - 10)
        c = rng.randrange(1, b // 2)
        out.append((a, b, c, rng.randrange(1, 4097), rng.randrange(1, 4097), rng.choice(which)))
    return out


# (1000, 500, 100, k, j, 0) is prime for these (k, j) (GMP BPSW)
PRIMES_1000 = [(1000, 500, 100, 32, 5, 0), (1000, 500, 100, 47, 57, 0), (1000, 500, 100, 54, 26, 0)]


def test_mixed_sizes_small_pool(gpu):
    """Four slots, short chunks: many round ends and refills across the
    2^11, 2^12 and 2^13 buckets."""
    rng = random.Random(7)
    cands = (mixed_candidates(rng, 40, 6) + mixed_candidates(rng, 300, 6)
             + mixed_candidates(rng, 1000, 4) + PRIMES_1000
             + mixed_candidates(rng, 9000, 2) + mixed_candidates(rng, 20000, 1))
    rng.shuffle(cands)
    exp = [expected(cd) for cd in cands]
    say(len(cands), "candidates;", sum(v for v, _ in exp), "probable primes expected")
    s = gpu.MrScheduler(BASES, max_active=4, chunk=64)
    s.submit(cands, tags=list(range(100, 100 + len(cands))))
    res = s.run_all()
    st = s.stats()
    say("chunks", st.chunks, "admin batches", st.admin_batches, "pools", st.n_pools,
        "hbm MiB", st.hbm_bytes >> 20)
    assert len(res) == len(cands)
    got = {r.tag - 100: (r.verdict, r.rounds) for r in res}
    for i, cd in enumerate(cands):
        assert got[i] == exp[i], (cd, got[i], exp[i])
    for r in res:
        assert r.t_submit <= r.t_admit <= r.t_done
    assert st.completed == len(cands) and st.active == 0 and st.pending == 0
    assert st.n_pools == 3
    s.close()


def test_batch_call_default_pool(gpu):
    rng = random.Random(11)
    cands = mixed_candidates(rng, 1000, 60) + PRIMES_1000
    exp = [expected(cd) for cd in cands]
    t0 = time.perf_counter()
    v, r = gpu.mr_test_batch(cands, BASES)
    say(len(cands), "x 1000 digits in", round(time.perf_counter() - t0, 3), "s")
    assert list(zip(v, r)) == exp


def test_serving_submission_while_running(gpu):
    """Requests arrive while earlier ones run (non-blocking pumps)."""
    rng = random.Random(13)
    waves = [mixed_candidates(rng, 2000, 5) for _ in range(4)]
    waves[2] += PRIMES_1000[:1]
    s = gpu.MrScheduler([2, 3], max_active=6, chunk=32)
    res, tag = [], 0
    all_cands = []
    for w in waves:
        s.submit(w, tags=list(range(tag, tag + len(w))))
        all_cands += w
        tag += len(w)
        for _ in range(3):
            s.pump(False)
            res += s.results()
            s.pump(True)
            res += s.results()
    res += s.run_all()
    assert sorted(r.tag for r in res) == list(range(len(all_cands)))
    for r in res:
        assert (r.verdict, r.rounds) == expected(all_cands[r.tag], [2, 3])
    s.close()


def test_invalid_submission_rejected(gpu):
    s = gpu.MrScheduler(BASES, max_active=2)
    with pytest.raises(ValueError):
        s.submit([(10, 20, 1, 1, 1, 0)])           # N < 0
    with pytest.raises(ValueError):
        s.submit([(200000, 5, 1, 1, 1, 0)])        # beyond 2^15 transforms
    assert s.stats().submitted == 0
    s.close()


# 600-digit primes of the candidate form (found with GMP, re-checked below)
PRIMES_600 = [(600, 475, 249, 1179, 1874, 0), (600, 360, 167, 77, 87, 0),
              (600, 410, 392, 2792, 3950, 0)]


def read_checkpoint(path):
    """Rows (tag, candidate, rounds, base index, exponent bit, residue) of a
    checkpoint file, and its bases."""
    with open(path) as f:
        head = f.readline().split()
        assert head == ["mr-sched-checkpoint", "1"]
        b = f.readline().split()
        bases = [int(x) for x in b[2:]]
        assert b[0] == "bases" and int(b[1]) == len(bases)
        cnt = f.readline().split()
        rows = []
        for line in f:
            v = line.split()
            tag, a, bb, c, k, j, w, rounds, bidx, bitpos, L = (int(x) for x in v[:11])
            res = [int(x, 16) for x in v[11:]]
            assert len(res) == L
            rows.append((tag, (a, bb, c, k, j, w), rounds, bidx, bitpos, L, res))
        assert cnt[0] == "candidates" and int(cnt[1]) == len(rows)
    return bases, rows


def test_checkpoint_and_resume(gpu, tmp_path):
    """Checkpoint a running scheduler mid-exponent, restore it into a fresh
    one and finish: every saved residue equals base^(N >> (bit + 1)) R mod N
    (GMP), and verdicts and round counts equal an uninterrupted run's
    (known primes included, so a damaged residue cannot pass as composite)."""
    from _support import gmp_powm, gmp_probab_prime
    for cd in PRIMES_600:
        assert gmp_probab_prime(gmp_build(*cd), 25)
    rng = random.Random(77)
    cands = PRIMES_600 + mixed_candidates(rng, 600, 5, which=(0, 1)) + \
        mixed_candidates(rng, 7500, 4, which=(0, 1))
    ref_v, ref_r = gpu.mr_test_batch(cands, BASES)
    assert [ref_v[i] for i in range(3)] == [1, 1, 1]
    s1 = gpu.MrScheduler(BASES, chunk=32)
    s1.submit(cands, tags=list(range(len(cands))))
    for _ in range(6):
        s1.pump(True)
    path = tmp_path / "sched.ckpt"
    n = s1.checkpoint(path)
    before = {r.tag: (r.verdict, r.rounds) for r in s1.results()}
    s1.close()
    say("checkpoint:", n, "undecided,", len(before), "decided before")
    assert n == len(cands) - len(before) and n > 0
    bases, rows = read_checkpoint(path)
    assert bases == BASES and len(rows) == n
    mid = 0
    for tag, cd, rounds, bidx, bitpos, L, res in rows:
        assert cd == tuple(cands[tag])
        if bidx < 0:
            continue
        N = gmp_build(*cd)
        R = 1 << (24 * L)
        r = sum(v << (24 * i) for i, v in enumerate(res))
        assert r == gmp_powm(BASES[bidx], N >> (bitpos + 1), N) * R % N, f"residue of {cd}"
        s = ((N - 1) & -(N - 1)).bit_length() - 1
        mid += s <= bitpos < N.bit_length() - 2
    say("checkpoint:", sum(r[3] >= 0 for r in rows), "admitted,", mid, "mid-exponent")
    assert mid >= 1
    s2 = gpu.MrScheduler(BASES, chunk=32)
    assert s2.restore(path) == n
    after = {r.tag: (r.verdict, r.rounds) for r in s2.run_all()}
    s2.close()
    assert set(before) | set(after) == set(range(len(cands))) and not set(before) & set(after)
    got = {**before, **after}
    for i in range(len(cands)):
        assert got[i] == (ref_v[i], ref_r[i]), f"candidate {cands[i]}"
    # refusals: other bases, a truncated file, no file
    s3 = gpu.MrScheduler([2, 3])
    with pytest.raises(ValueError):
        s3.restore(path)
    s3.close()
    cut = tmp_path / "cut.ckpt"
    cut.write_text("\n".join(path.read_text().split("\n")[:3]) + "\n")      # no rows
    s4 = gpu.MrScheduler(BASES)
    with pytest.raises(ValueError):
        s4.restore(cut)
    with pytest.raises(ValueError):
        s4.restore(tmp_path / "missing.ckpt")
    s4.close()


def _drain(s, watch):
    """Pump until done, sampling the active count after every pump."""
    res, peak = [], 0
    while True:
        left = s.pump(True)
        res += s.results()
        peak = max(peak, s.stats().active)
        watch.append(peak)
        if left == 0:
            return res + s.results(), peak


def test_admission_by_job_footprint(gpu):
    """A budget that holds three jobs of a size: at most three are admitted at
    once and the others wait in the queue (their slots' device state is
    what admission counts, mr_sched_job_bytes); every verdict and round
    count equals GMP's; the admitted footprint never exceeds the budget."""
    rng = random.Random(31)
    cands = mixed_candidates(rng, 1500, 9) + PRIMES_1000[:2]
    bits = int(1500 * 3.3219280948873623) + 2
    per = gpu.lib().mr_sched_job_bytes(bits)
    assert per > 0
    budget = 3 * per + per // 2
    s = gpu.MrScheduler(BASES, hbm_budget=budget)
    s.submit(cands, tags=list(range(len(cands))))
    watch = []
    res, peak = _drain(s, watch)
    st = s.stats()
    say("budget", budget, "per job", per, "peak active", peak, "peak bytes", st.peak_job_bytes)
    assert peak <= 3 and st.peak_job_bytes <= budget and st.hbm_bytes <= budget
    assert st.completed == len(cands) and st.job_bytes == 0
    got = {r.tag: (r.verdict, r.rounds) for r in res}
    assert [got[i] for i in range(len(cands))] == [expected(cd) for cd in cands]
    s.close()


def test_admission_rejects_jobs_larger_than_the_budget(gpu):
    """A budget below one 8 000-digit job (2^12 transforms): that job is
    rejected at submission, a 1 000-digit job (2^11, a smaller footprint)
    is admitted and decided."""
    per = gpu.lib().mr_sched_job_bytes(int(8000 * 3.3219280948873623) + 2)
    assert gpu.lib().mr_sched_job_bytes(int(1000 * 3.3219280948873623) + 2) < per
    s = gpu.MrScheduler(BASES, hbm_budget=per - 1)
    with pytest.raises(ValueError):
        s.submit([(8000, 2500, 300, 4001, 17, 0)])
    # a smaller job still fits and runs
    s.submit([PRIMES_1000[0]], tags=[0])
    res = s.run_all()
    assert [(r.verdict, r.rounds) for r in res] == [(1, 4)]
    assert gpu.lib().mr_sched_job_bytes(10 ** 7) == -1          # beyond this build's sizes
    s.close()


def test_admission_tight_budget_two_sizes(gpu):
    """A budget of two jobs of the larger size, jobs of two transform
    lengths: the first size's pool is released once idle so the other size
    can be admitted (no stall), and all verdicts are right."""
    rng = random.Random(32)
    small = mixed_candidates(rng, 1200, 4)
    large = mixed_candidates(rng, 8000, 3) + [PRIMES_1000[1]]     # 2^12 and 2^11 transforms
    per_large = gpu.lib().mr_sched_job_bytes(int(8000 * 3.3219280948873623) + 2)
    s = gpu.MrScheduler(BASES, hbm_budget=2 * per_large)
    cands = small + large
    s.submit(cands, tags=list(range(len(cands))))
    res, _ = _drain(s, [])
    st = s.stats()
    say("two sizes: pools", st.n_pools, "reclaimed", st.reclaimed_pools, "peak bytes",
        st.peak_job_bytes)
    assert st.peak_job_bytes <= 2 * per_large and st.hbm_bytes <= 2 * per_large
    assert st.reclaimed_pools >= 1
    got = {r.tag: (r.verdict, r.rounds) for r in res}
    assert [got[i] for i in range(len(cands))] == [expected(cd) for cd in cands]
    s.close()
