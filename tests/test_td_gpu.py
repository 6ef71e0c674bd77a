"""GPU parity of Phase-1 trial division (HIP library, called through the
C ABI) against the reference's golden bitmaps and the CPU oracle."""
import json
import random

import numpy as np
import pytest

from _support import (GOLDEN, bit_get, golden_cases, nwords, numpy_primes, oracle_check_pairs, oracle_sieve,
                      popcount, survivors)

pytestmark = pytest.mark.gpu

CONFIGS = json.loads((GOLDEN / "configs.json").read_text())


def say(*a):
    print("[td-gpu]", *a, flush=True)


def test_reference_abi_golden(gpu):
    """trial_div_gpu (K = J = 100) equals the reference's trial_div_cpu
    bitmaps committed under tests/golden."""
    for case, primes, words in golden_cases():
        say("golden", case["primes"], case["a"], case["b"], case["c"])
        rc, got = gpu.trial_div_gpu(primes, case["a"], case["b"], case["c"])
        assert rc == 0
        assert (got == words).all(), case


def test_reference_abi_timing_and_or_semantics(gpu):
    primes = numpy_primes(200000)
    rng = np.random.default_rng(3)
    pre = rng.integers(0, 2 ** 32, 313, dtype=np.uint64).astype(np.uint32)
    pre[312] &= 0xFFFF
    rc, got, t = gpu.trial_div_gpu(primes, 50001, 33334, 2749, bits=pre, timing=True)
    assert rc == 0
    ref = oracle_sieve(primes, 50001, 33334, 2749) | pre
    assert (got == ref).all()
    assert all(x >= 0.0 for x in t) and t[1] > 0.0
    # second call: device prime cache hit; same answer
    rc2, got2 = gpu.trial_div_gpu(primes, 50001, 33334, 2749, bits=pre)
    assert rc2 == 0 and (got2 == got).all()
    # a new buffer with other contents: cache must refresh
    p2 = primes.copy()
    p2[-1] = 3
    rc3, got3 = gpu.trial_div_gpu(p2, 50001, 33334, 2749)
    assert rc3 == 0
    assert (got3 == oracle_sieve(p2, 50001, 33334, 2749)).all()


def test_prime_cache_same_address_refresh(gpu):
    """The device prime list is reused while the caller's list is unchanged;
    contents changed in place (same pointer, same count) must be seen, at
    the first entry, the last and in the middle."""
    primes = numpy_primes(300000)
    a, b, c = 50001, 33334, 2749
    rc, first = gpu.trial_div_gpu(primes, a, b, c)
    assert rc == 0 and (first == oracle_sieve(primes, a, b, c)).all()
    addr = primes.ctypes.data
    for pos, val in ((len(primes) - 1, 3), (0, 7), (len(primes) // 2, 11)):
        primes[pos] = val
        rc, got = gpu.trial_div_gpu(primes, a, b, c)
        assert primes.ctypes.data == addr
        assert rc == 0
        assert (got == oracle_sieve(primes, a, b, c)).all(), pos


@pytest.mark.parametrize("K,J", [(1, 1), (9, 11), (100, 100), (150, 130), (1200, 37),
                                 (37, 1200), (999, 1001), (4096, 64), (64, 4096)])
def test_kj_grids_vs_oracle(gpu, K, J):
    primes = numpy_primes(150000)        # both kernels: below and above 2^16
    rng = random.Random(K * 7919 + J)
    for _ in range(2):
        a = rng.randint(50, 100000)
        b = rng.randint(1, a - 1)
        c = rng.randint(0, b - 1)
        say("grid", K, J, a, b, c)
        rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, K, J)
        assert rc == 0
        ref = oracle_sieve(primes, a, b, c, K, J)
        assert (got == ref).all(), (K, J, a, b, c)


def test_edge_exponents_vs_oracle(gpu):
    """a-b-l and a-c-l at and below zero; 2 and 5 in the list."""
    primes = numpy_primes(100000)
    for (a, b, c) in [(104, 101, 100), (103, 100, 1), (7, 4, 2), (6, 3, 1), (4, 3, 1),
                      (1005, 1001, 1000), (1004, 1000, 996), (10, 0, 0), (5000, 4999, 4998),
                      (1003, 1001, 1000), (1002, 1001, 1000), (50004, 50001, 50000)]:
        say("edge", a, b, c)
        for (K, J) in [(100, 100), (1500, 1500)]:
            rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, K, J)
            assert rc == 0
            assert (got == oracle_sieve(primes, a, b, c, K, J)).all(), (a, b, c, K, J)


def test_workload1_exact(gpu):
    """The 50k-digit K = J = 256 workload over all 78 498 primes < 10^6."""
    cfg = CONFIGS["1"]
    primes = numpy_primes(cfg["prime_limit"])
    assert len(primes) == 78498
    a, b, c, K, J = cfg["a"], cfg["b"], cfg["c"], cfg["K"], cfg["J"]
    rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, K, J)
    assert rc == 0
    ref = oracle_sieve(primes, a, b, c, K, J)
    assert (got == ref).all()
    surv = survivors(got, K, J)
    say("workload1 survivors", len(surv), "first 4", surv[:4])


def test_4096_sampled_primes_vs_oracle(gpu):
    """Full 4096 x 4096 grid, primes sampled from the 10^8 list: every prime
    below 2^16, every 499th above, and primes next to 2^16 and 2^31."""
    allp = gpu.primes_below(10 ** 8)
    assert len(allp) == 5761455
    big = allp[allp >= 65536][::499]
    extra = np.array([65537, 65539, 65521, 65519, 2147483647, 2147483629, 2147483587],
                     np.uint32)
    primes = np.concatenate([allp[allp < 65536], big, extra])
    for (a, b, c) in CONFIGS["2"]["triples"][:2]:
        say("sampled 4096", a, b, c, len(primes))
        rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, 4096, 4096)
        assert rc == 0
        ref = oracle_sieve(primes, a, b, c, 4096, 4096)
        assert (got == ref).all(), (a, b, c)


def test_max_grid_sampled_primes_vs_oracle(gpu):
    """K = J = 9999 (the largest grid): the lane kernel's fixed-point
    residue images carry the most accumulated rounding there (up to K + 1
    steps), so its candidate windows are widest; sampled primes from 2^16
    to 10^8 and next to 2^31 against the oracle, bit for bit."""
    allp = gpu.primes_below(10 ** 8)
    big = allp[allp >= 65536][::20011]
    extra = np.array([65537, 65539, 2147483647, 2147483629], np.uint32)
    primes = np.concatenate([allp[(allp > 60000) & (allp < 65536)][:8], big, extra])
    for (a, b, c) in [(60000, 31234, 100), (50001, 33334, 2749)]:
        say("max grid 9999", a, b, c, len(primes))
        rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, 9999, 9999)
        assert rc == 0
        ref = oracle_sieve(primes, a, b, c, 9999, 9999)
        assert (got == ref).all(), (a, b, c)


def test_edge_exponents_large_primes_vs_oracle(gpu):
    """Exponent-edge triples (separate j digit groups when a - c < D_j, rev
    terms with negative exponents) over primes from 2^24 up, which take the
    lane kernel's fixed-point residue path, and next to 2^31."""
    allp = gpu.primes_below(10 ** 8)
    primes = np.concatenate([allp[allp >= (1 << 24)][::9973],
                             np.array([16777259, 2147483647, 2147483629], np.uint32)])
    for (a, b, c) in [(1003, 1001, 1000), (1002, 1001, 1000), (5000, 4999, 4998),
                      (1005, 1001, 1000), (50004, 50001, 50000)]:
        say("edge large", a, b, c, len(primes))
        for (K, J) in [(1500, 1500), (4096, 4096)]:
            rc, got = gpu.trial_div_gpu_kj(primes, a, b, c, K, J)
            assert rc == 0
            assert (got == oracle_sieve(primes, a, b, c, K, J)).all(), (a, b, c, K, J)


def test_workload2_full_triple_properties(gpu):
    """One full workload-2 triple (5 761 455 primes, 4096 x 4096): sampled
    bits against full-list oracle verdicts, determinism, and the OR-split
    property bitmap(P) = bitmap(P1) | bitmap(P2)."""
    allp = gpu.primes_below(10 ** 8)
    a, b, c = CONFIGS["2"]["triples"][0]
    K = J = 4096
    rc, full = gpu.trial_div_gpu_kj(allp, a, b, c, K, J)
    assert rc == 0
    say("workload2 triple", a, b, c, "eliminated", popcount(full), "of", K * J)
    rc, again = gpu.trial_div_gpu_kj(allp, a, b, c, K, J)
    assert rc == 0 and (again == full).all()
    half = len(allp) // 2
    rc, h1 = gpu.trial_div_gpu_kj(allp[:half], a, b, c, K, J)
    assert rc == 0
    rc, h2 = gpu.trial_div_gpu_kj(allp[half:], a, b, c, K, J)
    assert rc == 0
    assert ((h1 | h2) == full).all()
    surv = survivors(full, K, J)
    rng = np.random.default_rng(11)
    pick_s = [surv[i] for i in rng.choice(len(surv), 64, replace=False)]
    elim = []
    while len(elim) < 64:
        k, j = int(rng.integers(1, K + 1)), int(rng.integers(1, J + 1))
        if bit_get(full, J, k, j):
            elim.append((k, j))
    pairs = pick_s + elim
    flags = oracle_check_pairs(allp, a, b, c, [p[0] for p in pairs], [p[1] for p in pairs])
    for (k, j), f in zip(pairs, flags):
        assert bit_get(full, J, k, j) == int(f), (k, j)


def test_workload2_all_triples_bitexact(gpu):
    """The full trial-division workload: all 8 seeded triples, K = J = 4096,
    every prime below 10^8 (5 761 455, resident on the device across
    calls).  Each GPU bitmap's SHA-256 equals the oracle's full-list bitmap
    (tests/golden/td_full_4096.json, made by make_td_full_golden.py)."""
    import hashlib
    gold = json.loads((GOLDEN / "td_full_4096.json").read_text())
    allp = gpu.primes_below(10 ** 8)
    assert len(allp) == 5761455
    assert [[t["a"], t["b"], t["c"]] for t in gold["triples"]] == CONFIGS["2"]["triples"]
    for t in gold["triples"]:
        rc, got = gpu.trial_div_gpu_kj(allp, t["a"], t["b"], t["c"], 4096, 4096)
        assert rc == 0
        say("workload2", t["a"], t["b"], t["c"], "survivors", 4096 * 4096 - popcount(got))
        assert popcount(got) == t["eliminated"]
        assert hashlib.sha256(got.tobytes()).hexdigest() == t["bits_sha256"], t


def test_device_resident_launch(gpu):
    """trial_div_launch_kj on device buffers, timed with events on a stream."""
    L = gpu.lib()
    primes = numpy_primes(300000)
    a, b, c, K, J = 60000, 30000, 100, 512, 512
    dp = gpu.DeviceBuffer.from_array(primes)
    db = gpu.DeviceBuffer(nwords(K, J) * 4)
    db.zero()
    st = L.pf_stream_create()
    assert st
    e0, e1 = L.pf_event_create(), L.pf_event_create()
    assert L.pf_event_record(e0, st) == 0
    assert L.trial_div_launch_kj(dp.ptr, len(primes), a, b, c, K, J, db.ptr, st) == 0
    assert L.pf_event_record(e1, st) == 0
    assert L.pf_stream_sync(st) == 0
    assert L.pf_event_elapsed_ms(e0, e1) > 0
    assert L.trial_div_check_error() == 0
    got = db.download()
    assert (got == oracle_sieve(primes, a, b, c, K, J)).all()
    L.pf_event_destroy(e0)
    L.pf_event_destroy(e1)
    L.pf_stream_destroy(st)


def test_errors_are_reported(gpu):
    primes = numpy_primes(1000)
    rc, bits = gpu.trial_div_gpu_kj(primes, 100, 50, 10, 0, 10)
    assert rc == -1
    bad = np.concatenate([primes, np.array([2147483659], np.uint32)])   # >= 2^31
    pre = np.full(313, 0x12345678, np.uint32)
    rc, out = gpu.trial_div_gpu(bad, 100, 50, 10, bits=pre)
    assert rc == -1
    # the error is cleared: a valid call afterwards succeeds
    rc, out = gpu.trial_div_gpu(primes, 100, 50, 10)
    assert rc == 0
    assert (out == oracle_sieve(primes, 100, 50, 10)).all()
