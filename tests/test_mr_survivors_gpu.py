"""Phase-1 survivors straight into the Miller-Rabin scheduler
(mr_sched_submit_survivors): the (k, j) list stays on the device and each
job's number is built from its list entry when a slot takes it.  Checked
against GMP strong-probable-prime rounds (the oracle's mpz_powm harness),
against the same candidates submitted from the host, on the reference's own
CPU-runnable case (workload 1: sieve on the GPU, survivors extracted on the
GPU, the first four tested as N and rev(N), worker.c:45-54 order), on the
host fallback for triples whose survivors may differ in length, and through
a checkpoint taken while device-listed jobs are still queued."""
import json
import random

import numpy as np
import pytest

from _support import ROOT, gmp_build, gmp_sprp, numpy_primes
from test_mr_sched_gpu import expected, read_checkpoint

pytestmark = pytest.mark.gpu

BASES = [2, 3, 5, 7]


def say(*a):
    print("[mr-surv]", *a, flush=True)


def device_list(gpu, pairs):
    packed = np.array([(k << 16) | j for k, j in pairs], dtype=np.uint32)
    return gpu.DeviceBuffer.from_array(packed)


def by_tag(results):
    return {r.tag: (r.verdict, r.rounds) for r in results}


def test_workload1_survivors_in_place(gpu):
    fx = json.loads((ROOT / "tests" / "golden" / "pipeline_config1.json").read_text())
    a, b, c, K, J = fx["a"], fx["b"], fx["c"], fx["K"], fx["J"]
    L = gpu.lib()
    primes = numpy_primes(fx["prime_limit"])
    d_primes = gpu.DeviceBuffer.from_array(primes)
    d_bits = gpu.DeviceBuffer(4 * ((K * J + 31) // 32))
    d_bits.zero()
    assert L.trial_div_launch_kj(d_primes.ptr, len(primes), a, b, c, K, J, d_bits.ptr, None) == 0
    d_surv = gpu.DeviceBuffer(4 * K * J)
    import ctypes as C
    n = C.c_uint32()
    assert L.trial_div_survivors(d_bits.ptr, K, J, d_surv.ptr, K * J, C.byref(n), None) == 0
    assert n.value == fx["survivor_count"]
    say("workload 1:", n.value, "survivors on the device; first 4 as N and rev N")
    s = gpu.MrScheduler([fx["mr_base"]])
    s.submit_survivors(a, b, c, d_surv, 4, which=0, tag_base=0)
    s.submit_survivors(a, b, c, d_surv, 4, which=1, tag_base=100)
    got = by_tag(s.run_all())
    s.close()
    rows = fx["mr_rows"]                       # (k, j) of the first 4 survivors, which 0 and 1
    for i in range(4):
        for which, base in ((0, 0), (1, 100)):
            r = [x for x in rows if x["which"] == which][i]
            assert (r["k"], r["j"]) == tuple(fx["first_survivors"][i])
            assert got[base + i] == (int(r["pass"]), 1), (r, got[base + i])
    for d in (d_surv, d_bits, d_primes):
        d.close()


def test_in_place_equals_host_submission_and_gmp(gpu):
    """Device-listed jobs (k, j up to 65 535, N, rev N and the formula,
    three sizes) against host submission of the same candidates and GMP."""
    rng = random.Random(2026)
    pairs = [(rng.randrange(1, 65536), rng.randrange(1, 65536)) for _ in range(5)] + \
            [(1, 1), (9999, 9999), (65535, 1)]
    triples = [(400, 200, 50), (1200, 700, 100), (6000, 3100, 100)]
    s_dev = gpu.MrScheduler(BASES, max_active=6, chunk=64)
    s_host = gpu.MrScheduler(BASES, max_active=6, chunk=64)
    lists, cands = [], {}
    tag = 0
    for (a, b, c) in triples:
        d = device_list(gpu, pairs)
        lists.append(d)
        for which in (0, 1, 2):
            s_dev.submit_survivors(a, b, c, d, len(pairs), which=which, tag_base=tag)
            cs = [(a, b, c, k, j, which) for k, j in pairs]
            s_host.submit(cs, tags=list(range(tag, tag + len(cs))))
            for i, cd in enumerate(cs):
                cands[tag + i] = cd
            tag += len(cs)
    dev = by_tag(s_dev.run_all())
    host = by_tag(s_host.run_all())
    s_dev.close()
    s_host.close()
    say(len(cands), "candidates in place and from the host")
    assert set(dev) == set(cands) == set(host)
    for t, cd in cands.items():
        assert dev[t] == host[t] == expected(cd), cd
    for d in lists:
        d.close()


def test_variable_length_triple_takes_host_path(gpu):
    """max(b, c) > a - 7: k 10^b may reach the leading digit (rev N then
    loses leading zeros), so the list is read back and queued per
    candidate; verdicts equal GMP's."""
    pairs = [(1, 1), (99, 3), (12, 34), (7, 9)]
    a, b, c = 300, 297, 10
    d = device_list(gpu, pairs)
    s = gpu.MrScheduler([2, 3])
    s.submit_survivors(a, b, c, d, len(pairs), which=0, tag_base=0)
    s.submit_survivors(a, b, c, d, len(pairs), which=1, tag_base=10)
    got = by_tag(s.run_all())
    s.close()
    d.close()
    for i, (k, j) in enumerate(pairs):
        for which, base in ((0, 0), (1, 10)):
            assert got[base + i] == expected((a, b, c, k, j, which), [2, 3])


def test_checkpoint_reads_queued_list_entries(gpu, tmp_path):
    """A checkpoint while device-listed jobs wait in the queue writes their
    (k, j) (read from the list then); restored into a fresh scheduler they
    give GMP's verdicts."""
    pairs = [(k, j) for k in (3, 17, 250) for j in (1, 44, 4001)]
    a, b, c = 700, 350, 100
    d = device_list(gpu, pairs)
    s1 = gpu.MrScheduler(BASES, max_active=2, chunk=32)
    s1.submit_survivors(a, b, c, d, len(pairs), which=1, tag_base=500)
    s1.pump(False)
    path = tmp_path / "surv.ckpt"
    n = s1.checkpoint(path)
    before = by_tag(s1.results())
    s1.close()
    d.close()                                   # the restore must not need the list
    bases, rows = read_checkpoint(path)
    assert bases == BASES and len(rows) == n == len(pairs) - len(before)
    for tag, cd, rounds, bidx, bitpos, L, res in rows:
        assert cd == (a, b, c) + pairs[tag - 500] + (1,)
    s2 = gpu.MrScheduler(BASES)
    assert s2.restore(path) == n
    after = by_tag(s2.run_all())
    s2.close()
    got = {**before, **after}
    assert set(got) == set(range(500, 500 + len(pairs)))
    for i, (k, j) in enumerate(pairs):
        assert got[500 + i] == expected((a, b, c, k, j, 1)), (k, j)


def test_invalid_arguments_rejected(gpu):
    d = device_list(gpu, [(1, 1)])
    s = gpu.MrScheduler([2])
    for args in ((0, 0, 0, d, 1, 0), (400, 200, 50, d, 1, 3), (400, -1, 50, d, 1, 0),
                 (400, 200, 50, 0, 1, 0)):
        with pytest.raises(ValueError):
            s.submit_survivors(*args)
    s.submit_survivors(400, 200, 50, d, 0)          # nothing to queue
    assert s.run_all() == []
    s.close()
    d.close()
