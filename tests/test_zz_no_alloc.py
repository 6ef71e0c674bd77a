"""The allocation-free trial-division launch (trial_div_launch_kj_noalloc)
against trial_div_launch_kj through the C ABI: two calls in a row on one
stream, the second with a smaller input (fewer primes, smaller grid), each
bitmap bit-exact against the allocating launch's and the oracle's; then the
error latch of its own scratch and the size limit."""
import numpy as np
import pytest

from _support import nwords, numpy_primes, oracle_sieve

pytestmark = pytest.mark.gpu


def say(*a):
    print("[no-alloc]", *a, flush=True)


def _run(L, fn, dp, n, a, b, c, K, J, st):
    from _support import load_pf
    pf = load_pf()
    db = pf.DeviceBuffer(nwords(K, J) * 4)
    db.zero()
    assert fn(dp.ptr, n, a, b, c, K, J, db.ptr, st) == 0
    assert L.pf_stream_sync(st) == 0
    out = db.download()
    db.close()
    return out


def test_noalloc_matches_allocating_launch(gpu):
    L = gpu.lib()
    st = L.pf_stream_create()
    assert st
    # primes below 2^16 (row kernel, scratch table and items), above (lane
    # kernel, rev decode table in scratch) and above 2^24 (fixed-point path)
    big = np.concatenate([numpy_primes(1_000_000),
                          np.array([16777259, 50331653, 99999989], np.uint32)])
    small = numpy_primes(30000)
    cases = [(big, 60001, 31234, 100, 1024, 1024), (small, 50000, 25000, 100, 256, 200)]
    for primes, a, b, c, K, J in cases:
        say("case", len(primes), "primes", (a, b, c), "grid", K, "x", J)
        dp = gpu.DeviceBuffer.from_array(primes)
        got = _run(L, L.trial_div_launch_kj_noalloc, dp, len(primes), a, b, c, K, J, st)
        want = _run(L, L.trial_div_launch_kj, dp, len(primes), a, b, c, K, J, st)
        assert L.trial_div_check_error() == 0
        assert (got == want).all()
        assert (got == oracle_sieve(primes, a, b, c, K, J)).all()
        dp.close()
    L.pf_stream_destroy(st)


def test_noalloc_limits_and_error_latch(gpu):
    L = gpu.lib()
    primes = numpy_primes(1000)
    dp = gpu.DeviceBuffer.from_array(primes)
    db = gpu.DeviceBuffer(nwords(100, 100) * 4)
    assert L.trial_div_launch_kj_noalloc(dp.ptr, len(primes), 100, 50, 10, 10000, 10, db.ptr,
                                         None) == -1
    assert L.trial_div_launch_kj_noalloc(dp.ptr, len(primes), 100, 50, 10, 10, 0, db.ptr,
                                         None) == -1
    bad = gpu.DeviceBuffer.from_array(np.concatenate([primes, np.array([2147483659], np.uint32)]))
    db.zero()
    assert L.trial_div_launch_kj_noalloc(bad.ptr, len(primes) + 1, 100, 50, 10, 100, 100, db.ptr,
                                         None) == 0
    assert L.trial_div_check_error() == 1
    assert L.trial_div_check_error() == 0          # cleared once read
    db.zero()
    assert L.trial_div_launch_kj_noalloc(dp.ptr, len(primes), 100, 50, 10, 100, 100, db.ptr,
                                         None) == 0
    assert L.trial_div_check_error() == 0
    assert (db.download() == oracle_sieve(primes, 100, 50, 10)).all()
    for x in (dp, db, bad):
        x.close()
