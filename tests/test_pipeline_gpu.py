"""End to end on the reference's own CPU-runnable case (workload 1: a =
50 000, b = 25 000, c = 100, k, j in 1..256, primes below 10^6): trial
division on the GPU, survivors extracted on the GPU (trial_div_survivors,
worker.c:45-54 order), then base-2 Miller-Rabin of N and rev(N) for the
first 4 survivors, all against the committed fixture written by
tests/golden/make_pipeline_golden.py from the oracle (sieve) and GMP."""
import hashlib
import json
import sys

import numpy as np
import pytest

from _support import ROOT, numpy_primes

pytestmark = pytest.mark.gpu
sys.set_int_max_str_digits(0)


def test_config1_pipeline(gpu):
    fx = json.loads((ROOT / "tests" / "golden" / "pipeline_config1.json").read_text())
    a, b, c, K, J = fx["a"], fx["b"], fx["c"], fx["K"], fx["J"]
    L = gpu.lib()
    primes = numpy_primes(fx["prime_limit"])
    d_primes = gpu.DeviceBuffer.from_array(primes)
    nw = (K * J + 31) // 32
    d_bits = gpu.DeviceBuffer(4 * nw)
    d_bits.zero()
    assert L.trial_div_launch_kj(d_primes.ptr, len(primes), a, b, c, K, J, d_bits.ptr, None) == 0
    assert L.pf_sync() == 0 and L.trial_div_check_error() == 0
    surv = gpu.survivors_gpu(d_bits, K, J)
    assert len(surv) == fx["survivor_count"]
    assert surv[:64].tolist() == fx["first_survivors"]
    packed = ((surv[:, 0] << 16) | surv[:, 1]).astype("<u4")
    assert hashlib.sha256(packed.tobytes()).hexdigest() == fx["survivors_sha256"]
    # Phase 2 on the first survivors (each as N and as its decimal reversal)
    rows = fx["mr_rows"]
    cands = [(a, b, c, r["k"], r["j"], r["which"]) for r in rows]
    ctx = gpu.MrContext(candidates=cands)
    mods = ctx.moduli()
    alive, res = ctx.run([fx["mr_base"]])
    for r, n, al, rr in zip(rows, mods, alive, res):
        nwr = r["n_words"]
        assert hashlib.sha256(n.to_bytes(4 * nwr, "little")).hexdigest() == r["n_sha256"]
        assert hashlib.sha256(rr[0].to_bytes(4 * nwr, "little")).hexdigest() == r["residue_sha256"]
        assert bool(al) == r["pass"]
    ctx.close()
    # the same verdicts through the scheduler
    v, rounds = gpu.mr_test_batch(cands, [fx["mr_base"]])
    assert v == [int(r["pass"]) for r in rows]
    d_bits.close()
    d_primes.close()


def test_survivors_edges(gpu):
    """Ragged grids (K*J not a multiple of 32), all-eliminated and
    nothing-eliminated bitmaps, and a cap below the count."""
    rng = np.random.default_rng(3)
    for K, J in ((1, 1), (3, 7), (17, 33), (100, 100), (257, 255)):
        nw = (K * J + 31) // 32
        bits = rng.integers(0, 2 ** 32, nw, dtype=np.uint64).astype(np.uint32)
        d = gpu.DeviceBuffer.from_array(bits)
        flat = np.unpackbits(bits.view(np.uint8), bitorder="little")[: K * J]
        idx = np.nonzero(flat == 0)[0]
        exp = np.stack([idx // J + 1, idx % J + 1], axis=1)
        got = gpu.survivors_gpu(d, K, J)
        assert got.tolist() == exp.tolist(), (K, J)
        if len(exp) > 3:
            assert gpu.survivors_gpu(d, K, J, cap=3).tolist() == exp[:3].tolist()
        d.close()
    K = J = 64
    full = gpu.DeviceBuffer.from_array(np.full(128, 0xFFFFFFFF, np.uint32))
    assert len(gpu.survivors_gpu(full, K, J)) == 0
    empty = gpu.DeviceBuffer.from_array(np.zeros(128, np.uint32))
    assert len(gpu.survivors_gpu(empty, K, J)) == K * J
