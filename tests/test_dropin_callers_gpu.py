"""The drop-in boundary exercised from compiled callers, the way the
reference's own programs call it:

* worker_dropin (C): prime_list_load, cuda_get_device_count /
  cuda_set_device (main.c:92-94), trial_div_gpu (worker.c:37-42), the
  survivor walk (worker.c:45-54) and miller_rabin_test (worker.c:84);
* bench_gpu_dropin (C++): trial_div_launch with C++ linkage on device
  buffers (tests/bench_gpu.cu:25-27,46,56).

Bitmaps are checked against the reference's own trial_div_cpu output
(tests/golden/td_ref_k100.json), codes against the strong-probable-prime
verdicts of N and the reference rev formula to bases 2, 3, 5, 7."""
import subprocess

import numpy as np
import pytest

from _support import PKG, bit_get, gmp_build, golden_cases, sprp_python, survivors, write_gap_file

pytestmark = pytest.mark.gpu


def _run(name, *args):
    exe = PKG / "lib" / name
    assert exe.exists(), f"{name} not built"
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def _bits(lines):
    row = next(l for l in lines if l.startswith("bits"))
    return np.array([int(w, 16) for w in row.split()[1:]], np.uint32)


def _code(a, b, c, k, j):
    """miller_rabin.c:76-118 order with this library's verdict (SPRP to 2, 3, 5, 7)."""
    sprp = lambda n: all(sprp_python(n, q)[0] for q in (2, 3, 5, 7)) if n > 7 else n in (2, 3, 5, 7)
    if not sprp(gmp_build(a, b, c, k, j, 0)):
        return 0
    return 2 if sprp(gmp_build(a, b, c, k, j, 2)) else 1


def test_worker_c_caller(gpu, tmp_path):
    for case, primes, words in golden_cases():
        a, b, c = case["a"], case["b"], case["c"]
        f = tmp_path / "primes.bin"
        write_gap_file(f, primes)
        # miller_rabin_test on the first survivors where a round takes
        # milliseconds (a 50 000-digit N takes ~20 s on one compute unit)
        n_mr = 3 if a < 2000 else 0
        out = _run("worker_dropin", f, a, b, c, n_mr)
        got = _bits(out)
        assert (got == words).all(), case
        surv = survivors(words, 100, 100)
        assert f"survivors {len(surv)}" in out
        mr = [l.split()[1:] for l in out if l.startswith("mr ")]
        assert len(mr) == min(n_mr, len(surv))
        for (k, j, code), (sk, sj) in zip(mr, surv):
            assert (int(k), int(j)) == (sk, sj)
            assert int(code) == _code(a, b, c, sk, sj), (case, sk, sj)


def test_bench_gpu_cpp_caller(gpu, tmp_path):
    for case, primes, words in golden_cases():
        f = tmp_path / "primes.bin"
        write_gap_file(f, primes)
        out = _run("bench_gpu_dropin", f, case["a"], case["b"], case["c"], 3)
        assert (_bits(out) == words).all(), case
        assert any(l.startswith("ms_per_run") for l in out)
