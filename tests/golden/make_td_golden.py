"""Writes tests/golden/td_ref_k100.json: Phase-1 bitmaps (K = J = 100, 313
words) produced by the reference's own CPU sieve, trial_div_cpu.c, compiled
from its sources into oracle/_ref/libref_td.so by oracle/Makefile.

Run from the repo root after `make -C oracle`:
    python tests/golden/make_td_golden.py

Inputs: the prime set of the reference's GPU test (tests/test_gpu.c:160-162)
with its triples (test_gpu.c:169-172, test_sieve.c:78-79), and all primes
below 20000 (2 included) with the reference's benchmark triple
(tests/runner.cu:49-51), the K = J = 100 slice of the 50k-digit workload,
and triples whose rev exponents reach 0 or go negative.
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from _support import numpy_primes, ref_td, ref_trial_div_cpu  # noqa: E402

TEST_GPU_PRIMES = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
                   73, 79, 83, 89, 97]


def main():
    if ref_td() is None:
        raise SystemExit("oracle/_ref/libref_td.so not built (needs the reference sources)")
    p20k = numpy_primes(20000)
    cases = []
    for (a, b, c) in [(15, 8, 3), (20, 10, 5), (25, 12, 6), (30, 15, 7), (12, 6, 3)]:
        cases.append(("test_gpu", a, b, c))
    for (a, b, c) in [(50001, 33334, 2749), (50000, 25000, 100), (104, 101, 100),
                      (103, 100, 1), (7, 4, 2), (6, 3, 1), (60123, 45000, 777)]:
        cases.append(("lt20000", a, b, c))
    out = {"source": "reference trial_div_cpu.c via oracle/_ref/libref_td.so",
           "K": 100, "J": 100, "prime_sets": {"test_gpu": "tests/test_gpu.c:160-162 list",
                                              "lt20000": "all primes < 20000, 2 included"},
           "cases": []}
    for (ps, a, b, c) in cases:
        primes = TEST_GPU_PRIMES if ps == "test_gpu" else p20k
        bits = ref_trial_div_cpu(primes, a, b, c)
        out["cases"].append({"primes": ps, "a": a, "b": b, "c": c,
                             "bits": "".join(f"{int(w):08x}" for w in bits)})
    (HERE / "td_ref_k100.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(out['cases'])} cases")


if __name__ == "__main__":
    main()
