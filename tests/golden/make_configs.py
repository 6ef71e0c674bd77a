"""Writes tests/golden/configs.json: the canonical synthetic inputs of the
five benchmark workloads.  The committed file is the input; this script only
documents how it was drawn.

Generator: SplitMix64 with seed 0x5EED20261202 + i for workload i (1-based);
a uniform integer in [lo, hi] is lo + ((draw * (hi - lo + 1)) >> 64)
(Lemire's multiply-shift); the log-uniform a of workload 4 is
round(50000 * 2**(draw / 2**64)).

  1  a=50000 b=25000 c=100, K=J=256, primes < 10^6          (fixed)
  2  8 triples: a ~ U[50000,100000], c = 100, b ~ U[c+1, a-1]; K=J=4096,
     primes < 10^8
  3  128 candidates: a = 50000, b ~ U[1000,49000], c ~ U[100,b-1],
     k, j ~ U[1,4096]; each is tested as N and rev(N); bases 2,3,5,7
  4  512 candidates: a log-uniform in [50000,100000], b, c, k, j as in 3
  5  10000 requests: a ~ U[50000,100000], b, c, k, j as in 3, and a coin
     choosing N or rev(N); bases 2,3,5,7
"""
import json
from pathlib import Path

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo, hi):
        return lo + ((self.next() * (hi - lo + 1)) >> 64)


SEED = 0x5EED20261202


def candidates(rng, n, a_draw):
    out = []
    for _ in range(n):
        a = a_draw()
        b = rng.uniform(1000, 49000)
        c = rng.uniform(100, b - 1)
        k = rng.uniform(1, 4096)
        j = rng.uniform(1, 4096)
        out.append([a, b, c, k, j])
    return out


def main():
    cfg = {"generator": "SplitMix64, seed 0x5EED20261202 + workload index; Lemire multiply-shift"}
    cfg["1"] = {"a": 50000, "b": 25000, "c": 100, "K": 256, "J": 256, "prime_limit": 10 ** 6,
                "mr_survivors": 4, "bases": [2]}
    r2 = SplitMix64(SEED + 2)
    triples = []
    for _ in range(8):
        a = r2.uniform(50000, 100000)
        c = 100
        b = r2.uniform(c + 1, a - 1)
        triples.append([a, b, c])
    cfg["2"] = {"triples": triples, "K": 4096, "J": 4096, "prime_limit": 10 ** 8}
    r3 = SplitMix64(SEED + 3)
    cfg["3"] = {"candidates": candidates(r3, 128, lambda: 50000), "bases": [2, 3, 5, 7],
                "fields": "a b c k j"}
    r4 = SplitMix64(SEED + 4)
    cfg["4"] = {"candidates": candidates(r4, 512,
                                         lambda: round(50000 * 2 ** (r4.next() / 2 ** 64))),
                "bases": [2, 3, 5, 7], "fields": "a b c k j"}
    r5 = SplitMix64(SEED + 5)
    reqs = []
    for _ in range(10000):
        a = r5.uniform(50000, 100000)
        b = r5.uniform(1000, 49000)
        c = r5.uniform(100, b - 1)
        k = r5.uniform(1, 4096)
        j = r5.uniform(1, 4096)
        rev = r5.next() >> 63
        reqs.append([a, b, c, k, j, rev])
    cfg["5"] = {"requests": reqs, "bases": [2, 3, 5, 7], "fields": "a b c k j rev",
                "load_factors": [0.70, 0.95], "hbm_budget_bytes": 64 * 10 ** 9}
    path = Path(__file__).resolve().parent / "configs.json"
    path.write_text(json.dumps(cfg, separators=(",", ":")) + "\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
