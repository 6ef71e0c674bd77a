"""Writes tests/golden/pipeline_config1.json: the end-to-end fixture of the
reference's own CPU-runnable case (workload 1: a = 50 000, b = 25 000,
c = 100, k, j in 1..256, primes below 10^6): the survivors of the oracle's
trial division (count, the first 64 and a SHA-256 of all, packed
(k << 16) | j, little-endian uint32), then GMP base-2 strong-probable-prime
residues of N and rev(N) (decimal reversal) for the first 4 survivors, as
SHA-256 digests of the little-endian uint32 words.

    python tests/golden/make_pipeline_golden.py
"""
import hashlib
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))


def digest(x: int, n_words: int) -> str:
    return hashlib.sha256(x.to_bytes(4 * n_words, "little")).hexdigest()


def one(args):
    from _support import gmp_build, gmp_sprp
    a, b, c, k, j, which = args
    n = gmp_build(a, b, c, k, j, which)
    nw = (n.bit_length() + 31) // 32
    v, r, s = gmp_sprp(n, 2)
    return {"k": k, "j": j, "which": which, "n_words": nw, "n_sha256": digest(n, nw), "s": s,
            "pass": bool(v == 1), "residue_sha256": digest(r, nw)}


def main():
    from _support import gmp_version, numpy_primes, oracle_sieve, survivors
    cfg = json.loads((HERE / "configs.json").read_text())["1"]
    a, b, c, K, J = cfg["a"], cfg["b"], cfg["c"], cfg["K"], cfg["J"]
    primes = numpy_primes(cfg["prime_limit"])
    bits = oracle_sieve(primes, a, b, c, K, J)
    surv = survivors(bits, K, J)
    packed = np.array([(k << 16) | j for k, j in surv], dtype="<u4")
    jobs = [(a, b, c, k, j, w) for k, j in surv[:cfg["mr_survivors"]] for w in (0, 1)]
    with ProcessPoolExecutor(max_workers=8) as ex:
        rows = list(ex.map(one, jobs))
    out = {"source": f"oracle/td_oracle.c sieve; GMP {gmp_version()} mpz_powm via oracle/mr_oracle.c",
           "a": a, "b": b, "c": c, "K": K, "J": J, "prime_limit": cfg["prime_limit"],
           "survivor_count": len(surv), "first_survivors": [list(x) for x in surv[:64]],
           "survivors_sha256": hashlib.sha256(packed.tobytes()).hexdigest(),
           "mr_base": 2, "mr_rows": rows}
    (HERE / "pipeline_config1.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"{len(surv)} survivors; {len(rows)} residues")


if __name__ == "__main__":
    main()
