"""Writes tests/golden/td_full_4096.json: the full-scale trial-division
workload (tests/golden/configs.json workload 2: 8 seeded triples, K = J =
4096, every prime below 10^8, 2 included) sieved by the oracle's
restatement of trial_div_cpu.c (oracle/td_oracle.c td_oracle_sieve_mt),
one SHA-256 of the 2 MiB bitmap (524 288 little-endian uint32 words) per
triple, with the number of eliminated pairs and a second digest of the
survivor list.  The prime list is an independent numpy sieve.  About a
minute per triple on 8 cores.

    python tests/golden/make_td_full_golden.py [threads]
"""
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from _support import numpy_primes, oracle_sieve, popcount  # noqa: E402


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 0
    cfg = json.loads((HERE / "configs.json").read_text())["2"]
    K, J, limit = cfg["K"], cfg["J"], cfg["prime_limit"]
    primes = numpy_primes(limit)
    rows = []
    for (a, b, c) in cfg["triples"]:
        t0 = time.time()
        bits = oracle_sieve(primes, a, b, c, K, J, threads=threads)
        flat = np.unpackbits(bits.view(np.uint8), bitorder="little")[: K * J]
        surv = np.nonzero(flat == 0)[0].astype(np.uint32)
        rows.append({"a": a, "b": b, "c": c,
                     "bits_sha256": hashlib.sha256(bits.tobytes()).hexdigest(),
                     "eliminated": popcount(bits),
                     "survivors": int(len(surv)),
                     "survivor_index_sha256": hashlib.sha256(surv.tobytes()).hexdigest()})
        print(f"({a},{b},{c}): {rows[-1]['survivors']} survivors, {time.time() - t0:.1f} s",
              flush=True)
    out = {"source": "oracle/td_oracle.c td_oracle_sieve_mt (restates trial_div_cpu.c:32-111)",
           "K": K, "J": J, "primes": f"all {len(primes)} primes below {limit}, 2 included",
           "digest": "sha256 of the bitmap as little-endian uint32 words; survivor index "
                     "= (k-1)*J + (j-1) as little-endian uint32, ascending",
           "triples": rows}
    (HERE / "td_full_4096.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} triples")


if __name__ == "__main__":
    main()
