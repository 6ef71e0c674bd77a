"""Writes tests/golden/mr_sched_golden.json: GMP verdicts for subsets of the
uneven-batch workload (512 N + 512 rev(N), a log-uniform in [50 000,
100 000] digits) and of the serving workload (10 000 requests, a uniform in
[50 000, 100 000]), both from tests/golden/configs.json.

Per bigint: (verdict, rounds) of per-base Miller-Rabin on bases 2, 3, 5, 7
with retirement at the first failing base (miller_rabin_runner.cu:281-359
semantics), the SHA-256 of N and of each computed residue a^d mod N (GMP
mpz_powm through oracle/mr_oracle.c), and the transform bucket the library
puts it in.  Subsets, chosen so a GPU test finishes in about a minute:

  uneven batch  the 4 candidates of smallest a in the 2^14 bucket and the 4
                of smallest a in the 2^15 bucket, N and rev(N) each (16);
  serving       the first 12 requests (in arrival order) of the 2^14 bucket
                and the first 4 of the 2^15 bucket with a <= 61 000 (16).

    python tests/golden/make_sched_golden.py [workers]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

from gmp_jobs import run_jobs  # noqa: E402

BASES = [2, 3, 5, 7]


def bucket(a: int) -> int:
    """log2 of the transform length the library gives an a-digit candidate
    (24-bit limbs, M >= 2L + 2, L = bits / 24 + 1)."""
    bits = int(a * 3.3219280948873623) + 2
    L = max(4, bits // 24 + 1)
    logm = 11
    while (1 << logm) < 2 * L + 2:
        logm += 1
    return logm


def select():
    cfg = json.loads((HERE / "configs.json").read_text())
    w4 = cfg["4"]["candidates"]
    order = sorted(range(len(w4)), key=lambda i: (w4[i][0], i))
    b14 = [i for i in order if bucket(w4[i][0]) == 14][:4]
    b15 = [i for i in order if bucket(w4[i][0]) == 15][:4]
    uneven = [{"index": i, "cand": list(w4[i]) + [w], "bucket": bucket(w4[i][0])}
              for i in b14 + b15 for w in (0, 1)]
    w5 = cfg["5"]["requests"]
    s14 = [i for i, r in enumerate(w5) if bucket(r[0]) == 14][:12]
    s15 = [i for i, r in enumerate(w5) if bucket(r[0]) == 15 and r[0] <= 61000][:4]
    serving = [{"index": i, "cand": list(w5[i]), "bucket": bucket(w5[i][0])}
               for i in sorted(s14 + s15)]
    return uneven, serving


def main():
    workers = int(sys.argv[1]) if len(sys.argv) > 1 else 7
    uneven, serving = select()
    jobs = [(tuple(r["cand"]), tuple(BASES), False) for r in uneven + serving]
    # longest first, for packing
    jobs.sort(key=lambda j: -j[0][0])
    done = run_jobs(jobs, HERE / "mr_sched_golden.partial.jsonl", workers)
    from gmp_jobs import job_key
    from _support import gmp_version
    for r in uneven + serving:
        row = done[job_key(tuple(r["cand"]), tuple(BASES), False)]
        r.update({k: row[k] for k in ("n_words", "n_bits", "n_sha256", "s", "verdict",
                                      "rounds", "residue_sha256")})
    out = {"source": f"GMP {gmp_version()} mpz_powm via oracle/mr_oracle.c",
           "bases": BASES,
           "cand_fields": "a b c k j which (0 = N, 1 = decimal reversal of N)",
           "digest": "sha256 of the little-endian uint32 words (n_words of them)",
           "uneven_batch": uneven, "serving": serving}
    (HERE / "mr_sched_golden.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(uneven)} + {len(serving)} rows")


if __name__ == "__main__":
    main()
