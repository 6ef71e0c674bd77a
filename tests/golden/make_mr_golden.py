"""Writes tests/golden/mr_50k_residues.json: GMP strong-probable-prime
residues a^d mod N (N - 1 = 2^s d) for the uniform-size workload (128
candidates with a = 50 000 digits and their 128 decimal reversals,
tests/golden/configs.json workload 3), as SHA-256 digests of the
little-endian uint32 words, with GMP's verdict per base.

Base 2 is computed for all 256 bigints; bases 3, 5 and 7 as well for the
bigints of the first `n_all` candidates (every base computed whatever the
earlier ones gave, as the reference runner does for every witness,
miller_rabin_runner.cu:305-354).  One GMP mpz_powm at this size takes about
three minutes on one core, so the jobs run in parallel processes and
resume from tests/golden/mr_50k_residues.partial.jsonl.

    python tests/golden/make_mr_golden.py [n_candidates] [n_all] [workers]
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

from gmp_jobs import job_key, run_jobs  # noqa: E402

ALL_BASES = (2, 3, 5, 7)


def main():
    ncand = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    n_all = int(sys.argv[2]) if len(sys.argv) > 2 else 4
    workers = int(sys.argv[3]) if len(sys.argv) > 3 else 7
    cfg = json.loads((HERE / "configs.json").read_text())["3"]["candidates"][:ncand]
    jobs = []
    for i, cd in enumerate(cfg):
        for w in (0, 1):
            if i < n_all:
                jobs.append((tuple(cd) + (w,), ALL_BASES, True))
            else:
                jobs.append((tuple(cd) + (w,), (2,), False))
    # the four-base chains first: they are the longest jobs
    jobs.sort(key=lambda j: -len(j[1]))
    done = run_jobs(jobs, HERE / "mr_50k_residues.partial.jsonl", workers)
    from _support import gmp_version
    rows = []
    for cd, bases, allb in sorted(jobs, key=lambda j: (cfg.index(list(j[0][:5])), j[0][5])):
        r = done[job_key(cd, bases, allb)]
        a, b, c, k, j, w = cd
        for base, dig, ok in zip(r["bases"], r["residue_sha256"], r["pass"]):
            rows.append({"a": a, "b": b, "c": c, "k": k, "j": j, "which": w, "base": base,
                         "n_words": r["n_words"], "n_sha256": r["n_sha256"], "s": r["s"],
                         "pass": ok, "residue_sha256": dig})
    out = {"source": f"GMP {gmp_version()} mpz_powm via oracle/mr_oracle.c",
           "digest": "sha256 of a^d mod N as n_words little-endian uint32 words",
           "all_bases_candidates": n_all,
           "rows": rows}
    (HERE / "mr_50k_residues.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} rows")


if __name__ == "__main__":
    main()
