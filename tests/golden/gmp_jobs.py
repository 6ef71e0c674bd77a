"""Shared runner for the GMP fixture generators (test infrastructure only).

A job is one strong-probable-prime chain on one bigint: (a, b, c, k, j,
which) and a list of bases, run in order with GMP mpz_powm through
oracle/mr_oracle.c, stopping at the first failing base unless `all_bases`.
Each finished job is appended as one JSON line to a ``.partial.jsonl`` file
beside the fixture, so an interrupted run resumes where it stopped (one
50 000-digit mpz_powm takes about three minutes on one core).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor, as_completed
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))


def digest(x: int, n_words: int) -> str:
    """SHA-256 of x as n_words little-endian uint32 words."""
    return hashlib.sha256(x.to_bytes(4 * n_words, "little")).hexdigest()


def job_key(cd, bases, all_bases) -> str:
    return json.dumps([list(cd), list(bases), bool(all_bases)])


def run_one(args):
    cd, bases, all_bases = args
    os.nice(19)
    from _support import gmp_build, gmp_sprp
    n = gmp_build(*cd)
    nw = (n.bit_length() + 31) // 32
    out = {"cand": list(cd), "n_words": nw, "n_bits": n.bit_length(),
           "n_sha256": digest(n, nw), "bases": [], "residue_sha256": [], "pass": []}
    verdict, rounds = 1, 0
    for base in bases:
        v, r, s = gmp_sprp(n, base)
        out["s"] = s
        out["bases"].append(base)
        out["residue_sha256"].append(digest(r, nw))
        out["pass"].append(bool(v == 1))
        if verdict:
            rounds += 1
        if v != 1:
            verdict = 0
            if not all_bases:
                break
    out["verdict"], out["rounds"] = verdict, rounds
    return job_key(cd, bases, all_bases), out


def run_jobs(jobs, partial: Path, workers: int):
    """jobs: list of (cand, bases, all_bases); returns {key: row}."""
    done = {}
    if partial.exists():
        for line in partial.read_text().splitlines():
            if line.strip():
                k, row = json.loads(line)
                done[k] = row
    todo = [j for j in jobs if job_key(*j) not in done]
    print(f"{len(jobs)} jobs, {len(jobs) - len(todo)} cached, {len(todo)} to run", flush=True)
    if todo:
        with ProcessPoolExecutor(max_workers=workers) as ex, partial.open("a") as f:
            futs = [ex.submit(run_one, j) for j in todo]
            for i, fu in enumerate(as_completed(futs)):
                k, row = fu.result()
                done[k] = row
                f.write(json.dumps([k, row]) + "\n")
                f.flush()
                print(f"  {i + 1}/{len(todo)} {row['cand']} rounds={row['rounds']}", flush=True)
    return done
