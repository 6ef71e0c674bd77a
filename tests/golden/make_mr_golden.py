"""Writes tests/golden/mr_50k_residues.json: GMP strong-probable-prime
residues a^d mod N (N - 1 = 2^s d) for the first candidates of workload 3
(a = 50 000 digits) and their decimal reversals, as SHA-256 digests of the
little-endian uint32 words, with GMP's verdict.  One GMP mpz_powm at this
size takes about three minutes, so the pairs run in parallel processes.

    python tests/golden/make_mr_golden.py [n_candidates] [bases...]
"""
import hashlib
import json
import sys
from concurrent.futures import ProcessPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))


def digest(x: int, n_words: int) -> str:
    return hashlib.sha256(x.to_bytes(4 * n_words, "little")).hexdigest()


def one(args):
    from _support import gmp_build, gmp_sprp
    a, b, c, k, j, which, base = args
    n = gmp_build(a, b, c, k, j, which)
    nw = (n.bit_length() + 31) // 32
    v, r, s = gmp_sprp(n, base)
    return {"a": a, "b": b, "c": c, "k": k, "j": j, "which": which, "base": base,
            "n_words": nw, "n_sha256": digest(n, nw), "s": s, "pass": bool(v == 1),
            "residue_sha256": digest(r, nw)}


def main():
    from _support import gmp_version
    ncand = int(sys.argv[1]) if len(sys.argv) > 1 else 4
    bases = [int(x) for x in sys.argv[2:]] or [2]
    cfg = json.loads((HERE / "configs.json").read_text())["3"]["candidates"][:ncand]
    jobs = [tuple(cd) + (w, base) for cd in cfg for w in (0, 1) for base in bases]
    with ProcessPoolExecutor(max_workers=8) as ex:
        rows = list(ex.map(one, jobs))
    out = {"source": f"GMP {gmp_version()} mpz_powm via oracle/mr_oracle.c",
           "digest": "sha256 of a^d mod N as n_words little-endian uint32 words",
           "rows": rows}
    (HERE / "mr_50k_residues.json").write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {len(rows)} rows")


if __name__ == "__main__":
    main()
