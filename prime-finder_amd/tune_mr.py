"""Per-bucket tuner of the Miller-Rabin scheduler, run on the device.

The reference fixes its Miller-Rabin kernel parameters when it is configured
(params.cmake:6-64, miller_rabin_gpu/config.h.in:11-95).  Here the transform
code of each size bucket is fixed at compile time (its passes, load batches
and the two-half split at 2^15 were chosen by measurement), and the
scheduler's runtime choices are tuned per bucket on the device itself:

  chunk         squaring steps per chunk, i.e. per HIP graph launch of the
                bucket's pool: fewer host round trips against a coarser
                admission grain;
  slots_per_cu  pool slots per compute unit: 1, or 2 where two workgroups
                of the bucket's transform fit one CU's LDS (up to 2^13).

For every bucket and setting it fills one pool with synthetic candidates of
the bucket's size (10^a - k 10^b - j 10^c - 1, base 2), lets the scheduler
run a few chunks, and times whole chunks between blocking pumps: slot-steps
per second, i.e. throughput.  The best setting per bucket (a setting within
1 % of the default counts as the default) goes into a JSON table with
every measurement beside it; MrScheduler(tuning=path) and
mr_sched_load_tuning() apply it.

    python prime-finder_amd/tune_mr.py [--out PATH] [--buckets 11,12,13,14,15]
        [--chunks 64,128,256,512,1024] [--slots 1,2] [--seconds 1.5]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import math
import random
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent

# decimal digits a of a candidate in each transform bucket (a 24-bit-limb
# number of L limbs needs 2^log_len >= 2 L + 2 points): near the top of the
# smaller buckets, so that a round (one squaring per bit) outlasts the
# measurement; the workloads' own sizes for 2^14 and 2^15
BUCKET_DIGITS = {11: 7000, 12: 14000, 13: 28000, 14: 50000, 15: 100000}
DEFAULT = {"chunk": 256, "slots_per_cu": 1}


def _pf():
    if "pf_amd" in sys.modules:
        return sys.modules["pf_amd"]
    spec = importlib.util.spec_from_file_location("pf_amd", HERE / "pf_amd.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pf_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def bucket_of_digits(digits: int) -> int:
    """Transform bucket (log2 of the length) of an a-digit candidate."""
    bits = math.ceil(digits * math.log2(10))
    limbs = -(-bits // 24)
    return max(11, math.ceil(math.log2(2 * limbs + 2)))


def measure(pf, log_len: int, chunk: int, spc: int, seconds: float, seed: int = 1) -> dict:
    """Slot-steps per second of one full pool of the bucket under (chunk,
    slots per CU): whole chunks timed between blocking pumps, after one
    warm-up chunk (graph capture, admission), up to `seconds` or until the
    rounds near their end (a chunk that may reach it is not counted)."""
    cus = pf.device_info()["cus"]
    n = spc * cus
    a = BUCKET_DIGITS[log_len]
    bits = int(a * math.log2(10)) - 1           # squarings of a round, at least
    rng = random.Random(seed)
    cands = [(a, a // 2, 100, rng.randrange(1, 4097), rng.randrange(1, 4097), 0)
             for _ in range(n)]
    s = pf.MrScheduler([2], max_active=n)
    try:
        s.set_bucket(log_len, chunk, n)
        s.submit(cands)
        s.pump(True)                      # admission
        while s.stats().chunks < 1:
            s.pump(True)                  # first chunk: graph capture
        st0, t0 = s.stats(), time.perf_counter()
        st1, t1 = st0, t0
        while st1.chunks - st0.chunks < 3 or t1 - t0 < seconds:
            s.pump(True)
            st, t = s.stats(), time.perf_counter()
            if st.chunks == st1.chunks or st.squarings > n * (bits - chunk):
                break                     # no chunk ran, or this one may reach the rounds' end
            st1, t1 = st, t
    finally:
        s.close()
    chunks = st1.chunks - st0.chunks
    if chunks < 1:
        raise RuntimeError(f"2^{log_len}: rounds end before a full chunk can be timed")
    steps = st1.squarings - st0.squarings
    return {"log_len": log_len, "chunk": chunk, "slots_per_cu": spc, "slots": n,
            "digits": a, "chunks": int(chunks), "steps_per_s": steps / (t1 - t0)}


def choose(rows: list[dict]) -> dict:
    """Best throughput; the default setting wins ties within 1 %."""
    best = max(rows, key=lambda r: r["steps_per_s"])
    dflt = [r for r in rows if r["chunk"] == DEFAULT["chunk"]
            and r["slots_per_cu"] == DEFAULT["slots_per_cu"]]
    if dflt and dflt[0]["steps_per_s"] >= 0.99 * best["steps_per_s"]:
        best = dflt[0]
    out = {k: best[k] for k in ("log_len", "chunk", "slots_per_cu", "digits", "steps_per_s")}
    if dflt:
        out["vs_default"] = best["steps_per_s"] / dflt[0]["steps_per_s"]
    return out


def tune(buckets=(11, 12, 13, 14, 15), chunks=(64, 128, 256, 512, 1024), slots=(1, 2),
         seconds: float = 1.5, out: Path | None = None, log=print) -> dict:
    pf = _pf()
    info = pf.device_info()
    meas, table = [], []
    for lg in buckets:
        rows = []
        for spc in slots:
            for ch in chunks:
                r = measure(pf, lg, ch, spc, seconds)
                log(f"[tune] 2^{lg} chunk {ch:5d} slots/CU {spc}: {r['steps_per_s']:.4g} slot-steps/s")
                rows.append(r)
        meas += rows
        table.append(choose(rows))
        log(f"[tune] 2^{lg}: chunk {table[-1]['chunk']}, slots/CU {table[-1]['slots_per_cu']}")
    doc = {"device": info["name"], "arch": info["arch"], "cus": info["cus"], "bases": [2],
           "objective": "slot-steps per second of one full pool (throughput)",
           "seconds_per_setting": seconds, "buckets": table, "measurements": meas}
    if out is not None:
        Path(out).write_text(json.dumps(doc, indent=1) + "\n")
    return doc


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(HERE / "mr_tuning.json"))
    ap.add_argument("--buckets", default="11,12,13,14,15")
    ap.add_argument("--chunks", default="64,128,256,512,1024")
    ap.add_argument("--slots", default="1,2")
    ap.add_argument("--seconds", type=float, default=1.5)
    args = ap.parse_args()
    ints = lambda s: tuple(int(x) for x in s.split(","))
    doc = tune(ints(args.buckets), ints(args.chunks), ints(args.slots), args.seconds,
               Path(args.out), log=lambda m: print(m, flush=True))
    print(json.dumps(doc["buckets"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
