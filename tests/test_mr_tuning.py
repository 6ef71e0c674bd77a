"""Per-bucket tuning of the Miller-Rabin scheduler.

CPU: the library's parser of tuning tables (mr_tuning_read) reads what
prime-finder_amd/tune_mr.py writes and rejects malformed tables; the
tuner's digit counts land in the buckets it names.
GPU: the tuner measures a small grid on the device and writes a table; a
scheduler that loads it, and one forced to a non-default chunk and two slots
per compute unit, give every candidate the verdict and round count GMP
gives."""
import importlib.util
import json
import random

import pytest

from _support import PKG, gmp_build, gmp_sprp

BASES = [2, 3, 5, 7]


def _tuner():
    spec = importlib.util.spec_from_file_location("tune_mr", PKG / "tune_mr.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _table(buckets, **extra):
    doc = {"device": "x", "buckets": buckets,
           "measurements": [{"log_len": 11, "chunk": 1, "slots_per_cu": 9, "steps_per_s": 1.0}]}
    doc.update(extra)
    return doc


def test_tuning_table_roundtrip(pf, tmp_path):
    rows = [{"log_len": 11, "chunk": 512, "slots_per_cu": 2, "digits": 6000, "steps_per_s": 1e9},
            {"log_len": 14, "chunk": 128, "slots_per_cu": 1, "vs_default": 1.02}]
    p = tmp_path / "t.json"
    p.write_text(json.dumps(_table(rows), indent=1))
    got = pf.tuning_read(p)
    assert got == [{"log_len": 11, "chunk": 512, "slots_per_cu": 2},
                   {"log_len": 14, "chunk": 128, "slots_per_cu": 1}]
    p.write_text(json.dumps(_table([])))
    assert pf.tuning_read(p) == []


@pytest.mark.parametrize("bad", [
    {"device": "x"},                                                          # no buckets
    _table([{"log_len": 11, "chunk": 64}]),                                   # field missing
    _table([{"log_len": 10, "chunk": 64, "slots_per_cu": 1}]),                # bucket too small
    _table([{"log_len": 16, "chunk": 64, "slots_per_cu": 1}]),                # bucket too large
    _table([{"log_len": 12, "chunk": -1, "slots_per_cu": 1}]),                # negative chunk
    _table([{"log_len": 12, "chunk": 64, "slots_per_cu": 1, "x": [1, 2]}]),   # nested value
])
def test_tuning_table_rejects_malformed(pf, tmp_path, bad):
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        pf.tuning_read(p)
    with pytest.raises(ValueError):
        pf.tuning_read(tmp_path / "missing.json")


def test_tuner_digits_match_buckets():
    t = _tuner()
    for lg, d in t.BUCKET_DIGITS.items():
        assert t.bucket_of_digits(d) == lg
    # the default setting wins a tie within 1 %
    rows = [{"log_len": 12, "chunk": 256, "slots_per_cu": 1, "digits": 12000, "steps_per_s": 100.0},
            {"log_len": 12, "chunk": 512, "slots_per_cu": 1, "digits": 12000, "steps_per_s": 100.5},
            {"log_len": 12, "chunk": 64, "slots_per_cu": 2, "digits": 12000, "steps_per_s": 99.0}]
    assert t.choose(rows)["chunk"] == 256
    rows[2]["steps_per_s"] = 110.0
    c = t.choose(rows)
    assert (c["chunk"], c["slots_per_cu"]) == (64, 2) and abs(c["vs_default"] - 1.1) < 1e-9


def _gmp_outcome(cd):
    n = gmp_build(*cd)
    for i, b in enumerate(BASES):
        if gmp_sprp(n, b)[0] != 1:
            return 0, i + 1
    return 1, len(BASES)


@pytest.mark.gpu
def test_tuned_scheduler_matches_gmp(gpu, tmp_path):
    t = _tuner()
    out = tmp_path / "tuning.json"
    print("[tune] bucket 2^11, chunks 64/128, 1-2 slots per CU", flush=True)
    doc = t.tune(buckets=(11,), chunks=(64, 128), slots=(1, 2), seconds=0.3, out=out,
                 log=lambda m: print(m, flush=True))
    assert len(doc["measurements"]) == 4
    assert all(m["steps_per_s"] > 0 for m in doc["measurements"])
    b = doc["buckets"][0]
    assert gpu.tuning_read(out) == [{"log_len": 11, "chunk": b["chunk"],
                                     "slots_per_cu": b["slots_per_cu"]}]
    rng = random.Random(11)
    cands = [(4000, rng.randrange(200, 3800), 100, rng.randrange(1, 4097),
              rng.randrange(1, 4097), rng.choice((0, 1))) for _ in range(12)]
    want = [_gmp_outcome(cd) for cd in cands]
    cus = gpu.device_info()["cus"]
    for label, setup in (("tuned", lambda s: None),
                         ("chunk 64, 2 slots/CU", lambda s: s.set_bucket(11, 64, 2 * cus))):
        s = gpu.MrScheduler(BASES, tuning=out if label == "tuned" else None)
        setup(s)
        s.submit(cands, tags=list(range(len(cands))))
        got = {r.tag: (r.verdict, r.rounds) for r in s.run_all()}
        st = s.stats()
        with pytest.raises(ValueError):      # the bucket's pool is built
            s.set_bucket(11, 128, cus)
        s.close()
        print(f"[tune] {label}: {st.chunks} chunks", flush=True)
        assert [got[i] for i in range(len(cands))] == want, label
