"""Miller-Rabin at the workloads' real sizes on the GPU, against GMP
fixtures made in the oracle's container (tests/golden/make_mr_golden.py,
tests/golden/make_sched_golden.py):

* uniform batch (50 000 digits, 128 N + 128 rev N): every base-2 residue
  a^d mod N in one launch, and every (number, base) test of bases 2, 3, 5,
  7 of a subset in one launch, the way the reference runner computes all
  witnesses (miller_rabin_runner.cu:305-354);
* uneven batch (a log-uniform in [50 000, 100 000]): a subset spanning the
  2^14 and 2^15 transform buckets through the continuous-batching scheduler
  with its default slots and chunk (captured squaring graphs, both big
  pools), verdicts and round counts as miller_rabin_runner.cu:281-359 with
  retirement;
* serving (a uniform in [50 000, 100 000]): a seeded Poisson arrival of a
  request subset with non-blocking pumps, each verdict equal to GMP's;
* alone vs in the mix: smaller candidates run one per scheduler and all
  together give the same verdicts and round counts.
"""
import hashlib
import json
import random
import time

import pytest

from _support import GOLDEN, gmp_build, gmp_sprp

pytestmark = pytest.mark.gpu

BASES = [2, 3, 5, 7]


def say(*a):
    print("[mr-work]", *a, flush=True)


def sha(x: int, nw: int) -> str:
    return hashlib.sha256(x.to_bytes(4 * nw, "little")).hexdigest()


def _golden(name):
    path = GOLDEN / name
    if not path.exists():
        pytest.fail(f"{path} missing: run tests/golden/{name.replace('.json', '')} generator")
    return json.loads(path.read_text())


def test_uniform_batch_base2_all_256(gpu):
    """All 256 bigints of the uniform batch (a = 50 000 digits) built on the
    device, one base-2 round in one launch: N and every residue equal GMP's
    digests."""
    rows = [g for g in _golden("mr_50k_residues.json")["rows"] if g["base"] == 2]
    say("uniform batch:", len(rows), "bigints, base 2, one launch")
    cands = [(g["a"], g["b"], g["c"], g["k"], g["j"], g["which"]) for g in rows]
    ctx = gpu.MrContext(candidates=cands)
    mods = ctx.moduli()
    t0 = time.perf_counter()
    alive, res = ctx.run([2])
    say("uniform batch: round done in", round(time.perf_counter() - t0, 1), "s")
    for i, (g, n, ok, r) in enumerate(zip(rows, mods, alive, res)):
        nw = g["n_words"]
        assert sha(n, nw) == g["n_sha256"], g
        assert ctx.s[i] == g["s"]
        assert sha(r[0], nw) == g["residue_sha256"], g
        assert ok == g["pass"]
    ctx.close()


def test_uniform_batch_all_witnesses_one_launch(gpu):
    """Every (number, base) pair of bases 2, 3, 5, 7 on the fixture's subset,
    one slot per pair in a single launch (mr_run_per_slot): each residue and
    each round's verdict equals GMP's, whatever the earlier bases gave."""
    rows = [g for g in _golden("mr_50k_residues.json")["rows"]]
    keyed = {}
    for g in rows:
        keyed.setdefault((g["a"], g["b"], g["c"], g["k"], g["j"], g["which"]), {})[g["base"]] = g
    subset = [(cd, per) for cd, per in keyed.items() if set(per) == set(BASES)]
    assert len(subset) >= 8, "fixture lacks the all-bases subset"
    slots = [(cd, b, per[b]) for cd, per in subset for b in BASES]
    say("all witnesses:", len(subset), "bigints x", len(BASES), "bases =", len(slots), "slots")
    ctx = gpu.MrContext(candidates=[s[0] for s in slots])
    ok, res = ctx.run_per_slot([s[1] for s in slots])
    for (cd, b, g), o, r in zip(slots, ok, res):
        assert sha(r, g["n_words"]) == g["residue_sha256"], (cd, b)
        assert o == g["pass"], (cd, b)
    ctx.close()


def test_run_all_bases_without_retirement(gpu):
    """mr_run_ex with MR_RUN_ALL_BASES: residues for every base even after a
    failure (the retiring mode leaves those rows zero), verdict = AND."""
    rng = random.Random(5)
    nums = [gmp_build(1500, rng.randrange(600, 1400), rng.randrange(1, 500),
                      rng.randrange(1, 4097), rng.randrange(1, 4097), rng.choice((0, 1)))
            for _ in range(6)]
    nums += [2 ** 1279 - 1, 3215031751, 2047 * (2 ** 700 + 1) + 2]
    nums = [n | 1 for n in nums]
    ctx = gpu.MrContext(nums)
    alive, res, passes = ctx.run(BASES, all_bases=True, passes=True)
    alive_r, res_r, passes_r = ctx.run(BASES, passes=True)
    for i, n in enumerate(nums):
        want = [gmp_sprp(n, b) for b in BASES]
        for bi, (v, r, _) in enumerate(want):
            assert res[i][bi] == r, (n, BASES[bi])
            assert passes[i][bi] == (v == 1)
        assert alive[i] == all(v == 1 for v, _, _ in want)
        assert alive_r[i] == alive[i]
        first_fail = next((bi for bi, (v, _, _) in enumerate(want) if v != 1), None)
        for bi in range(len(BASES)):
            if first_fail is None or bi <= first_fail:
                assert res_r[i][bi] == want[bi][1]
            else:
                assert res_r[i][bi] == 0
            assert passes_r[i][bi] == (first_fail is None or bi < first_fail)
    ctx.close()


def _check_sched_rows(res, rows):
    got = {r.tag: (r.verdict, r.rounds) for r in res}
    assert sorted(got) == list(range(len(rows)))
    for t, g in enumerate(rows):
        assert got[t] == (g["verdict"], g["rounds"]), g["cand"]


def test_uneven_batch_subset_big_pools(gpu):
    """The uneven-batch subset (8 bigints in the 2^14 bucket, 8 in 2^15)
    through the scheduler with its default slots and chunk length: both big
    pools and their captured chunk graphs run; verdicts and rounds equal
    GMP's, and the device-built numbers equal GMP's digests."""
    gold = _golden("mr_sched_golden.json")
    rows = gold["uneven_batch"]
    assert gold["bases"] == BASES
    assert sorted({g["bucket"] for g in rows}) == [14, 15]
    cands = [tuple(g["cand"]) for g in rows]
    for lg in (14, 15):
        part = [i for i, g in enumerate(rows) if g["bucket"] == lg]
        ctx = gpu.MrContext(candidates=[cands[i] for i in part])
        assert ctx.log_len == lg
        for i, n in zip(part, ctx.moduli()):
            assert sha(n, rows[i]["n_words"]) == rows[i]["n_sha256"]
        ctx.close()
    say("uneven batch:", len(cands), "bigints, a =", min(c[0] for c in cands), "..",
        max(c[0] for c in cands), "digits")
    s = gpu.MrScheduler(BASES)
    t0 = time.perf_counter()
    s.submit(cands, tags=list(range(len(cands))))
    res = s.run_all()
    st = s.stats()
    say("uneven batch: done in", round(time.perf_counter() - t0, 1), "s;", st.chunks, "chunks,",
        st.n_pools, "pools,", st.hbm_bytes >> 20, "MiB")
    _check_sched_rows(res, rows)
    assert st.n_pools == 2 and st.completed == len(cands)
    s.close()


def test_serving_poisson_subset(gpu):
    """Serving: the request subset arrives as a seeded Poisson stream while
    earlier requests run (non-blocking pumps, both big pools, the 64 GB
    admission budget); every verdict and round count equals GMP's, and time
    to verdict is reported."""
    gold = _golden("mr_sched_golden.json")
    rows = gold["serving"]
    rng = random.Random(0x5E12)
    gaps = [rng.expovariate(1.0 / 0.6) for _ in rows]      # mean 0.6 s between arrivals
    s = gpu.MrScheduler(BASES, hbm_budget=64_000_000_000)
    t0 = time.perf_counter()
    due, nxt, res = 0.0, 0, []
    while nxt < len(rows) or len(res) < len(rows):
        now = time.perf_counter() - t0
        while nxt < len(rows) and now >= due:
            s.submit([tuple(rows[nxt]["cand"])], tags=[nxt])
            nxt += 1
            due += gaps[nxt - 1]
        s.pump(False)
        res += s.results()
        if nxt >= len(rows):
            s.pump(True)
            res += s.results()
        else:
            time.sleep(0.02)
    _check_sched_rows(res, rows)
    ttv = sorted(r.t_done - r.t_submit for r in res)
    say("serving:", len(res), "requests; time to verdict p50", round(ttv[len(ttv) // 2], 1),
        "s, max", round(ttv[-1], 1), "s")
    st = s.stats()
    assert st.n_pools == 2
    s.close()


def test_alone_equals_mix(gpu):
    """Each candidate run alone (its own scheduler) and all of them together
    (one scheduler, requests arriving while others run) give the same
    verdict and round count; primes of the form included."""
    rng = random.Random(21)
    cands = [(1000, 500, 100, 32, 5, 0), (1000, 500, 100, 47, 57, 0)]
    for d in (800, 1500, 3000, 6000):
        for _ in range(2):
            cands.append((d, rng.randrange(d // 3, d - 10), rng.randrange(1, d // 4),
                          rng.randrange(1, 4097), rng.randrange(1, 4097), rng.choice((0, 1))))
    alone = []
    for cd in cands:
        s = gpu.MrScheduler(BASES)
        s.submit([cd], tags=[0])
        r = s.run_all()
        alone.append((r[0].verdict, r[0].rounds))
        s.close()
    s = gpu.MrScheduler(BASES, max_active=5, chunk=64)
    res = []
    for i, cd in enumerate(cands):
        s.submit([cd], tags=[i])
        s.pump(False)
        res += s.results()
    res += s.run_all()
    mix = {r.tag: (r.verdict, r.rounds) for r in res}
    assert [mix[i] for i in range(len(cands))] == alone
    assert alone[0] == (1, 4) and alone[1] == (1, 4)
    s.close()
