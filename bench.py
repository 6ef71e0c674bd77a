"""Benchmark of the Phase-1 trial-division hot path on one MI355X.

Workload (the full-scale trial-division configuration): eight seeded (a,b,c)
triples (tests/golden/configs.json, workload 2), k and j in 1..4096
(16 777 216 pairs per triple), all 5 761 455 primes below 10^8.  One step
is one pass over the eight triples: each bitmap is cleared and sieved by
every prime.  Primes and bitmaps are resident in HBM before the timed region.

Prints one JSON line:  value = pair*primes/s over the whole step
(8 * 4096^2 * 5 761 455 / seconds per step).  "roofline" rates the dominant
kernel (the lattice search of the primes from trial_div_lattice_min() up;
the row kernel and the split k loop of the primes below run beside it on a
second stream) from its own HIP-event durations on the stream it runs on,
against its algorithmic INT32 ops (units counted on the host, ops per unit
in TD_OPS_PER_UNIT);
"cpu_baseline" times the CPU restatement of the reference's trial_div_cpu.c
on up to eight processes, one triple each, on a bounded sample of the
workload's primes.
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
K = J = 4096
DJ = 4
# Algorithmic INT32 ops per unit of work of the per-prime kernels for primes
# >= 2^16 (the units are counted exactly on the host by trial_div_work, with
# the lattice kernel's own per-prime routine):
#   modmul  32-bit Montgomery multiplication: 4 multiply-class + 4 add-class;
#   modmul_lazy  the same without the final subtraction (the powers of ten
#           of the setup, primes below 2^30): 4 multiply-class + 2 add-class;
#   level   one continued-fraction level of a first-hit search: two unsigned
#           divisions by the same divisor (10 ops each: reciprocal estimate,
#           multiply-back, correct) and 26 compare/select/add ops (the
#           search-window bound included);
#   query   start and finish of a first-hit search (residue, bound check,
#           next start): 10;
#   hit     k and j decoded from a hit and the bit marked: 12;
#   step    one direct step of the one-digit rev class: modular subtraction
#           and compare, 4;
#   gapstep one return-time step from a hit to the next (two compares, the
#           selects of the step and of the landing residue, two adds, the
#           end test): 8;
#   gapcf   one division step of the return-time computation (a run of the
#           slow continued fraction: the run length and the first value
#           below the window, two divisions by the same divisor at 10 each,
#           and 18 compare/select/multiply-add ops): 38;
#   kstep   one k of the split k loop (primes from 2^16 below the lattice
#           range, beside the lattice kernel on the second stream): two
#           residues advanced and compared, 8.
TD_OPS_PER_UNIT = {"modmul": 8, "modmul_lazy": 6, "level": 46, "query": 10, "hit": 12, "step": 4,
                   "gapstep": 8, "gapcf": 38, "kstep": 8}
INT32_VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12      # 39.32 Tops/s (256 CUs, wave64, 2.4 GHz)
# HBM bytes per launch of the lattice kernel on this workload, measured with
# rocprofv3 FETCH_SIZE and WRITE_SIZE (scaled by a 256 MiB copy kernel's
# readings: x2.0 and x1.0): 23.29 MB read (the prime list) + 478.65 MB
# written (memory-side atomic ORs of the hits of the primes from 2^17).
TD_TRAFFIC_BYTES = 23.29e6 + 478.65e6
# Same for one batched squaring step of the 50 000-digit Miller-Rabin batch
# (256 bigints, 2^14 transforms): 91.7 MB read + 14.4 MB written.
MR_TRAFFIC_BYTES_PER_STEP = 91.7e6 + 14.4e6


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


_CPU_PRIMES = None      # set in the parent before the fork pool starts


def cpu_share() -> int:
    """Host cores this run may use: the process's affinity set, capped at
    16 (one GPU's share of a shared box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _td_cpu_worker(job):
    """One process: the oracle (single-threaded restatement of
    trial_div_cpu.c) on every stride-th prime of the whole list (small
    primes included), sized to ~budget_s after a calibration probe."""
    (a, b, c), offset, budget_s = job
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import oracle_sieve
    primes = _CPU_PRIMES
    probe = primes[offset:: max(1, len(primes) // 300)][:300]
    t0 = time.perf_counter()
    oracle_sieve(probe, a, b, c, K, J, threads=1)
    per_prime = (time.perf_counter() - t0) / len(probe)
    n = int(max(50, min(len(primes), budget_s / max(per_prime, 1e-9))))
    stride = max(1, len(primes) // n)
    sample = primes[offset % stride::stride][:n]
    t0 = time.perf_counter()
    oracle_sieve(sample, a, b, c, K, J, threads=1)
    dt = time.perf_counter() - t0
    return len(sample), stride, dt


def _mr_cpu_worker(job):
    """One process: GMP mpz_powm of base 2 over the top sample_bits bits of
    d for candidates of the workload in turn (~budget_s), scaled to the
    whole exponent: tests/hour of this core."""
    cands, sample_bits, budget_s = job
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import gmp_build, gmp_powm
    spent, scaled, done, bits = 0.0, 0.0, 0, 0
    for cd in cands:
        n = gmp_build(*cd)
        d = n - 1
        while d % 2 == 0:
            d //= 2
        bits = d.bit_length()
        e = d >> max(0, bits - sample_bits)
        t0 = time.perf_counter()
        gmp_powm(2, e, n)
        dt = time.perf_counter() - t0
        spent += dt
        scaled += dt * bits / sample_bits
        done += 1
        if spent >= budget_s:
            break
    return done, scaled, spent, bits


def _bpsw_cpu_worker(job):
    """One process: mpz_probab_prime_p(N, 5) (what the reference worker runs,
    miller_rabin.c:91) and a full base-2 mpz_powm on the same N."""
    cd = job
    sys.path.insert(0, str(ROOT / "tests"))
    import math
    from _support import gmp_build, gmp_powm, gmp_probab_prime, numpy_primes
    # a Phase-2 input survived trial division: take the first k from cd's
    # on whose N has no prime factor below 10^5, so BPSW runs in full
    small = math.prod(int(q) for q in numpy_primes(100000))
    a, b, c, k, j, w = cd
    while True:
        n = gmp_build(a, b, c, k, j, w)
        if math.gcd(n, small) == 1:
            break
        k += 1
    t0 = time.perf_counter()
    gmp_probab_prime(n, 5)
    t_bpsw = time.perf_counter() - t0
    d = n - 1
    while d % 2 == 0:
        d //= 2
    t0 = time.perf_counter()
    gmp_powm(2, d, n)
    return t_bpsw, time.perf_counter() - t0


def cpu_baselines(budget_s: float) -> dict:
    """The reference's CPU path timed on this host before the GPU is used,
    one process per core (fork pool): trial division on min(8, cores)
    processes, one triple each (the reference farms one triple per rank,
    master.c:587-615); Miller-Rabin as GMP mpz_powm on every core; and
    mpz_probab_prime_p (BPSW) at 10 000 digits beside a full mpz_powm on the
    same numbers, the ratio carrying it to 50 000 digits."""
    import multiprocessing as mp
    global _CPU_PRIMES
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import numpy_primes
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())
    _CPU_PRIMES = numpy_primes(cfg["2"]["prime_limit"])
    cores = cpu_share()
    ctx = mp.get_context("fork")
    out = {}
    triples = cfg["2"]["triples"]
    n_td = min(len(triples), cores)
    t0 = time.perf_counter()
    with ctx.Pool(n_td) as pool:
        rows = pool.map(_td_cpu_worker, [(tuple(triples[i]), i, budget_s) for i in range(n_td)])
    wall = time.perf_counter() - t0
    rate = sum(K * J * n / dt for n, _, dt in rows)
    out["td"] = {"value": rate, "unit": "pair*primes/s", "cores": n_td, "kind": "port",
                 "per_triple_1core_s": float(np.mean([dt * len(_CPU_PRIMES) / n
                                                      for n, _, dt in rows])),
                 "sample": (f"{n_td} processes, one triple of the workload each, K=J={K}; each "
                            f"sieves every {rows[0][1]}th prime of the {len(_CPU_PRIMES)} below "
                            f"10^8 (small primes included, {rows[0][0]} primes, "
                            f"{min(dt for _, _, dt in rows):.1f}-{max(dt for _, _, dt in rows):.1f} "
                            f"s per process); oracle/td_oracle.c restating trial_div_cpu.c, "
                            f"single-threaded per triple; {wall:.1f} s wall with calibration")}
    w3 = [tuple(x) + (w,) for x in cfg["3"]["candidates"] for w in (0, 1)]
    jobs = [(w3[i::cores], 4096, budget_s) for i in range(cores)]
    bp = [(10000, 5000 + 97 * i, 300 + i, 4001 + 2 * i, 17 + 4 * i, i & 1) for i in range(cores)]
    with ctx.Pool(cores) as pool:
        mr_rows = pool.map(_mr_cpu_worker, jobs)
        bpsw_rows = pool.map(_bpsw_cpu_worker, bp)
    tph = sum(3600.0 * done / scaled for done, scaled, _, _ in mr_rows)
    t_test_50k = float(np.mean([scaled / done for done, scaled, _, _ in mr_rows]))
    ratio = float(np.mean([tb / tp for tb, tp in bpsw_rows]))
    out["mr"] = {"value": tph, "unit": "tests/hour", "cores": cores, "kind": "port",
                 "s_per_test_1core": t_test_50k,
                 "sample": (f"{cores} processes; workload 3 bigints (50 000 digits), GMP "
                            f"mpz_powm of base 2 over the top 4096 of {mr_rows[0][3]} bits of "
                            f"d, scaled to the whole exponent (extrapolated); "
                            f"{sum(r[0] for r in mr_rows)} powms in "
                            f"{max(r[2] for r in mr_rows):.1f} s per process")}
    out["mr_reference_verdict"] = {
        "what": "mpz_probab_prime_p(N, 5) per bigint (BPSW, GMP 6.2), the call of "
                "miller_rabin.c:91 that worker.c:84 makes",
        "s_per_bigint_10k_digits": float(np.mean([tb for tb, _ in bpsw_rows])),
        "bpsw_over_powm_10k_digits": ratio,
        "s_per_bigint_50k_digits_extrapolated": ratio * t_test_50k,
        "bigints_per_hour_extrapolated": cores * 3600.0 / (ratio * t_test_50k),
        "cores": cores,
        "sample": f"{cores} numbers of 10 000 digits without prime factors below 10^5 (as "
                  f"Phase-1 survivors), one per process, each timed with mpz_probab_prime_p "
                  f"and a full base-2 mpz_powm; the ratio times the 50 000-digit powm time "
                  f"above"}
    return out


def mr_ops_per_squaring(logm: int) -> int:
    """INT32 VALU ops of one Montgomery squaring step on one candidate, by
    the transform-count model T(M) = (M/2) log2(M) 12 per transform and
    prime.  Per prime: r^2 and T_lo N' are products of M points (a forward
    and an inverse transform each, the other operand precomputed), m N is a
    wrapped product of M/2 points (its low half is known), so
    ops = 2 (4 T(M) + 2 T(M/2) + 20 M) + 50 M, the 20 M and 50 M being the
    pointwise products and the CRT/carry per coefficient, at this build's
    transform length M = 2^logm (24-bit limbs: 2^14 at 50 000 digits)."""
    M = 1 << logm
    T = (M // 2) * logm * 12
    Th = (M // 4) * (logm - 1) * 12
    return 2 * (4 * T + 2 * Th + 20 * M) + 50 * M


def mr_leg(pf, n_bits: int, digits: int = 50000) -> dict:
    """Batched Miller-Rabin at 50 000 digits (workload 3: 128 seeded
    candidates and their decimal reversals, built on the device): time the
    first n_bits squaring steps of a base-2 round over the whole batch with
    HIP events on the context's stream and scale to the full exponent
    length.  digits != 50000 keeps the workload's b, c, k, j with a = digits."""
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["3"]
    cands = [(digits,) + tuple(x[1:]) + (w,) for x in cfg["candidates"] for w in (0, 1)]
    t0 = time.perf_counter()
    ctx = pf.MrContext(candidates=cands)
    setup_s = time.perf_counter() - t0
    ctx.time_squarings(2, 64)                         # warm-up
    ms = ctx.time_squarings(2, n_bits)
    bits = int(digits * 3.3219280948873623) - 1       # squarings per test (bits of d)
    step_s = ms / 1e3 / n_bits
    test_s = step_s * bits
    ops = mr_ops_per_squaring(ctx.log_len) * len(cands)
    achieved = ops / step_s / 1e12
    res = {"workload": f"256 bigints (128 N + 128 rev N) at {digits} digits, base 2",
           "limbs": ctx.limbs, "transform_len": 1 << ctx.log_len,
           "us_per_squaring_batch": step_s * 1e6,
           "tests_per_hour_extrapolated": 3600.0 * len(cands) / test_s,
           "sample": f"{n_bits} of {bits} squaring steps timed (HIP events), scaled",
           "setup_s": setup_s,
           "roofline": {"bound": "valu", "achieved": achieved, "peak": INT32_VALU_PEAK_TOPS,
                        "unit": "Tops/s", "frac": achieved / INT32_VALU_PEAK_TOPS,
                        "traffic": MR_TRAFFIC_BYTES_PER_STEP if ctx.log_len == 14 else None,
                        "traffic_note": "HBM bytes per batched squaring step (rocprofv3 "
                                        "FETCH_SIZE + WRITE_SIZE); algorithmic 50.3e6 "
                                        "(NTT(N') and NTT(N) at M/2 points per bigint)",
                        "kernel": f"exp_kernel<{ctx.log_len}>",
                        "ops_per_squaring_step": ops}}
    ctx.close()
    return res


def td_reference_shape(pf, primes: np.ndarray, triple, calls: int = 10) -> dict:
    """The call the reference's worker makes (worker.c:37-42): trial_div_gpu
    at its own grid, K = J = 100, on a host prime list (here every prime
    below 10^8), timed per call end to end and by its gpu_timing_t.  The
    first call uploads the list; later calls compare the host list with the
    copy kept from the upload and reuse the device list."""
    a, b, c = triple
    t0 = time.perf_counter()
    rc, _, first = pf.trial_div_gpu(primes, a, b, c, timing=True)
    first_s = time.perf_counter() - t0
    assert rc == 0
    walls, tims = [], []
    for _ in range(calls):
        t0 = time.perf_counter()
        rc, _, tm = pf.trial_div_gpu(primes, a, b, c, timing=True)
        walls.append(time.perf_counter() - t0)
        tims.append(tm)
        assert rc == 0
    tm = np.mean(np.array(tims), axis=0)
    wall = float(np.mean(walls))
    return {"workload": f"trial_div_gpu, K = J = 100, triple ({a},{b},{c}), "
                        f"{len(primes)} primes below 10^8 (host list)",
            "first_call_ms": first_s * 1e3,
            "first_call_timing_ms": {"upload": first[0] * 1e3, "kernel": first[1] * 1e3,
                                     "download": first[2] * 1e3},
            "ms_per_call": wall * 1e3,
            "timing_ms": {"upload": tm[0] * 1e3, "kernel": tm[1] * 1e3, "download": tm[2] * 1e3},
            "host_ms_per_call": (wall - float(tm.sum())) * 1e3,
            "pair_primes_per_s": 100 * 100 * len(primes) / wall,
            "calls": calls}


def mr_full_round(pf) -> dict:
    """Workload 3 as specified: the 256 bigints (128 N + 128 rev N, 50 000
    digits) built on the device, one base-2 strong-probable-prime round over
    all of them in one launch sequence, wall-clocked end to end (no
    extrapolation).  Every bigint of the workload is composite at base 2, so
    this is also the whole retiring test of bases 2, 3, 5, 7."""
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["3"]
    cands = [tuple(x) + (w,) for x in cfg["candidates"] for w in (0, 1)]
    t0 = time.perf_counter()
    ctx = pf.MrContext(candidates=cands)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    alive, _ = ctx.run([2], residues=False)
    wall = time.perf_counter() - t0
    ctx.close()
    return {"workload": "256 bigints (128 N + 128 rev N) at 50 000 digits, base 2, one launch",
            "wall_s": wall, "build_setup_s": build_s,
            "tests_per_hour": 3600.0 * len(cands) / wall,
            "probable_primes": int(sum(alive))}


def mr_step_leg(pf, cands, n_bits: int, total_bits: int, label: str) -> dict:
    """Seconds per batched squaring step (HIP events on the context's
    stream) of a base-2 round over `cands`, the kernel's roofline by the
    transform-count model, and the same for one candidate alone (the plain
    one-candidate-per-launch path)."""
    ctx = pf.MrContext(candidates=cands)
    ctx.time_squarings(2, 64)
    step_s = ctx.time_squarings(2, n_bits) / 1e3 / n_bits
    logm = ctx.log_len
    ops = mr_ops_per_squaring(logm) * len(cands)
    ctx.close()
    one = pf.MrContext(candidates=cands[:1])
    one.time_squarings(2, 64)
    one_s = one.time_squarings(2, min(n_bits, 1024)) / 1e3 / min(n_bits, 1024)
    one.close()
    achieved = ops / step_s / 1e12
    tph = 3600.0 * len(cands) / (step_s * total_bits)
    plain = 3600.0 / (one_s * total_bits)
    return {"workload": label, "transform_len": 1 << logm,
            "us_per_squaring_batch": step_s * 1e6,
            "tests_per_hour_extrapolated": tph,
            "sample": f"{n_bits} of {total_bits} squaring steps timed (HIP events), scaled",
            "plain_one_candidate": {"us_per_squaring": one_s * 1e6, "tests_per_hour": plain,
                                    "batched_over_plain": tph / plain},
            "roofline": {"bound": "valu", "achieved": achieved, "peak": INT32_VALU_PEAK_TOPS,
                         "unit": "Tops/s", "frac": achieved / INT32_VALU_PEAK_TOPS,
                         "kernel": f"exp_kernel<{logm}>", "ops_per_squaring_step": ops}}


def mr_cpu_baseline(cands, budget_s: float = 15.0, sample_bits: int = 4096) -> dict:
    """GMP mpz_powm (the oracle's harness, oracle/mr_oracle.c) per (bigint,
    base 2) on one core: candidates of the workload in turn, each over the
    top `sample_bits` bits of its d, until about budget_s seconds, scaled
    to the whole exponent."""
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import gmp_build, gmp_powm
    spent, scaled, done, bits = 0.0, 0.0, 0, 0
    for cd in cands:
        n = gmp_build(*cd)
        d = n - 1
        while d % 2 == 0:
            d //= 2
        bits = d.bit_length()
        e = d >> max(0, bits - sample_bits)
        t0 = time.perf_counter()
        gmp_powm(2, e, n)
        dt = time.perf_counter() - t0
        spent += dt
        scaled += dt * bits / sample_bits
        done += 1
        if spent >= budget_s:
            break
    return {"value": 3600.0 * done / scaled, "unit": "tests/hour", "cores": 1, "kind": "port",
            "sample": (f"{done} candidates of the workload, GMP mpz_powm of base 2 over the top "
                       f"{sample_bits} of {bits} bits of d, scaled to the whole exponent "
                       f"(extrapolated); {spent:.1f} s on 1 core")}


def workload3(pf, args) -> dict:
    """Workload 3 on its own line: tests/hour of the batched path at 50 000
    digits, extrapolated from the timed squaring steps (every candidate of
    the workload is composite at base 2, so one round per bigint)."""
    leg = mr_leg(pf, args.mr_bits or 2048, args.digits)
    # the plain path for comparison: one candidate per launch, same kernel
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["3"]
    one = pf.MrContext(candidates=[(args.digits,) + tuple(cfg["candidates"][0][1:]) + (0,)])
    one.time_squarings(2, 64)
    one_step_s = one.time_squarings(2, 1024) / 1e3 / 1024
    one.close()
    plain_tph = 3600.0 / (one_step_s * (int(args.digits * 3.3219280948873623) - 1))
    return {
        "metric": f"Miller-Rabin tests/hour (batched, {args.digits} digits)",
        "value": leg["tests_per_hour_extrapolated"], "unit": "tests/hour", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": leg["us_per_squaring_batch"] / 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded candidates, tests/golden/configs.json workload 3)",
        "config": {"workload": leg["workload"], "transform_len": leg["transform_len"],
                   "sample": leg["sample"], "parallelism": "single GPU"},
        "roofline": leg["roofline"],
        "plain_one_candidate": {"us_per_squaring": one_step_s * 1e6,
                                "tests_per_hour": plain_tph,
                                "batched_over_plain": leg["tests_per_hour_extrapolated"] / plain_tph},
        "cpu_baseline": (mr_cpu_baseline([tuple(x) + (w,) for x in cfg["candidates"] for w in (0, 1)])
                         if not args.no_cpu_baseline else None),
        "device": pf.device_info(),
    }


def _mr_cands(rows, coin=False):
    """(a, b, c, k, j, which) tuples: each row as N and rev(N), or (coin)
    one of the two, chosen by the row's sixth field."""
    out = []
    for r in rows:
        if coin:
            out.append(tuple(r[:5]) + (int(r[5]),))
        else:
            out += [tuple(r[:5]) + (0,), tuple(r[:5]) + (1,)]
    return out


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload4(pf, args) -> dict:
    """512 candidates, a log-uniform in [50 000, 100 000] digits, each as N
    and rev(N) (1 024 bigints across the 2^14 and 2^15 transform buckets),
    bases 2, 3, 5, 7 with retirement at the first failing base, through the
    continuous-batching scheduler.  Wall time from submission to the last
    verdict."""
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["4"]
    cands = _mr_cands(cfg["candidates"])
    if args.limit:
        cands = cands[: args.limit]
    s = pf.MrScheduler(cfg["bases"], chunk=args.chunk)
    t0 = time.perf_counter()
    s.submit(cands)
    res, last = [], t0
    while True:
        left = s.pump(True)
        res += s.results()
        now = time.perf_counter()
        if now - last > 30:
            _progress(f"workload 4: {len(res)}/{len(cands)} decided after {now - t0:.0f} s")
            last = now
        if left == 0:
            break
    wall = time.perf_counter() - t0
    st = s.stats()
    s.close()
    rounds = sum(r.rounds for r in res)
    tests_h = 3600.0 * rounds / wall
    return {
        "metric": "Miller-Rabin tests/hour (batched, mixed sizes)",
        "value": tests_h, "unit": "tests/hour", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded candidates, tests/golden/configs.json workload 4)",
        "config": {"workload": f"{len(cands)} bigints (N and rev N), a log-uniform in "
                               f"[50000, 100000] digits, bases 2,3,5,7, retire on first failure",
                   "candidates": len(cands), "parallelism": "single GPU"},
        "candidates_per_hour": 3600.0 * len(cands) / wall,
        "rounds": rounds, "probable_primes": sum(r.verdict for r in res),
        "scheduler": {"chunks": st.chunks, "admin_batches": st.admin_batches,
                      "slot_steps": st.squarings, "hbm_bytes": st.hbm_bytes,
                      "max_active": st.max_active, "pools": st.n_pools, "chunk": args.chunk},
        "device": pf.device_info(),
    }


def capacity_estimate(pf, rows, bases_n=1) -> dict:
    """Requests/s the scheduler sustains on this request mix: measured
    seconds per squaring step of a full slot pool (one candidate per compute
    unit) at each transform length, times the mean exponent length of each
    bucket, assuming one round per request (composites retire at base 2)."""
    cus = pf.device_info()["cus"]
    per_step = {}
    for digits in (50000, 100000):
        probe = [(digits, 20000, 300, 4001, 17, w) for w in (0, 1)] * (cus // 2)
        ctx = pf.MrContext(candidates=probe)
        ctx.time_squarings(2, 16)
        per_step[ctx.log_len] = ctx.time_squarings(2, 256) / 1e3 / 256
        ctx.close()
    busy = 0.0
    for r in rows:
        bits = int(r[0] * 3.3219280948873623)
        L = (bits + 2) // 24 + 1
        logm = 14 if 2 * L + 2 <= 1 << 14 else 15
        busy += bits * per_step[logm]
    mean_s = busy / len(rows)               # one slot busy per request
    return {"requests_per_s": cus / mean_s, "mean_test_s": mean_s,
            "s_per_step": {str(1 << k): v for k, v in per_step.items()}}


def workload5(pf, args, cap=None) -> dict:
    """Serving: requests of tests/golden/configs.json workload 5 (a uniform in
    [50 000, 100 000] digits, N or rev(N) by the row's coin, bases 2, 3, 5, 7)
    arrive as a seeded Poisson process at `load` times the measured capacity
    for `serve_seconds`; the scheduler admits them first come, first served
    into free slots.  Reports time-to-verdict percentiles over the requests
    that arrived in the window (all are drained)."""
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["5"]
    rows = cfg["requests"]
    cap = cap or capacity_estimate(pf, rows)
    rate = args.load * cap["requests_per_s"]
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(rows)))
    n_req = int(np.searchsorted(arrivals, args.serve_seconds))
    cands = _mr_cands(rows[:n_req], coin=True)
    _progress(f"workload 5: capacity {cap['requests_per_s']:.3f} req/s, load {args.load}, "
              f"{n_req} requests over {args.serve_seconds} s")
    if n_req == 0:
        raise RuntimeError("workload 5: no request arrives in the window")
    s = pf.MrScheduler(cfg["bases"], chunk=args.chunk)
    t0 = time.perf_counter()
    mono0 = time.clock_gettime(time.CLOCK_MONOTONIC)
    nxt, res, last = 0, [], t0
    while True:
        now = time.perf_counter() - t0
        k = nxt
        while k < n_req and arrivals[k] <= now:
            k += 1
        if k > nxt:
            s.submit(cands[nxt:k], tags=list(range(nxt, k)))
            nxt = k
        left = s.pump(True)
        res += s.results()
        if time.perf_counter() - last > 30:
            last = time.perf_counter()
            _progress(f"workload 5: t={last - t0:.0f} s, {nxt} arrived, {len(res)} decided")
        if nxt == n_req and left == 0:
            break
        if nxt < n_req and left == 0:        # idle until the next arrival
            time.sleep(max(0.0, min(0.05, arrivals[nxt] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    st = s.stats()
    s.close()
    ttv = np.array([r.t_done - (mono0 + arrivals[r.tag]) for r in res])
    rounds = sum(r.rounds for r in res)
    return {
        "metric": "Miller-Rabin serving tests/hour",
        "value": 3600.0 * rounds / wall, "unit": "tests/hour", "n_gpus": 1, "steps": 1,
        "warmup": 0, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded requests, tests/golden/configs.json workload 5)",
        "config": {"workload": f"serving: Poisson arrivals at {args.load} x capacity for "
                               f"{args.serve_seconds} s, a in [50000, 100000] digits, bases 2,3,5,7",
                   "requests": n_req, "load": args.load, "parallelism": "single GPU"},
        "p50_s": float(np.percentile(ttv, 50)), "p99_s": float(np.percentile(ttv, 99)),
        "mean_s": float(ttv.mean()), "max_s": float(ttv.max()),
        "capacity": cap, "arrival_rate_per_s": rate,
        "scheduler": {"chunks": st.chunks, "admin_batches": st.admin_batches,
                      "max_active": st.max_active, "pools": st.n_pools,
                      "hbm_bytes": st.hbm_bytes, "hbm_budget": 64 << 30},
        "device": pf.device_info(),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="seconds of work per process in each CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mr-bits", type=int, default=2048,
                    help="squaring steps timed per Miller-Rabin sample (0 = skip the MR leg)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain copy kernel of this many MiB (counter calibration)")
    ap.add_argument("--workload", type=int, default=2, choices=(2, 3, 4, 5),
                    help="2: trial-division headline; 3: batched Miller-Rabin at 50 000 digits; "
                         "4: batched Miller-Rabin of mixed sizes; 5: Miller-Rabin serving")
    ap.add_argument("--limit", type=int, default=0, help="workload 4: first N bigints only")
    ap.add_argument("--digits", type=int, default=50000,
                    help="workload 3: decimal digits a (the workload's b, c, k, j kept)")
    ap.add_argument("--chunk", type=int, default=256, help="squaring steps per scheduler chunk")
    ap.add_argument("--load", type=float, default=0.70, help="workload 5: arrival rate / capacity")
    ap.add_argument("--serve-seconds", type=float, default=300.0,
                    help="workload 5: length of the arrival window")
    ap.add_argument("--serve-window", type=float, default=30.0,
                    help="default line: arrival window of each serving sample at 0.70 and "
                         "0.95 x capacity (0 = no serving sample)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not "
              f"supported (the triples shard across GPUs with no exchange)", file=sys.stderr)
        return 2

    pf = _load("pf_amd", ROOT / "prime-finder_amd" / "pf_amd.py")
    cpu = None
    if args.workload == 2 and not args.no_cpu_baseline:
        # before anything touches the GPU: the pool forks this process
        _progress("CPU baselines (trial division, GMP Miller-Rabin, GMP BPSW)")
        cpu = cpu_baselines(args.cpu_seconds)
    if args.workload in (3, 4, 5):
        out = {3: workload3, 4: workload4, 5: workload5}[args.workload](pf, args)
        print(json.dumps(out), flush=True)
        return 0
    L = pf.lib()
    assert L.cuda_get_device_count() >= 1, "no HIP device"
    cfg = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["2"]
    triples = cfg["triples"]
    primes = pf.primes_below(cfg["prime_limit"])
    n = len(primes)
    _progress("algorithmic work of the per-prime kernels, counted on the host")
    works = [pf.trial_div_work(primes, a, b, c, K, J, cpu_share()) for (a, b, c) in triples]
    work_avg = {f: float(np.mean([w[f] for w in works])) for f in pf.TD_WORK_FIELDS}
    # the events bracket the lattice kernel; the k loop below it runs beside
    # it on the second stream, with the row kernel
    alg_ops = sum(TD_OPS_PER_UNIT[f] * work_avg[f] for f in TD_OPS_PER_UNIT if f != "kstep")
    nw = (K * J + 31) // 32

    d_primes = pf.DeviceBuffer.from_array(primes)
    d_bits = [pf.DeviceBuffer(nw * 4) for _ in triples]
    evs = [[(L.pf_event_create(), L.pf_event_create()) for _ in triples]
           for _ in range(args.steps)]

    def step(events=None):
        for t, (a, b, c) in enumerate(triples):
            if L.pf_memset(d_bits[t].ptr, 0, nw * 4) != 0:
                raise RuntimeError("pf_memset failed")
            e0, e1 = events[t] if events else (None, None)
            rc = L.trial_div_launch_kj_ev(d_primes.ptr, n, a, b, c, K, J, d_bits[t].ptr,
                                          None, e0, e1)
            if rc != 0:
                raise RuntimeError("trial_div_launch_kj_ev failed")

    if args.copy_mib:
        nb = args.copy_mib << 20
        src, dst = pf.DeviceBuffer(nb), pf.DeviceBuffer(nb)
        src.zero()
        assert L.pf_device_copy(dst.ptr, src.ptr, nb, None) == 0
        assert L.pf_sync() == 0
        src.close()
        dst.close()

    for _ in range(args.warmup):
        step()
    assert L.pf_sync() == 0
    assert L.trial_div_check_error() == 0
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(evs[s])
    assert L.pf_sync() == 0
    wall = time.perf_counter() - t0
    assert L.trial_div_check_error() == 0

    lane_ms = [L.pf_event_elapsed_ms(e0, e1) for row in evs for (e0, e1) in row]
    for row in evs:
        for e0, e1 in row:
            L.pf_event_destroy(e0)
            L.pf_event_destroy(e1)
    avg_lane_s = float(np.mean(lane_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3
    work = len(triples) * K * J * n
    value = work / (wall / args.steps)
    achieved = alg_ops / avg_lane_s / 1e12

    bits0 = d_bits[0].download()
    survivors0 = K * J - int(np.unpackbits(bits0.view(np.uint8)).sum())

    out = {
        "metric": "trial-div pair*primes/s",
        "value": value,
        "unit": "pair*primes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded triples, tests/golden/configs.json workload 2)",
        "config": {"workload": "trial division: 8 triples, K=J=4096, primes < 1e8",
                   "triples": len(triples), "K": K, "J": J, "primes": n,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": INT32_VALU_PEAK_TOPS,
                     "unit": "Tops/s", "frac": achieved / INT32_VALU_PEAK_TOPS,
                     "traffic": TD_TRAFFIC_BYTES,
                     "traffic_note": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE); "
                                     "algorithmic 25.1e6 (prime list + bitmap), the rest are "
                                     f"hit atomics; {TD_TRAFFIC_BYTES / avg_lane_s / 1e12:.2f} "
                                     "TB/s, far below the 8 TB/s roof",
                     "kernel": "td_lattice_kernel<4>",
                     "kernel_avg_ms": avg_lane_s * 1e3,
                     "ops_per_launch": alg_ops,
                     "work_per_launch": work_avg,
                     "ops_per_unit": TD_OPS_PER_UNIT,
                     "lattice_min": int(L.trial_div_lattice_min())},
        "survivors_first_triple": survivors0,
        "device": pf.device_info(),
    }
    _progress("trial division at the reference's grid (K = J = 100)")
    out["td_reference_shape"] = td_reference_shape(pf, primes, triples[0])
    if cpu:
        out["cpu_baseline"] = cpu["td"]
    if args.mr_bits:
        cfg3 = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["3"]
        w3 = [tuple(x) + (w,) for x in cfg3["candidates"] for w in (0, 1)]
        w3_100k = [(100000,) + tuple(x[1:]) + (w,) for x in cfg3["candidates"] for w in (0, 1)]
        _progress("Miller-Rabin: full base-2 round of workload 3 (256 bigints, 50 000 digits)")
        mr = {"uniform_50k_full_round": mr_full_round(pf)}
        _progress("Miller-Rabin: squaring-step rates at 50 000 and 100 000 digits")
        mr["step_50k"] = mr_step_leg(pf, w3, args.mr_bits, 166096,
                                     "256 bigints (128 N + 128 rev N) at 50 000 digits, base 2")
        mr["step_100k"] = mr_step_leg(pf, w3_100k, max(256, args.mr_bits // 2), 332192,
                                      "256 bigints at 100 000 digits (workload 3's b, c, k, j "
                                      "with a = 100 000; N and rev N), base 2")
        if args.serve_window > 0:
            cfg5 = json.loads((ROOT / "tests" / "golden" / "configs.json").read_text())["5"]
            cap = capacity_estimate(pf, cfg5["requests"])
            mr["serving"] = {}
            for load in (0.70, 0.95):
                _progress(f"Miller-Rabin serving window at {load} x capacity")
                sargs = argparse.Namespace(**vars(args))
                sargs.load, sargs.serve_seconds = load, args.serve_window
                w5 = workload5(pf, sargs, cap)
                mr["serving"][f"{load:.2f}"] = {
                    k: w5[k] for k in ("value", "unit", "p50_s", "p99_s", "mean_s", "max_s",
                                       "arrival_rate_per_s", "ms_per_step")}
                mr["serving"][f"{load:.2f}"]["requests"] = w5["config"]["requests"]
                mr["serving"][f"{load:.2f}"]["window_s"] = args.serve_window
            mr["serving"]["capacity"] = cap
        if cpu:
            mr["cpu_baseline"] = cpu["mr"]
            mr["cpu_reference_verdict"] = cpu["mr_reference_verdict"]
        out["mr"] = mr
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
