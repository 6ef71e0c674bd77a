"""CPU checks of the C ABI: the library loads without a GPU and exports
every function declared in include/*.h; the host-side prime tables match an
independent sieve and the reference's file format."""
import ctypes as C
import re
import struct
import subprocess

import numpy as np
import pytest

from _support import PKG, ROOT, numpy_primes, write_gap_file

INCLUDE = ROOT / "include"


def declared_functions():
    names = {}
    for h in sorted(INCLUDE.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names[name] = h.name
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", str(path)], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_exports_every_declared_function(pf):
    lib = pf.lib()   # loads without a GPU
    assert lib is not None
    syms = exported_symbols(pf.LIB_PATH)
    decl = declared_functions()
    assert "trial_div_gpu" in decl and "cuda_set_device" in decl
    assert "miller_rabin_test" in decl and "mr_run" in decl
    missing = []
    for name, hdr in decl.items():
        if name == "trial_div_launch":
            # C++ linkage, as in the reference (trial_div.cu:127)
            if not any(s.startswith("_Z16trial_div_launch") for s in syms):
                missing.append((name, hdr))
        elif name not in syms:
            missing.append((name, hdr))
    assert not missing, missing


def test_reference_abi_struct_layouts(pf):
    assert C.sizeof(pf.PrimeList) == 16          # {uint32_t *, uint32_t} on LP64
    assert C.sizeof(pf.GpuTiming) == 24
    assert C.sizeof(pf.MrTiming) == 32          # four doubles, miller_rabin.h:3-8


def test_prime_sieve_matches_numpy(pf):
    for limit in (2, 3, 10, 1000, 65537, 10 ** 6):
        assert (pf.primes_below(limit) == numpy_primes(limit)).all(), limit
    p = pf.primes_below(10 ** 6)
    assert len(p) == 78498
    q = pf.primes_below(10 ** 6, lo=3)
    assert len(q) == 78497 and q[0] == 3


@pytest.mark.slow
def test_prime_sieve_1e8_count(pf):
    p = pf.primes_below(10 ** 8)
    assert len(p) == 5761455 and p[-1] == 99999989


def test_prime_list_load_reference_format(pf, tmp_path):
    L = pf.lib()
    primes = numpy_primes(200000)[1:]                  # 2 excluded, as gen_primes writes
    # append gaps > 255 (up to the format's u16 limit) to exercise the escape;
    # the loader does not test primality
    last = int(primes[-1])
    primes = np.concatenate([primes, np.array([last + 282, last + 282 + 65535, last + 65817 + 256],
                                              np.uint32)])
    f = tmp_path / "p.bin"
    write_gap_file(f, primes)
    pl = pf.PrimeList()
    assert L.prime_list_load(C.byref(pl), str(f).encode()) == 0
    got = np.ctypeslib.as_array(pl.primes, shape=(pl.count,)).copy()
    L.prime_list_free(C.byref(pl))
    assert (got == primes).all()
    # truncated file -> error
    f.write_bytes(f.read_bytes()[:-2])
    assert L.prime_list_load(C.byref(pl), str(f).encode()) == -1


def test_scheduler_rejects_bad_arguments_without_gpu(pf):
    L = pf.mr_lib()
    import ctypes as C
    bad = (C.c_uint32 * 2)(2, 1)                 # base 1 is not allowed
    assert not L.mr_sched_create(bad, 2, 0, 0, 0, 0)
    assert not L.mr_sched_create(None, 0, 0, 0, 0, 0)
    assert L.mr_sched_pump(None, 0) == -1
    assert L.mr_sched_submit(None, None, None, 0) == -1


def test_cpp_runner_header_compiles():
    """include/batch_mod_ctx.hpp (the reference's C++ runner surface) is
    self-contained C++17."""
    src = "#include \"batch_mod_ctx.hpp\"\nint main() { return DEFAULT_WITNESSES.size() == 16 ? 0 : 1; }\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{INCLUDE}",
                        "-x", "c++", "-"], input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_trial_div_work_counts(pf):
    """Host-side work counts used to rate the kernels: the lattice takes the
    primes from trial_div_lattice_min() up, the k loop those from 2^16 below
    it (or all from 2^16 when a - c < D_j), the counts do not depend on the
    thread count, and invalid arguments are rejected."""
    primes = numpy_primes(3_000_000)
    lmin = int(pf.lib().trial_div_lattice_min())
    assert 1 << 16 <= lmin <= 1 << 24
    a, b, c, K, J = 60000, 21000, 100, 512, 700
    w1 = pf.trial_div_work(primes, a, b, c, K, J, 1)
    w4 = pf.trial_div_work(primes, a, b, c, K, J, 4)
    assert w1 == w4
    assert w1["prime"] == int((primes >= lmin).sum())
    assert w1["kstep"] == K * int(((primes >= 1 << 16) & (primes < lmin)).sum())
    # every lattice prime searches the N condition at least once, and each
    # query that is not decided at once takes at least one level
    assert w1["query"] >= w1["prime"] and w1["level"] >= w1["query"] - w1["hit"]
    assert w1["modmul"] + w1["modmul_lazy"] > 50 * w1["prime"]
    # the setup's powers take lazily reduced products below 2^30
    assert w1["modmul_lazy"] > 30 * w1["prime"]
    # j digit groups that do not merge: every prime from 2^16 on the k loop
    w = pf.trial_div_work(primes, 60000, 21000, 59998, K, J, 2)
    assert w["prime"] == 0 and w["kstep"] == K * int((primes >= 1 << 16).sum())
    with pytest.raises(ValueError):
        pf.trial_div_work(primes, 100, 200, 10, K, J)
    with pytest.raises(ValueError):
        pf.trial_div_work(primes, 60000, 21000, 100, 0, J)
