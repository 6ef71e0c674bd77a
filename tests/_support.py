"""Test-side bindings: the product library (through the package's ctypes
module) and the CPU oracle (oracle/_build, oracle/_ref).  The oracle is only
ever used here as the checker."""
from __future__ import annotations

import ctypes as C
import importlib.util
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "prime-finder_amd"
ORACLE_LIB = ROOT / "oracle" / "_build" / "liboracle.so"
REF_TD_LIB = ROOT / "oracle" / "_ref" / "libref_td.so"
REF_MR_LIB = ROOT / "oracle" / "_ref" / "libref_mr.so"
GOLDEN = ROOT / "tests" / "golden"

u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
ll = C.c_longlong


def load_pf():
    """The product package's binding module (prime-finder_amd/pf_amd.py)."""
    if "pf_amd" in sys.modules:
        return sys.modules["pf_amd"]
    spec = importlib.util.spec_from_file_location("pf_amd", PKG / "pf_amd.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pf_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_oracle = None


def oracle() -> C.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing; run __graft_entry__.build()")
        L = C.CDLL(str(ORACLE_LIB))
        for name in ("td_oracle_sieve", "td_oracle_brute"):
            getattr(L, name).argtypes = [u32p, C.c_uint32, ll, ll, ll, C.c_uint32, C.c_uint32, u32p]
            getattr(L, name).restype = C.c_int
        L.td_oracle_sieve_mt.argtypes = [u32p, C.c_uint32, ll, ll, ll, C.c_uint32, C.c_uint32,
                                         u32p, C.c_int]
        L.td_oracle_sieve_mt.restype = C.c_int
        L.td_oracle_check_pairs.argtypes = [u32p, C.c_uint32, ll, ll, ll, u32p, u32p, C.c_uint32,
                                            u8p, C.c_int]
        L.td_oracle_check_pairs.restype = C.c_int
        _oracle = L
    return _oracle


TEST_GPU_PRIMES = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71,
                   73, 79, 83, 89, 97]


def golden_cases():
    """(case, primes, bitmap words) of tests/golden/td_ref_k100.json: bitmaps
    of the reference's own trial_div_cpu.c at K = J = 100."""
    import json
    data = json.loads((GOLDEN / "td_ref_k100.json").read_text())
    p20k = numpy_primes(20000)
    for case in data["cases"]:
        primes = np.array(TEST_GPU_PRIMES, np.uint32) if case["primes"] == "test_gpu" else p20k
        words = np.array([int(case["bits"][i:i + 8], 16) for i in range(0, len(case["bits"]), 8)],
                         np.uint32)
        yield case, primes, words


def nwords(K: int, J: int) -> int:
    return (K * J + 31) // 32


def oracle_sieve(primes, a, b, c, K=100, J=100, threads: int = 0) -> np.ndarray:
    p = np.ascontiguousarray(primes, dtype=np.uint32)
    bits = np.zeros(nwords(K, J), np.uint32)
    L = oracle()
    if threads == 1:
        rc = L.td_oracle_sieve(p.ctypes.data_as(u32p), len(p), a, b, c, K, J,
                               bits.ctypes.data_as(u32p))
    else:
        rc = L.td_oracle_sieve_mt(p.ctypes.data_as(u32p), len(p), a, b, c, K, J,
                                  bits.ctypes.data_as(u32p), threads)
    assert rc == 0
    return bits


def oracle_brute(primes, a, b, c, K=100, J=100) -> np.ndarray:
    p = np.ascontiguousarray(primes, dtype=np.uint32)
    bits = np.zeros(nwords(K, J), np.uint32)
    assert oracle().td_oracle_brute(p.ctypes.data_as(u32p), len(p), a, b, c, K, J,
                                    bits.ctypes.data_as(u32p)) == 0
    return bits


def oracle_check_pairs(primes, a, b, c, ks, js, threads: int = 0) -> np.ndarray:
    p = np.ascontiguousarray(primes, dtype=np.uint32)
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    js = np.ascontiguousarray(js, dtype=np.uint32)
    flags = np.zeros(len(ks), np.uint8)
    assert oracle().td_oracle_check_pairs(p.ctypes.data_as(u32p), len(p), a, b, c,
                                          ks.ctypes.data_as(u32p), js.ctypes.data_as(u32p),
                                          len(ks), flags.ctypes.data_as(u8p), threads) == 0
    return flags


class RefPrimeList(C.Structure):
    _fields_ = [("primes", u32p), ("count", C.c_uint32)]


_ref_td = None


def have_ref_td() -> bool:
    """True when oracle/_ref holds the reference's own trial_div_cpu (built in
    the container that has the reference sources; never shipped to a GPU
    run).  Checks the file only: nothing is loaded."""
    return REF_TD_LIB.exists()


def ref_td():
    """The reference's own trial_div_cpu (K = J = 100) built from its sources,
    loaded on first use, or None when oracle/_ref was not built."""
    global _ref_td
    if _ref_td is None and REF_TD_LIB.exists():
        L = C.CDLL(str(REF_TD_LIB))
        L.trial_div_cpu.argtypes = [C.POINTER(RefPrimeList), ll, ll, ll, u32p]
        L.trial_div_cpu.restype = None
        _ref_td = L
    return _ref_td


def ref_trial_div_cpu(primes, a, b, c) -> np.ndarray:
    p = np.ascontiguousarray(primes, dtype=np.uint32)
    bits = np.zeros(313, np.uint32)
    pl = RefPrimeList(p.ctypes.data_as(u32p), len(p))
    ref_td().trial_div_cpu(C.byref(pl), a, b, c, bits.ctypes.data_as(u32p))
    return bits


def numpy_primes(limit: int) -> np.ndarray:
    """Independent sieve (numpy) used to cross-check the library's lists."""
    s = np.ones(limit, dtype=bool)
    s[:2] = False
    for i in range(2, int(limit ** 0.5) + 1):
        if s[i]:
            s[i * i::i] = False
    return np.nonzero(s)[0].astype(np.uint32)


def bit_get(bits: np.ndarray, J: int, k: int, j: int) -> int:
    b = (k - 1) * J + (j - 1)
    return int((int(bits[b >> 5]) >> (b & 31)) & 1)


def survivors(bits: np.ndarray, K: int, J: int):
    """Sorted (k, j) survivors, lexicographic as worker.c:45-54."""
    flat = np.unpackbits(bits.view(np.uint8), bitorder="little")[: K * J]
    idx = np.nonzero(flat == 0)[0]
    return [(int(i // J) + 1, int(i % J) + 1) for i in idx]


def popcount(bits: np.ndarray) -> int:
    return int(np.unpackbits(bits.view(np.uint8)).sum())


# ------------------------------------------------------------- GMP oracle
_mr_oracle_ready = False


def _mr_oracle():
    global _mr_oracle_ready
    L = oracle()
    if not _mr_oracle_ready:
        sz = C.c_size_t
        L.mr_oracle_build.argtypes = [ll, ll, ll, C.c_ulong, C.c_ulong, C.c_int, u32p, sz]
        L.mr_oracle_build.restype = C.c_long
        L.mr_oracle_sprp.argtypes = [u32p, sz, C.c_ulong, u32p, C.POINTER(C.c_int)]
        L.mr_oracle_sprp.restype = C.c_int
        L.mr_oracle_probab_prime.argtypes = [u32p, sz, C.c_int]
        L.mr_oracle_probab_prime.restype = C.c_int
        L.mr_oracle_powm.argtypes = [u32p, sz, u32p, u32p, sz, u32p]
        L.mr_oracle_powm.restype = C.c_int
        L.mr_oracle_gmp_version.restype = C.c_char_p
        _mr_oracle_ready = True
    return L


def int_words(x: int, n: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype=np.uint32).copy()


def words_int(w) -> int:
    return int.from_bytes(np.ascontiguousarray(w, dtype=np.uint32).tobytes(), "little")


def gmp_build(a, b, c, k, j, mode=0) -> int:
    """N (mode 0), decimal reversal of N (1) or the reference rev formula (2)."""
    cap = (int(a * 3.33) + 64) // 32 + 4
    out = np.zeros(cap, np.uint32)
    n = _mr_oracle().mr_oracle_build(a, b, c, k, j, mode, out.ctypes.data_as(u32p), cap)
    if n < 0:
        raise ValueError("mr_oracle_build failed")
    return words_int(out)


def gmp_sprp(N: int, base: int):
    """(verdict, base^d mod N, s) from GMP mpz_powm."""
    nw = (N.bit_length() + 31) // 32
    w = int_words(N, nw)
    res = np.zeros(nw, np.uint32)
    s = C.c_int()
    v = _mr_oracle().mr_oracle_sprp(w.ctypes.data_as(u32p), nw, base, res.ctypes.data_as(u32p),
                                    C.byref(s))
    return v, words_int(res), s.value


def gmp_probab_prime(N: int, reps: int = 5) -> int:
    nw = (N.bit_length() + 31) // 32
    w = int_words(N, nw)
    return _mr_oracle().mr_oracle_probab_prime(w.ctypes.data_as(u32p), nw, reps)


def gmp_version() -> str:
    return _mr_oracle().mr_oracle_gmp_version().decode()


def gmp_powm(base: int, e: int, N: int) -> int:
    """base^e mod N with GMP mpz_powm (the oracle's harness)."""
    nw = (N.bit_length() + 31) // 32
    ne = max(1, (e.bit_length() + 31) // 32)
    out = np.zeros(nw, np.uint32)
    Nw, Bw, Ew = int_words(N, nw), int_words(base % N, nw), int_words(e, ne)
    rc = _mr_oracle().mr_oracle_powm(Nw.ctypes.data_as(u32p), nw, Bw.ctypes.data_as(u32p),
                                     Ew.ctypes.data_as(u32p), ne, out.ctypes.data_as(u32p))
    if rc != 0:
        raise ValueError("mr_oracle_powm failed")
    return words_int(out)


def inv_pow2_mod(n: int, R: int) -> int:
    """R^-1 mod n for odd n and R a power of two: u = n^-1 mod R by Newton
    (y <- y (2 - n y), doubling the bits), n u = 1 + k R, so R (-k) == 1
    (mod n).  Far faster than pow(R, -1, n) at 10^5 bits."""
    kb, y = 1, 1
    while (1 << kb) < R:
        kb *= 2
        y = y * (2 - n * y) % (1 << kb)
    k = (n * (y % R) - 1) // R
    return (-k) % n


def sprp_python(N: int, base: int):
    """Textbook strong-probable-prime round with Python integers."""
    d, s = N - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    r = pow(base, d, N)
    if base % N == 0:          # a multiple of N says nothing about N
        return True, r, s
    if r == 1 or r == N - 1:
        return True, r, s
    x = r
    for _ in range(s - 1):
        x = x * x % N
        if x == N - 1:
            return True, r, s
        if x == 1:
            break
    return False, r, s


def write_gap_file(path, primes):
    """A prime list in the reference's gap-encoded file format
    (gen_primes.c:98-127): u32 count, u32 first prime, then one byte per
    gap, a gap above 255 written as 0 and a little-endian u16."""
    import struct
    out = bytearray(struct.pack("<I", len(primes)))
    if len(primes):
        out += struct.pack("<I", int(primes[0]))
    for i in range(1, len(primes)):
        gap = int(primes[i]) - int(primes[i - 1])
        out += bytes([gap]) if gap <= 255 else bytes([0, gap & 0xFF, gap >> 8])
    path.write_bytes(bytes(out))
