"""ctypes binding of ``lib/libpfamd.so`` (the gfx950 library).

The Python names mirror the reference's C entry points for the hot path
(``trial_div_gpu``, ``cuda_get_device_count``, ``cuda_set_device``;
reference trial_div.h:26-32) and add the K x J and device-resident forms.
Loading fails loudly when the library has not been built: there is no
Python or CPU fallback for any computation here.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "lib" / "libpfamd.so"

u32p = C.POINTER(C.c_uint32)


class PrimeList(C.Structure):
    """``prime_list_t`` (reference prime_list.h:4-7)."""
    _fields_ = [("primes", u32p), ("count", C.c_uint32)]


class GpuTiming(C.Structure):
    """``gpu_timing_t`` (reference trial_div.h:10-14), seconds."""
    _fields_ = [("upload", C.c_double), ("kernel", C.c_double), ("download", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """Load the library once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    ll, u32, vp, sz = C.c_longlong, C.c_uint32, C.c_void_p, C.c_size_t
    sig = {
        "trial_div_gpu": (C.c_int, [C.POINTER(PrimeList), ll, ll, ll, u32p, C.POINTER(GpuTiming)]),
        "trial_div_gpu_kj": (C.c_int, [C.POINTER(PrimeList), ll, ll, ll, u32, u32, u32p, sz,
                                       C.POINTER(GpuTiming)]),
        "trial_div_launch_kj": (C.c_int, [vp, u32, ll, ll, ll, u32, u32, vp, vp]),
        "trial_div_launch_kj_ev": (C.c_int, [vp, u32, ll, ll, ll, u32, u32, vp, vp, vp, vp]),
        "trial_div_launch_kj_noalloc": (C.c_int, [vp, u32, ll, ll, ll, u32, u32, vp, vp]),
        "trial_div_check_error": (C.c_int, []),
        "trial_div_survivors": (C.c_int, [vp, u32, u32, vp, u32, C.POINTER(C.c_uint32), vp]),
        "trial_div_lattice_min": (u32, []),
        "trial_div_work": (C.c_int, [u32p, u32, ll, ll, ll, u32, u32, u32,
                                     C.POINTER(C.c_uint64)]),
        "cuda_get_device_count": (C.c_int, []),
        "cuda_set_device": (C.c_int, [C.c_int]),
        "prime_list_load": (C.c_int, [C.POINTER(PrimeList), C.c_char_p]),
        "prime_list_free": (None, [C.POINTER(PrimeList)]),
        "prime_list_sieve": (C.c_int, [C.POINTER(PrimeList), u32, u32]),
        "pf_device_count": (C.c_int, []),
        "pf_device_name": (C.c_char_p, []),
        "pf_device_arch": (C.c_int, [C.c_char_p, sz]),
        "pf_device_cu_count": (C.c_int, []),
        "pf_malloc": (vp, [sz]),
        "pf_free": (C.c_int, [vp]),
        "pf_memcpy_htod": (C.c_int, [vp, vp, sz]),
        "pf_memcpy_dtoh": (C.c_int, [vp, vp, sz]),
        "pf_memcpy_dtod": (C.c_int, [vp, vp, sz]),
        "pf_memset": (C.c_int, [vp, C.c_int, sz]),
        "pf_sync": (C.c_int, []),
        "pf_device_copy": (C.c_int, [vp, vp, sz, vp]),
        "pf_stream_create": (vp, []),
        "pf_stream_destroy": (C.c_int, [vp]),
        "pf_stream_sync": (C.c_int, [vp]),
        "pf_event_create": (vp, []),
        "pf_event_destroy": (C.c_int, [vp]),
        "pf_event_record": (C.c_int, [vp, vp]),
        "pf_event_sync": (C.c_int, [vp]),
        "pf_event_elapsed_ms": (C.c_float, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _u32(arr: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    return a


def primes_below(limit: int, lo: int = 2) -> np.ndarray:
    """All primes lo <= p < limit (uint32), from the library's host sieve."""
    L = lib()
    pl = PrimeList()
    if L.prime_list_sieve(C.byref(pl), lo, limit) != 0:
        raise MemoryError("prime_list_sieve failed")
    try:
        out = np.ctypeslib.as_array(pl.primes, shape=(pl.count,)).copy() if pl.count else \
            np.zeros(0, np.uint32)
    finally:
        L.prime_list_free(C.byref(pl))
    return out


def n_words(K: int, J: int) -> int:
    return (K * J + 31) // 32


def trial_div_gpu(primes: np.ndarray, a: int, b: int, c: int, bits: np.ndarray | None = None,
                  timing: bool = False):
    """Reference-ABI call (K = J = 100, 313 words).  Returns (rc, bits[, timing])."""
    L = lib()
    p = _u32(primes)
    pl = PrimeList(p.ctypes.data_as(u32p), len(p))
    out = np.zeros(313, np.uint32) if bits is None else _u32(bits).copy()
    t = GpuTiming()
    rc = L.trial_div_gpu(C.byref(pl), a, b, c, out.ctypes.data_as(u32p), C.byref(t))
    if timing:
        return rc, out, (t.upload, t.kernel, t.download)
    return rc, out


def trial_div_gpu_kj(primes: np.ndarray, a: int, b: int, c: int, K: int, J: int,
                     bits: np.ndarray | None = None, timing: bool = False):
    """K x J grid.  Returns (rc, bits[, timing])."""
    L = lib()
    p = _u32(primes)
    pl = PrimeList(p.ctypes.data_as(u32p), len(p))
    nw = n_words(K, J)
    out = np.zeros(nw, np.uint32) if bits is None else _u32(bits).copy()
    t = GpuTiming()
    rc = L.trial_div_gpu_kj(C.byref(pl), a, b, c, K, J, out.ctypes.data_as(u32p), len(out),
                            C.byref(t))
    if timing:
        return rc, out, (t.upload, t.kernel, t.download)
    return rc, out


TD_WORK_FIELDS = ("prime", "modmul", "query", "level", "hit", "step", "kstep", "gapstep",
                  "gapcf", "modmul_lazy")


def trial_div_work(primes: np.ndarray, a: int, b: int, c: int, K: int, J: int,
                   threads: int = 1) -> dict:
    """Algorithmic work of one launch on (a, b, c) over `primes`, counted on
    the host (trial_div_work): lattice primes, Montgomery multiplications
    (fully, and in "modmul_lazy" lazily, reduced),
    first-hit queries, continued-fraction levels, hits, one-digit-class
    steps, k-loop prime x k steps and return-time steps between hits."""
    p = _u32(primes)
    out = (C.c_uint64 * len(TD_WORK_FIELDS))()
    if lib().trial_div_work(p.ctypes.data_as(u32p), len(p), a, b, c, K, J, threads, out) != 0:
        raise ValueError("trial_div_work: invalid arguments")
    return dict(zip(TD_WORK_FIELDS, (int(x) for x in out)))


def survivors_gpu(d_bits, K: int, J: int, cap: int | None = None) -> np.ndarray:
    """Survivor pairs of a device bitmap (trial_div_survivors): an (n, 2)
    array of (k, j), lexicographic.  d_bits: a DeviceBuffer or pointer."""
    ptr = d_bits.ptr if hasattr(d_bits, "ptr") else d_bits
    cap = K * J if cap is None else cap
    out = DeviceBuffer(4 * max(cap, 1))
    n = C.c_uint32()
    if lib().trial_div_survivors(ptr, K, J, out.ptr, cap, C.byref(n), None) != 0:
        raise RuntimeError("trial_div_survivors failed")
    got = min(n.value, cap)
    packed = out.download(count=got) if got else np.zeros(0, np.uint32)
    out.close()
    return np.stack([packed >> 16, packed & 0xFFFF], axis=1).astype(np.int64)


class DeviceBuffer:
    """A device allocation owned by Python (freed on close / GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().pf_malloc(self.nbytes)
        if not self.ptr:
            raise MemoryError(f"pf_malloc({self.nbytes}) failed")

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        if lib().pf_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError("pf_memcpy_htod failed")

    def download(self, dtype=np.uint32, count: int | None = None) -> np.ndarray:
        n = self.nbytes // np.dtype(dtype).itemsize if count is None else count
        out = np.empty(n, dtype)
        if lib().pf_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes) != 0:
            raise RuntimeError("pf_memcpy_dtoh failed")
        return out

    def zero(self) -> None:
        if lib().pf_memset(self.ptr, 0, self.nbytes) != 0:
            raise RuntimeError("pf_memset failed")

    def close(self) -> None:
        if self.ptr:
            lib().pf_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_info() -> dict:
    L = lib()
    arch = C.create_string_buffer(64)
    L.pf_device_arch(arch, 64)
    return {"name": (L.pf_device_name() or b"").decode(), "arch": arch.value.decode(),
            "cus": L.pf_device_cu_count(), "count": L.pf_device_count()}


# ---------------------------------------------------------------- Miller-Rabin

class MrCandidate(C.Structure):
    """``mr_candidate_t`` (include/mr_gpu.h)."""
    _fields_ = [("a", C.c_int64), ("b", C.c_int64), ("c", C.c_int64),
                ("k", C.c_uint32), ("j", C.c_uint32), ("which", C.c_int32)]


def _mr_sigs(L):
    vp, i32, u32 = C.c_void_p, C.c_int, C.c_uint32
    sig = {
        "mr_ctx_create": (vp, [u32p, i32, i32, i32]),
        "mr_ctx_create_limbs16": (vp, [C.POINTER(C.c_uint64), i32, i32, i32]),
        "mr_ctx_create_sparse": (vp, [C.POINTER(MrCandidate), i32, i32]),
        "mr_run": (i32, [vp, u32p, i32, C.POINTER(C.c_uint8), u32p]),
        "mr_run_per_slot": (i32, [vp, u32p, C.POINTER(C.c_uint8), u32p]),
        "mr_run_ex": (i32, [vp, u32p, i32, i32, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), u32p]),
        "mr_ctx_info": (i32, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                              C.POINTER(C.c_int32)]),
        "mr_ctx_moduli": (i32, [vp, u32p]),
        "mr_ctx_destroy": (None, [vp]),
        "mr_debug_product": (i32, [vp, u32p, u32p, u32p]),
        "mr_debug_montmul": (i32, [vp, u32p, u32p, u32p]),
        "mr_debug_mulsmall": (i32, [vp, u32p, u32, u32p]),
        "mr_time_squarings": (i32, [vp, u32, i32, C.POINTER(C.c_float)]),
        "mr_set_min_log_len": (i32, [i32]),
        "mr_ctx_modmul_dev": (i32, [vp, vp, vp, vp, i32, vp]),
        "mr_ctx_modsq_dev": (i32, [vp, vp, vp, i32, vp]),
        "mr_ctx_check_passed_dev": (i32, [vp, vp, i32, vp, vp]),
        "mr_ctx_to_residue": (i32, [vp, C.POINTER(C.c_uint64), i32, C.POINTER(C.c_uint64), i32]),
        "mr_ctx_from_residue_dev": (i32, [vp, vp, i32, C.POINTER(C.c_uint64), i32]),
        "miller_rabin_test": (i32, [C.c_longlong, C.c_longlong, C.c_longlong, i32, i32,
                                    C.POINTER(MrTiming)]),
        "mr_sched_create": (vp, [C.POINTER(u32), i32, i32, i32, i32, C.c_int64]),
        "mr_sched_job_bytes": (C.c_int64, [C.c_int64]),
        "mr_sched_submit": (i32, [vp, C.POINTER(MrCandidate), C.POINTER(C.c_int64), i32]),
        "mr_sched_submit_survivors": (i32, [vp, C.c_longlong, C.c_longlong, C.c_longlong, vp,
                                            u32, i32, C.c_int64]),
        "mr_sched_pump": (i32, [vp, i32]),
        "mr_sched_results": (i32, [vp, C.POINTER(MrResult), i32]),
        "mr_sched_stats": (i32, [vp, C.POINTER(MrSchedStats)]),
        "mr_sched_destroy": (None, [vp]),
        "mr_sched_checkpoint": (i32, [vp, C.c_char_p]),
        "mr_sched_restore": (i32, [vp, C.c_char_p]),
        "mr_sched_set_bucket": (i32, [vp, i32, i32, i32]),
        "mr_tuning_read": (i32, [C.c_char_p, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), i32]),
        "mr_sched_load_tuning": (i32, [vp, C.c_char_p]),
        "mr_test_batch": (i32, [C.POINTER(MrCandidate), i32, C.POINTER(u32), i32, i32,
                                C.POINTER(C.c_uint8), C.POINTER(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def int_to_words(x: int, n_words: int) -> np.ndarray:
    if x < 0 or x.bit_length() > 32 * n_words:
        raise ValueError("integer does not fit")
    return np.frombuffer(x.to_bytes(4 * n_words, "little"), dtype=np.uint32).copy()


def words_to_int(w: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(w, dtype=np.uint32).tobytes(), "little")


class MrResult(C.Structure):
    """mr_result_t (include/mr_gpu.h)."""
    _fields_ = [("tag", C.c_int64), ("verdict", C.c_int32), ("rounds", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class MrSchedStats(C.Structure):
    """mr_sched_stats_t (include/mr_gpu.h)."""
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("pending", C.c_int64),
                ("active", C.c_int64), ("chunks", C.c_int64), ("admin_batches", C.c_int64),
                ("squarings", C.c_int64), ("hbm_bytes", C.c_int64),
                ("max_active", C.c_int32), ("n_pools", C.c_int32),
                ("job_bytes", C.c_int64), ("peak_job_bytes", C.c_int64),
                ("reclaimed_pools", C.c_int64)]


def _cands(cands):
    arr = (MrCandidate * len(cands))()
    for i, cd in enumerate(cands):
        a, b, c, k, j, w = cd
        arr[i] = MrCandidate(a, b, c, k, j, w)
    return arr


# The tuning table tune_mr.py writes for this package (per-bucket chunk and
# slots measured on the device).
TUNING_PATH = Path(__file__).resolve().parent / "mr_tuning.json"


def tuning_read(path) -> list[dict]:
    """The buckets of a tuning table as the library parses it
    (mr_tuning_read; no GPU call)."""
    L = mr_lib()
    cap = 16
    lg, ch, spc = (C.c_int32 * cap)(), (C.c_int32 * cap)(), (C.c_int32 * cap)()
    n = L.mr_tuning_read(os.fsencode(str(path)), lg, ch, spc, cap)
    if n < 0:
        raise ValueError(f"mr_tuning_read rejected {path}")
    return [{"log_len": lg[i], "chunk": ch[i], "slots_per_cu": spc[i]} for i in range(min(n, cap))]


class MrScheduler:
    """Continuous batching over mixed-size candidates (mr_sched_* C ABI)."""

    def __init__(self, bases, device: int = 0, max_active: int = 0, chunk: int = 0,
                 hbm_budget: int = 0, tuning=None):
        """tuning: a tuning table (path) whose per-bucket chunk and slots
        apply (mr_sched_load_tuning); "default" = TUNING_PATH when it exists."""
        L = mr_lib()
        self.L = L
        b = (C.c_uint32 * len(bases))(*bases)
        self.ptr = L.mr_sched_create(b, len(bases), device, max_active, chunk, hbm_budget)
        if not self.ptr:
            raise RuntimeError("mr_sched_create failed")
        self.tuned = 0
        if tuning == "default":
            tuning = TUNING_PATH if TUNING_PATH.exists() else None
        if tuning is not None:
            self.tuned = self.load_tuning(tuning)

    def set_bucket(self, log_len: int, chunk: int = 0, slots: int = 0):
        """Chunk and pool slots for the bucket of transform length 2^log_len
        (before its first admission)."""
        if self.L.mr_sched_set_bucket(self.ptr, log_len, chunk, slots) != 0:
            raise ValueError(f"mr_sched_set_bucket({log_len}, {chunk}, {slots}) rejected")

    def load_tuning(self, path) -> int:
        n = self.L.mr_sched_load_tuning(self.ptr, os.fsencode(str(path)))
        if n < 0:
            raise ValueError(f"mr_sched_load_tuning rejected {path}")
        return n

    def submit(self, cands, tags=None):
        arr = _cands(cands)
        t = None
        if tags is not None:
            t = (C.c_int64 * len(cands))(*tags)
        if self.L.mr_sched_submit(self.ptr, arr, t, len(cands)) != 0:
            raise ValueError("mr_sched_submit rejected the candidates")

    def submit_survivors(self, a, b, c, d_list, n, which=0, tag_base=0):
        """Queue the first n entries of a device survivor list ((k << 16) | j
        per entry, a device pointer or DeviceBuffer) as candidates of the
        triple (a, b, c), with tags tag_base + i; the list must outlive
        the jobs."""
        ptr = d_list.ptr if isinstance(d_list, DeviceBuffer) else d_list
        rc = self.L.mr_sched_submit_survivors(self.ptr, a, b, c, ptr, n, which, tag_base)
        if rc != 0:
            raise ValueError(f"mr_sched_submit_survivors rejected the list ({rc})")

    def pump(self, block: bool = True) -> int:
        left = self.L.mr_sched_pump(self.ptr, 1 if block else 0)
        if left < 0:
            raise RuntimeError(f"mr_sched_pump failed ({left})")
        return left

    def results(self, cap: int = 4096):
        out = (MrResult * cap)()
        k = self.L.mr_sched_results(self.ptr, out, cap)
        if k < 0:
            raise RuntimeError("mr_sched_results failed")
        return [out[i] for i in range(k)]

    def stats(self) -> MrSchedStats:
        st = MrSchedStats()
        if self.L.mr_sched_stats(self.ptr, C.byref(st)) != 0:
            raise RuntimeError("mr_sched_stats failed")
        return st

    def checkpoint(self, path) -> int:
        """Write the undecided candidates (and the progress of admitted
        ones) to path; returns how many."""
        n = self.L.mr_sched_checkpoint(self.ptr, os.fsencode(path))
        if n < 0:
            raise RuntimeError(f"mr_sched_checkpoint failed ({n})")
        return n

    def restore(self, path) -> int:
        """Queue the candidates of a checkpoint file; returns how many."""
        n = self.L.mr_sched_restore(self.ptr, os.fsencode(path))
        if n < 0:
            raise ValueError(f"mr_sched_restore rejected {path}")
        return n

    def run_all(self):
        res = []
        while True:
            left = self.pump(True)
            res += self.results()
            if left == 0:
                return res + self.results()

    def close(self):
        if getattr(self, "ptr", None):
            self.L.mr_sched_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mr_test_batch(cands, bases, device: int = 0):
    """verdicts (list of 0/1) and rounds run per candidate."""
    L = mr_lib()
    n = len(cands)
    v = (C.c_uint8 * n)()
    r = (C.c_int32 * n)()
    b = (C.c_uint32 * len(bases))(*bases)
    rc = L.mr_test_batch(_cands(cands), n, b, len(bases), device, v, r)
    if rc != 0:
        raise RuntimeError(f"mr_test_batch failed ({rc})")
    return list(v), list(r)


class MrTiming(C.Structure):
    """mr_timing_t (include/miller_rabin.h; reference miller_rabin.h:3-8)."""
    _fields_ = [("build_n", C.c_double), ("mr_n", C.c_double),
                ("build_revn", C.c_double), ("mr_revn", C.c_double)]


def miller_rabin_test(a: int, b: int, c: int, k: int, j: int):
    """The reference's Phase-2 call (miller_rabin.h:24-25) on the GPU:
    returns (code, MrTiming) with code 0 / 1 / 2 as the reference, -1 on
    a candidate it cannot form or a HIP error."""
    t = MrTiming()
    code = mr_lib().miller_rabin_test(a, b, c, k, j, C.byref(t))
    return code, t


def mr_lib():
    L = lib()
    if not hasattr(L, "_mr_ready"):
        _mr_sigs(L)
        L._mr_ready = True
    return L


class MrContext:
    """A batch of moduli resident on the GPU (include/mr_gpu.h).
    min_log_len forces a transform of at least 2^min_log_len points."""

    def __init__(self, moduli=None, n_words: int | None = None, device: int = 0,
                 candidates=None, min_log_len: int = 11):
        L = mr_lib()
        self.L = L
        if L.mr_set_min_log_len(min_log_len) != 0:
            raise ValueError("min_log_len out of range")
        if candidates is not None:
            arr = (MrCandidate * len(candidates))(*[MrCandidate(*c) for c in candidates])
            self.ptr = L.mr_ctx_create_sparse(arr, len(candidates), device)
        else:
            moduli = [int(m) for m in moduli]
            nw = n_words or max(1, max((m.bit_length() + 31) // 32 for m in moduli))
            w = np.concatenate([int_to_words(m, nw) for m in moduli])
            self.ptr = L.mr_ctx_create(w.ctypes.data_as(u32p), nw, len(moduli), device)
        L.mr_set_min_log_len(11)
        if not self.ptr:
            raise RuntimeError("mr_ctx_create failed")
        nw_, limbs, logm = C.c_int(), C.c_int(), C.c_int()
        L.mr_ctx_info(self.ptr, C.byref(nw_), C.byref(limbs), C.byref(logm), None)
        self.n_words, self.limbs, self.log_len = nw_.value, limbs.value, logm.value
        self.n = len(candidates) if candidates is not None else len(moduli)
        sv = (C.c_int32 * self.n)()
        L.mr_ctx_info(self.ptr, None, None, None, sv)
        self.s = list(sv)

    @property
    def R(self) -> int:
        return 1 << (24 * self.limbs)

    def moduli(self):
        out = np.zeros(self.n * self.n_words, np.uint32)
        assert self.L.mr_ctx_moduli(self.ptr, out.ctypes.data_as(u32p)) == 0
        return [words_to_int(out[i * self.n_words:(i + 1) * self.n_words]) for i in range(self.n)]

    def run(self, bases, residues: bool = True, all_bases: bool = False, passes: bool = False):
        """Rounds of each base (mr_run; all_bases: mr_run_ex with
        MR_RUN_ALL_BASES, every base on every candidate).  Returns (alive,
        residues or None) and, with passes, the per-base pass flags too."""
        bases = np.asarray(bases, dtype=np.uint32)
        alive = np.zeros(self.n, np.uint8)
        res = np.zeros(self.n * len(bases) * self.n_words, np.uint32) if residues else None
        pas = np.zeros(self.n * max(1, len(bases)), np.uint8)
        u8 = C.POINTER(C.c_uint8)
        if all_bases or passes:
            rc = self.L.mr_run_ex(self.ptr, bases.ctypes.data_as(u32p), len(bases),
                                  1 if all_bases else 0, alive.ctypes.data_as(u8),
                                  pas.ctypes.data_as(u8),
                                  res.ctypes.data_as(u32p) if residues else None)
        else:
            rc = self.L.mr_run(self.ptr, bases.ctypes.data_as(u32p), len(bases),
                               alive.ctypes.data_as(u8), res.ctypes.data_as(u32p) if residues else None)
        if rc != 0:
            raise RuntimeError(f"mr_run failed ({rc})")
        rr = None
        if residues:
            nw, nb = self.n_words, len(bases)
            rr = [[words_to_int(res[(i * nb + b) * nw:(i * nb + b + 1) * nw]) for b in range(nb)]
                  for i in range(self.n)]
        if passes:
            nb = len(bases)
            pp = [[bool(pas[i * nb + b]) for b in range(nb)] for i in range(self.n)]
            return alive.astype(bool).tolist(), rr, pp
        return alive.astype(bool).tolist(), rr

    def run_per_slot(self, bases, residues: bool = True):
        """One round, candidate i with base bases[i] (mr_run_per_slot).
        Returns (pass flags, residues or None)."""
        bases = np.asarray(bases, dtype=np.uint32)
        assert len(bases) == self.n
        ok = np.zeros(self.n, np.uint8)
        res = np.zeros(self.n * self.n_words, np.uint32) if residues else None
        rc = self.L.mr_run_per_slot(self.ptr, bases.ctypes.data_as(u32p),
                                    ok.ctypes.data_as(C.POINTER(C.c_uint8)),
                                    res.ctypes.data_as(u32p) if residues else None)
        if rc != 0:
            raise RuntimeError(f"mr_run_per_slot failed ({rc})")
        rr = None
        if residues:
            nw = self.n_words
            rr = [words_to_int(res[i * nw:(i + 1) * nw]) for i in range(self.n)]
        return ok.astype(bool).tolist(), rr

    def _words(self, xs):
        return np.concatenate([int_to_words(int(x), self.n_words) for x in xs])

    def debug_product(self, a, b):
        out = np.zeros(self.n * 2 * self.n_words, np.uint32)
        A, B = self._words(a), self._words(b)
        assert self.L.mr_debug_product(self.ptr, A.ctypes.data_as(u32p), B.ctypes.data_as(u32p),
                                       out.ctypes.data_as(u32p)) == 0
        w = 2 * self.n_words
        return [words_to_int(out[i * w:(i + 1) * w]) for i in range(self.n)]

    def debug_montmul(self, a, b):
        out = np.zeros(self.n * self.n_words, np.uint32)
        A, B = self._words(a), self._words(b)
        assert self.L.mr_debug_montmul(self.ptr, A.ctypes.data_as(u32p), B.ctypes.data_as(u32p),
                                       out.ctypes.data_as(u32p)) == 0
        w = self.n_words
        return [words_to_int(out[i * w:(i + 1) * w]) for i in range(self.n)]

    def debug_mulsmall(self, a, s: int):
        out = np.zeros(self.n * self.n_words, np.uint32)
        A = self._words(a)
        assert self.L.mr_debug_mulsmall(self.ptr, A.ctypes.data_as(u32p), s,
                                        out.ctypes.data_as(u32p)) == 0
        w = self.n_words
        return [words_to_int(out[i * w:(i + 1) * w]) for i in range(self.n)]

    def time_squarings(self, base: int, n_bits: int) -> float:
        """Milliseconds for the first n_bits squaring steps of a round."""
        ms = C.c_float()
        if self.L.mr_time_squarings(self.ptr, base, n_bits, C.byref(ms)) != 0:
            raise RuntimeError("mr_time_squarings failed")
        return ms.value

    def close(self):
        if getattr(self, "ptr", None):
            self.L.mr_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
