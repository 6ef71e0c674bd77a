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
        "trial_div_check_error": (C.c_int, []),
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
