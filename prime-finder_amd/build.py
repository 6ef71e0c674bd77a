"""Build recipe for the gfx950 library ``lib/libpfamd.so``.

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only, every
``csrc/*.c`` file by gcc, every ``csrc/*.cpp`` file (host-only code) by g++,
and the objects are linked into one shared library
whose exported C symbols are declared in ``include/*.h``.  The compiled
callers of ``callers/`` (``worker_dropin``, ``bench_gpu_dropin``,
``mr_cpp_check``: the reference's worker, bench and runner call sites) are
linked against it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
CALLERS = PKG / "callers"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libpfamd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled in every HIP object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-Wall", "-Wextra", f"-I{ROOT / 'include'}"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build(verbose: bool = False) -> Path:
    """Compile all sources and link ``lib/libpfamd.so``; returns its path."""
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OBJ / (src.stem + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted(CSRC.glob("*.c")):
        obj = OBJ / (src.stem + ".c.o")
        jobs.append(["gcc", *C_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJ / (src.stem + ".cpp.o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    if verbose:
        for j in jobs:
            print(" ".join(j))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs) or 1)) as ex:
        list(ex.map(_run, jobs))
    objs = [str(OBJ / (s.stem + ".o")) for s in sorted(CSRC.glob("*.hip"))]
    objs += [str(OBJ / (s.stem + ".c.o")) for s in sorted(CSRC.glob("*.c"))]
    objs += [str(OBJ / (s.stem + ".cpp.o")) for s in sorted(CSRC.glob("*.cpp"))]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs, "-pthread"])
    shutil.move(str(tmp), str(LIB))
    # compiled callers of the drop-in boundary (callers/): the reference's
    # worker (C), its tests/bench_gpu.cu (C++, trial_div_launch) and its
    # runner surface (C++, include/batch_mod_ctx.hpp)
    link = [f"-L{LIB.parent}", "-lpfamd", "-Wl,-rpath,$ORIGIN"]
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", f"-I{ROOT / 'include'}",
          str(CALLERS / "worker_dropin.c"), "-o", str(LIB.parent / "worker_dropin"), *link])
    for name in ("bench_gpu_dropin", "mr_cpp_check"):
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}",
              str(CALLERS / f"{name}.cpp"), "-o", str(LIB.parent / name), *link])
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
