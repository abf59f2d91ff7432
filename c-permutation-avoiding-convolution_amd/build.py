"""Build libpacg.so: hand-written gfx950 HIP kernels + the C host layer.

    python build.py [--jobs N] [--incremental]

1. gen_variants.py writes the pass-kernel instantiation units into csrc/gen/;
2. every .hip unit is compiled by hipcc for gfx950 (parallel jobs);
3. the C host layer (plans, drop-in fft.h entry points, helpers) by gcc;
4. hipcc links the shared library lib/libpacg.so;
5. gcc builds the verify/bench driver programs into bin/ from drivers/bench_driver.c,
   linked against lib/libpacg.so through the fft.h names.
--incremental skips objects newer than their sources (developer loop only).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
GEN = os.path.join(CSRC, "gen")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
DRIVER_SRC = os.path.join(PKG, "drivers", "bench_driver.c")
# program name -> defines; the names and command lines of the reference's driver programs
DRIVERS = {
    "bench1d_convolution": ["-DPACG_DIM=1", "-DPACG_KIND=PACG_CONV"],
    "bench1d_forward_fft": ["-DPACG_DIM=1", "-DPACG_KIND=PACG_FWD"],
    "bench1d_backward_fft": ["-DPACG_DIM=1", "-DPACG_KIND=PACG_BWD"],
    "bench2d_convolution": ["-DPACG_DIM=2", "-DPACG_KIND=PACG_CONV"],
    "bench2d_convolution_anisotropic": ["-DPACG_DIM=2", "-DPACG_KIND=PACG_CONV", "-DPACG_ANISO"],
    "bench2d_forward_fft": ["-DPACG_DIM=2", "-DPACG_KIND=PACG_FWD"],
    "bench2d_backward_fft": ["-DPACG_DIM=2", "-DPACG_KIND=PACG_BWD"],
    "bench3d_convolution": ["-DPACG_DIM=3", "-DPACG_KIND=PACG_CONV"],
    "bench3d_convolution_anisotropic": ["-DPACG_DIM=3", "-DPACG_KIND=PACG_CONV", "-DPACG_ANISO"],
    "bench3d_forward_fft": ["-DPACG_DIM=3", "-DPACG_KIND=PACG_FWD"],
    "bench3d_backward_fft": ["-DPACG_DIM=3", "-DPACG_KIND=PACG_BWD"],
}
INCLUDE = os.path.join(ROOT, "include")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
             # keep hipcc from forming packed FP32 ops on its own
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", CSRC, "-I", INCLUDE]
C_FLAGS = ["-O2", "-std=c11", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I", CSRC, "-I", INCLUDE]
HOST_C = ["pacg_plan.c", "pacg_host.c", "pacg_dropin.c", "pacg_batch.c", "pacg_search.c"]
HIP_SRC = ["pacg_aux.hip", "pacg_batch.hip", "pacg_mfma.hip"]


def _newer(src: str, obj: str, deps) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src, *deps])


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(jobs: int | None = None, incremental: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    sys.path.insert(0, CSRC)
    try:
        import gen_variants  # noqa: E402
    finally:
        sys.path.pop(0)
    gen_variants.main(GEN)
    hdrs = [os.path.join(CSRC, h) for h in ("pacg_pass.hpp", "pacg_internal.h", "pacg_host.h",
                                            "pacg_batch.h")]
    hdrs += [os.path.join(INCLUDE, h) for h in ("pacg.h", "pacg_fft.h")]
    tasks = []
    objs = []
    hip_units = [os.path.join(CSRC, s) for s in HIP_SRC]
    hip_units += sorted(os.path.join(GEN, f) for f in os.listdir(GEN) if f.endswith(".hip"))
    for src in hip_units:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        objs.append(obj)
        if not incremental or _newer(src, obj, hdrs):
            tasks.append([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    c_units = [os.path.join(CSRC, s) for s in HOST_C] + [os.path.join(GEN, "pacg_variants_index.c")]
    for src in c_units:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        objs.append(obj)
        if not incremental or _newer(src, obj, hdrs):
            tasks.append(["gcc", *C_FLAGS, "-c", src, "-o", obj])
    jobs = jobs or min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in cf.as_completed([ex.submit(_run, t) for t in tasks]):
            f.result()
    out = os.path.join(LIB, "libpacg.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-lm"])
    os.replace(tmp, out)
    build_drivers(jobs)
    if verbose:
        print("built %s (%d units compiled) and %d driver programs in %s" % (out, len(tasks), len(DRIVERS), BIN))
    return out


def build_drivers(jobs: int | None = None) -> None:
    """The verify/bench programs; each finds lib/libpacg.so through its rpath."""
    os.makedirs(BIN, exist_ok=True)
    tasks = []
    for name, defs in DRIVERS.items():
        tasks.append(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", "-I", INCLUDE, *defs, DRIVER_SRC,
                      "-o", os.path.join(BIN, name), "-L", LIB, "-lpacg", "-Wl,-rpath,$ORIGIN/../lib", "-lm"])
    with cf.ThreadPoolExecutor(max_workers=jobs or min(16, os.cpu_count() or 4)) as ex:
        for f in cf.as_completed([ex.submit(_run, t) for t in tasks]):
            f.result()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    a = ap.parse_args()
    build(a.jobs, a.incremental)
