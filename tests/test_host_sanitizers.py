"""The host C layer (pacg_host.c; pacg_batch.c's items, dealing and autotuner
with the HIP glue stubbed) built with gcc under AddressSanitizer and
UndefinedBehaviorSanitizer and run on the CPU: tests/host_sanitize.c."""
from __future__ import annotations

import ctypes as C
import os
import shutil
import subprocess

import pytest
from conftest import PKG_DIR, ROOT

CSRC = os.path.join(PKG_DIR, "csrc")


class _Variant(C.Structure):
    _fields_ = [("log2n", C.c_int), ("variant", C.c_int), ("nstages", C.c_int), ("radix", C.c_int * 7),
                ("tpl", C.c_int), ("cost", C.c_int), ("cost_mixed", C.c_int)]


def test_host_layer_under_asan_and_ubsan(pacg, tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    n = C.c_int.in_dll(pacg.lib, "pacg_num_batch_variants").value
    rows = (_Variant * n).in_dll(pacg.lib, "pacg_batch_variants")
    assert n >= 14
    lines = ["/* written by tests/test_host_sanitizers.py from the built library */",
             "const pacg_batch_variant_t pacg_batch_variants[] = {"]
    for v in rows:
        lines.append("  {%d, %d, %d, {%s}, %d, %d, %d}," % (v.log2n, v.variant, v.nstages,
                                                          ", ".join(str(r) for r in v.radix), v.tpl, v.cost,
                                                          v.cost_mixed))
    lines += ["};", "const int pacg_num_batch_variants = %d;" % n]
    (tmp_path / "pacg_batch_variants.h").write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "host_sanitize")
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", str(tmp_path), "-I", CSRC, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host_sanitize.c"), os.path.join(CSRC, "pacg_host.c"), "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "OK" in r.stdout, (r.stdout[-2000:], r.stderr[-4000:])
