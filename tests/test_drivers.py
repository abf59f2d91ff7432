"""The verify/bench driver programs in bin/, the counterparts of the reference's
bench{1,2,3}d_*.c programs, linked against libpacg.so through the fft.h names.

CPU: the command-line handling of the reference programs (usage line, unknown
mode, unknown option, all exit status 1) and a loud failure when no GPU is
visible.  GPU: every program's verify mode at small sizes, whose maximum error
against the programs' own CPU FFT must stay at rounding level, and bench mode,
whose lines must have the reference's wording and value column."""
from __future__ import annotations

import os
import re
import subprocess

import pytest
from conftest import PKG_DIR

BIN = os.path.join(PKG_DIR, "bin")

# column where the value starts in the "Average ..." lines of each reference program
# (bench1d_convolution.c:264, bench2d_convolution.c:237, bench3d_convolution.c:246,
# bench1d_forward_fft.c:161, bench1d_backward_fft.c:173, bench2d_forward_fft.c:154,
# bench2d_backward_fft.c:162, bench3d_forward_fft.c:163, bench3d_backward_fft.c:171)
WIDTH = {
    "bench1d_convolution": 46, "bench2d_convolution": 46, "bench2d_convolution_anisotropic": 46,
    "bench3d_convolution": 49, "bench3d_convolution_anisotropic": 49,
    "bench1d_forward_fft": 38, "bench1d_backward_fft": 38,
    "bench2d_forward_fft": 40, "bench2d_backward_fft": 38,
    "bench3d_forward_fft": 40, "bench3d_backward_fft": 40,
}
AVG = re.compile(r"^(Average radix-(\d) ([a-zA-Z -]+) time:) +(\d+\.\d{6}) \+/- (\d+\.\d{6}) seconds$")
ERR = re.compile(r"^Max error (FFT-based|Permutation-avoiding|FFT) radix-(\d) vs CPU FFT: (\d\.\d{3}e[+-]\d\d)$")


def run(name, *args, timeout=120):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail("%s is not built (run the package build)" % path)
    print("driver:", name, *args, flush=True)
    return subprocess.run([path, *args], capture_output=True, text=True, timeout=timeout)


def npts(name, t):
    d = int(name[5])
    if d == 1:
        return 1 << t
    rest = 2 if "anisotropic" in name else t
    return 1 << (t + rest * (d - 1))


# ---- CPU ---------------------------------------------------------------------------------


@pytest.mark.parametrize("name", sorted(WIDTH))
def test_command_line_errors(pacg, name):
    r = run(name)
    assert r.returncode == 1 and r.stderr.startswith("Usage: ") and "[verify|bench] [-t power] [-r timing_runs]" \
        in r.stderr
    r = run(name, "check")
    assert r.returncode == 1 and r.stderr == "Unknown mode: check\n"
    r = run(name, "bench", "-q")
    assert r.returncode == 1 and r.stderr == "Unknown or malformed option: -q\n"
    r = run(name, "bench", "-t")  # value missing: malformed
    assert r.returncode == 1 and r.stderr == "Unknown or malformed option: -t\n"


def test_no_gpu_fails_loudly(pacg):
    if pacg.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = run("bench1d_convolution", "verify", "-t", "6")
    assert r.returncode == 1 and "failed:" in r.stderr
    assert "Max error" not in r.stdout


# ---- GPU ---------------------------------------------------------------------------------

VERIFY = [
    ("bench1d_convolution", 12, [2, 4, 8]),
    ("bench1d_convolution", 7, [2]),
    ("bench1d_forward_fft", 12, [2, 4, 8]),
    ("bench1d_backward_fft", 9, [2, 8]),
    ("bench2d_convolution", 6, [2, 4]),
    ("bench2d_convolution_anisotropic", 7, [2]),
    ("bench2d_convolution_anisotropic", 8, [2, 4]),
    ("bench2d_forward_fft", 6, [2, 4]),
    ("bench2d_backward_fft", 5, [2]),
    ("bench3d_convolution", 4, [2, 4]),
    ("bench3d_convolution_anisotropic", 6, [2, 4]),
    ("bench3d_forward_fft", 4, [2, 4]),
    ("bench3d_backward_fft", 4, [2, 4]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,t,radices", VERIFY)
def test_driver_verify(gpu, name, t, radices):
    r = run(name, "verify", "-t", str(t))
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    conv = "convolution" in name
    want = [(w, rr) for rr in radices for w in (("FFT-based", "Permutation-avoiding") if conv else ("FFT",))]
    assert len(lines) == len(want), r.stdout
    # values are sums of up to npts terms of size <= 1 (inputs uniform in [0, 1])
    tol = 1e-13 * npts(name, t) * max(t, 4)
    for line, (who, rr) in zip(lines, want):
        m = ERR.match(line)
        assert m, line
        assert (m.group(1), int(m.group(2))) == (who, rr), line
        assert float(m.group(3)) < tol, line


BENCH = [
    ("bench1d_convolution", 12, [2, 4, 8], ["FFT-based", "permutation-avoiding"]),
    ("bench2d_forward_fft", 6, [2, 4], ["FFT", "permutation", "unordered FFT"]),
    ("bench1d_backward_fft", 9, [2, 8], ["FFT", "permutation", "unordered FFT"]),
    ("bench3d_convolution_anisotropic", 4, [2, 4], ["FFT", "permutation-avoiding"]),
]


@pytest.mark.gpu
@pytest.mark.parametrize("name,t,radices,steps", BENCH)
def test_driver_bench_lines(gpu, name, t, radices, steps):
    r = run(name, "bench", "-t", str(t), "-r", "3")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert len(lines) == len(radices) * len(steps), r.stdout
    k = 0
    for rr in radices:
        for step in steps:
            if name == "bench3d_convolution_anisotropic" and rr == 4 and step == "permutation-avoiding":
                step = "permutation-avoidning"  # the reference program's own wording at :314
            m = AVG.match(lines[k])
            assert m, lines[k]
            assert int(m.group(2)) == rr and m.group(3) == step, lines[k]
            assert m.start(4) == WIDTH[name], lines[k]
            mu, sd = float(m.group(4)), float(m.group(5))
            assert 0.0 <= mu < 10.0 and sd >= 0.0, lines[k]  # a host permutation of 512 points prints 0.000000
            k += 1
