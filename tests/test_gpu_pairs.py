"""Pair launches: passes of two axes in one kernel (3D plans).

With pair kernels a 3D convolution takes three launches: [top stages of axis 1
+ all of axis 2], [all of axis 0 + bottom stage of axis 1] fused with the
spectral product, and the conj(A) mirror of the first.  Checked against the
oracle (FP64 <= 1e-12, FP32 <= 2e-6 relative L2) at sizes it finishes in
seconds, and against the one-axis-per-launch plan of the same schedules."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2
from oracle import A, CONJ, FFT, FFT2, IFFT, TRANS, UIFFT

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6

CASES = [
    ((64, 256, 64), [[8, 8], [8, 8, 4], [8, 8]]),
    ((64, 256, 64), [[8, 8], [8, 4, 8], [8, 8]]),
    ((256, 256, 64), [[4, 8, 8], [8, 4, 8], [8, 8]]),
]
# FP32 only: 16-value tiles for [axis 1 top 8-8 + axis 2], and radix-4 bottom
# stages of axis 1 beside 8 or 16 columns of axis 2
CASES32 = [
    ((256, 256, 256), [[8, 8, 4], [4, 8, 8], [4, 8, 8]]),
    ((256, 128, 256), [[4, 8, 8], [4, 8, 4], [8, 8, 4]]),
]


def rnd(seed, shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


@pytest.mark.parametrize("dims,scheds", CASES + CASES32)
def test_pair_plan_has_three_launches(gpu, dims, scheds):
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, pairs=True)
    assert pl.launches_per_convolve() == 3, pl.describe()
    assert "pair" in pl.describe()
    # one axis per launch: at least 2d - 1 = 5 (7 where axis 1 needs two passes)
    assert gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32).launches_per_convolve() >= 5
    # FP64 pairs on tiles up to 4096 points (an outer axis of 64 here)
    if dims[0] == 64:
        assert gpu.Plan(dims, schedules=scheds, pairs=True).launches_per_convolve() == 3


@pytest.mark.parametrize("dims,scheds", CASES)
def test_pair_convolve_fp64(gpu, oracle, dims, scheds):
    x = rnd(11, dims)
    k = rnd(12, dims)
    y_ref, s_ref = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, pairs=True)
    pl.set_kernel(k)
    assert rel_l2(pl.spectrum(), s_ref) < TOL64
    dx = gpu.DeviceComplex.from_numpy(x)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < TOL64
    pl.convolve(dx)  # in place
    assert rel_l2(dx.numpy(), y_ref) < TOL64


@pytest.mark.parametrize("dims,scheds", CASES[:2])
def test_pair_operators_fp64(gpu, oracle, dims, scheds):
    x = rnd(13, dims)
    pl = gpu.Plan(dims, schedules=scheds, pairs=True)
    ops = [(gpu.OP_TRANS, lambda: oracle.butterflies_nd(x, scheds, TRANS)),
           (gpu.OP_CONJ, lambda: oracle.butterflies_nd(x, scheds, CONJ)),
           (gpu.OP_FWD, lambda: oracle.butterflies_nd(x, scheds, A)),
           (gpu.OP_FFT, lambda: oracle.transform(x, scheds, FFT)),
           (gpu.OP_FFT2, lambda: oracle.transform(x, scheds, FFT2)),
           (gpu.OP_IFFT, lambda: oracle.transform(x, scheds, IFFT)),
           (gpu.OP_UIFFT, lambda: oracle.transform(x, scheds, UIFFT))]
    for op, ref in ops:
        dx = gpu.DeviceComplex.from_numpy(x)
        dy = pl.alloc()
        pl.execute(op, dx, dy)
        assert rel_l2(dy.numpy(), ref()) < TOL64, op


@pytest.mark.parametrize("sprec", [0, 1])
@pytest.mark.parametrize("dims,scheds", CASES)
def test_pair_convolve_fp32(gpu, oracle, dims, scheds, sprec):
    x = rnd(14, dims)
    k = rnd(15, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=sprec, pairs=True)
    assert pl.launches_per_convolve() == 3
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < TOL32


@pytest.mark.parametrize("dims,scheds", CASES32)
def test_pair_wide_tiles_fp32(gpu, oracle, dims, scheds):
    """Against the FP64 plan (itself at 1e-12 of the oracle on the smaller cases) for the
    convolution and the three operators, on the 16-value and multi-line tiles."""
    x = rnd(20, dims)
    k = rnd(21, dims)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=gpu.PREC_F32, pairs=True)
    assert pl.launches_per_convolve() == 3, pl.describe()
    ref = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F64)
    pl.set_kernel(k)
    ref.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    rx = gpu.DeviceComplex.from_numpy(x)
    ry = ref.alloc()
    ref.convolve(rx, ry)
    assert rel_l2(dy.numpy(), ry.numpy()) < TOL32
    for op in (gpu.OP_TRANS, gpu.OP_CONJ, gpu.OP_FWD):
        pl.execute(op, dx, dy)
        ref.execute(op, rx, ry)
        assert rel_l2(dy.numpy(), ry.numpy()) < TOL32, op


def test_pairs_match_one_axis_per_launch(gpu):
    dims, scheds = CASES[0]
    x = rnd(16, dims)
    k = rnd(17, dims)
    out = []
    for pairs in (True, False):
        pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=gpu.PREC_F32, pairs=pairs)
        pl.set_kernel(k)
        dx = gpu.DeviceComplex.from_numpy(x, np.float32)
        pl.convolve(dx)
        out.append(dx.numpy())
    assert rel_l2(out[0], out[1]) < 1e-6


def test_pair_graph_replay(gpu, oracle):
    dims, scheds = CASES[0]
    x = rnd(18, dims)
    k = rnd(19, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=gpu.PREC_F32, pairs=True)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.capture(dx, dy)
    for _ in range(3):
        dy.upload(np.zeros(dims, complex))
        pl.replay()
        pl.sync()
        assert rel_l2(dy.numpy(), y_ref) < TOL32
        pl.convolve(dx, dy)  # eager launch between replays
        pl.sync()
