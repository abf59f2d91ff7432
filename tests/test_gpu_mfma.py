"""Radix-8 stages on the matrix cores (F_8 (x) I, split FP16, FP32 accumulation)
against the FP64 oracle.  FP32 bar: rel-L2 <= 2e-6."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2
from oracle import A, CONJ, TRANS

pytestmark = pytest.mark.gpu

TOL32 = 2e-6

CASES = [
    ((512,), [[8, 8, 8]]),
    ((64, 256), [[8, 8], [8, 8, 4]]),
    ((8, 64, 32), [[8], [8, 8], [4, 8]]),
    ((4096, 16), [[8, 8, 8, 8], [4, 4]]),
    ((32, 8), [[4, 8], [8]]),
    ((256, 64), [[4, 8, 8], [2, 8, 4]]),
    ((16, 512), [[2, 8], [8, 8, 8]]),
]


def test_mfma_plan_one_pass_per_axis(gpu):
    """4096 x 256 (8-8-8-8 / 8-8-4): the MFMA plan cuts the axes like the VALU plan (5 launches)."""
    scheds = [[8, 8, 8, 8], [8, 8, 4]]
    mf = gpu.Plan((4096, 256), schedules=scheds, precision=gpu.PREC_F32, mfma=True)
    va = gpu.Plan((4096, 256), schedules=scheds, precision=gpu.PREC_F32)
    assert mf.launches_per_convolve() == va.launches_per_convolve() == 5, mf.describe()


def test_mfma_plan_autotune_keeps_one_load_policy(gpu, oracle):
    """MFMA pass kernels have no streamed-input form: the autotuner leaves the first launch's load
    policy alone on an MFMA plan whose first pass has an MFMA stage, and the tuned plan keeps parity."""
    dims, scheds = (512, 256), [[8, 8, 8], [8, 8, 4]]  # first launch: axis 1, one MFMA pass 8-8-4
    x = rnd(35, dims)
    k = rnd(36, dims)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, mfma=True)
    assert pl.describe().split("]")[0].count("mfma") == 1, pl.describe()
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.autotune(dx, dy, 2)
    assert "x:nt" not in pl.describe(), pl.describe()
    pl.convolve(dx, dy)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    assert rel_l2(dy.numpy(), y_ref) < TOL32


def rnd(seed, shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


@pytest.mark.parametrize("dims,scheds", CASES)
@pytest.mark.parametrize("sprec", [0, 1])
def test_mfma_convolution(gpu, oracle, dims, scheds, sprec):
    x = rnd(31, dims)
    k = rnd(32, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=sprec, mfma=True)
    assert "mfma" in pl.describe()
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < TOL32


@pytest.mark.parametrize("dims,scheds", CASES)
def test_mfma_operators(gpu, oracle, dims, scheds):
    x = rnd(33, dims)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, mfma=True)
    for op, kind in ((gpu.OP_TRANS, TRANS), (gpu.OP_CONJ, CONJ), (gpu.OP_FWD, A)):
        dx = gpu.DeviceComplex.from_numpy(x, np.float32)
        dy = pl.alloc()
        pl.execute(op, dx, dy)
        assert rel_l2(dy.numpy(), oracle.butterflies_nd(x, scheds, kind)) < TOL32, op


def test_mfma_block_scaling_handles_wide_range(gpu, oracle):
    """Values from 1e-30 to 1e30 in one array: per-butterfly scaling keeps FP16 in range."""
    dims = (64, 64)
    scheds = [[8, 8], [8, 8]]
    g = np.random.default_rng(5)
    x = rnd(34, dims) * 10.0 ** g.integers(-30, 30, dims)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, mfma=True)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    pl.execute(gpu.OP_TRANS, dx)
    ref = oracle.butterflies_nd(x.astype(np.complex64).astype(np.complex128), scheds, TRANS)
    assert rel_l2(dx.numpy(), ref) < TOL32


@pytest.mark.parametrize("scale", [1.5 * 2.0 ** 127, 2.0 ** 127, 1.0])
def test_mfma_block_scaling_near_float_max(gpu, scale):
    """A delta of magnitude up to near FLT_MAX: every butterfly of A^T sees one nonzero input, so
    the exact result is that value at every position; the block scale must stay a finite FP32
    factor (2^128 is not)."""
    n = 512
    x = np.zeros(n, complex)
    x[0] = scale * (1 - 0.5j)
    pl = gpu.Plan((n,), schedules=[[8, 8, 8]], precision=gpu.PREC_F32, mfma=True)
    assert "mfma" in pl.describe()
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    pl.execute(gpu.OP_TRANS, dx)
    y = dx.numpy()
    want = np.full(n, np.complex64(x[0]), dtype=complex)
    assert np.all(np.isfinite(y))
    assert rel_l2(y, want) < 2e-6, rel_l2(y, want)
    pl.close()


def test_anisotropic_4096x256_full_size(gpu, synth):
    """Config 4: 4096 x 256, U(-1,1), seed 3, radix 8-8-8-8 and 8-8-4, MFMA radix-8 stages."""
    x, k = synth.workload_inputs("2d_4096x256")
    pl = gpu.Plan(x.shape, schedules=[[8, 8, 8, 8], [8, 8, 4]], precision=gpu.PREC_F32,
                  spectrum_precision=gpu.PREC_F32, mfma=True)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    ref = np.fft.ifft2(np.fft.fft2(x) * np.fft.fft2(k))
    assert rel_l2(dy.numpy(), ref) < TOL32
