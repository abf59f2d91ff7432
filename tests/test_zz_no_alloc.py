"""The allocation-free entry point pacg_convolve_fft_based_noalloc, through the
C-ABI, against the allocating pacg_convolve_fft_based on the same inputs (FP64
bar 1e-12 relative L2; the same kernels run, so the results are expected equal
bit for bit) and against the oracle: two calls in a row, the second on a
smaller plan, then a plan too large for the fixed scratch, which must be
rejected before any GPU work."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def rnd(seed, shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


def run(gpu, fn, pl, x, h, dtype):
    dx = gpu.DeviceComplex.from_numpy(x, dtype)
    dh = gpu.DeviceComplex.from_numpy(h, dtype)
    rc = fn(pl.ptr, dx.re.ptr, dx.im.ptr, dh.re.ptr, dh.im.ptr)
    assert rc == 0, gpu.last_error()
    pl.sync()
    return dx.numpy(), dh.numpy()


# (dims, schedules, precision): the second case is smaller than the first, so
# the second call reads scratch the first one left behind
CASES = [((64, 256), [[4, 4, 4], [8, 8, 4]], 0), ((512,), [[8, 8, 8]], 0), ((32, 16, 8), [[8, 4], [4, 4], [8]], 1)]


def test_noalloc_matches_the_allocating_entry_point(gpu, oracle):
    noalloc = gpu.lib.pacg_convolve_fft_based_noalloc
    alloc = gpu.lib.pacg_convolve_fft_based
    for seed, (dims, scheds, prec) in enumerate(CASES):
        x, h = rnd(10 + seed, dims), rnd(20 + seed, dims)
        dtype = np.float64 if prec == 0 else np.float32
        pl = gpu.Plan(dims, schedules=scheds, precision=prec)
        ya, ha = run(gpu, alloc, pl, x, h, dtype)
        yb, hb = run(gpu, noalloc, pl, x, h, dtype)
        tol = TOL64 if prec == 0 else 2e-6
        assert rel_l2(yb, ya) <= (TOL64 if prec == 0 else 1e-7), (dims, rel_l2(yb, ya))
        assert rel_l2(hb, ha) <= (TOL64 if prec == 0 else 1e-7), (dims, rel_l2(hb, ha))
        if prec == 0:
            assert np.array_equal(yb, ya) and np.array_equal(hb, ha), dims
        y_ref, _ = oracle.fft_based_convolution(x, h, scheds)
        assert rel_l2(yb, y_ref) < tol, (dims, rel_l2(yb, y_ref))
        assert rel_l2(hb, np.fft.fftn(h)) < tol
        pl.close()


def test_noalloc_rejects_a_plan_beyond_its_scratch(gpu):
    # 2^22 FP64 points need 2 x 32 MiB of scratch, twice the fixed 32 MiB
    pl = gpu.Plan((1 << 11, 1 << 11), schedules=[[4, 4, 4, 4, 4, 2]] * 2, precision=0)
    dx = gpu.DeviceComplex((4,), np.float64)  # never touched: the size check comes first
    rc = gpu.lib.pacg_convolve_fft_based_noalloc(pl.ptr, dx.re.ptr, dx.im.ptr, dx.re.ptr, dx.im.ptr)
    assert rc == 2 and "scratch" in gpu.last_error()
    pl.close()
