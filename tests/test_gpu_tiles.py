"""Every compiled tile shape of every launch gives the same convolution.

The autotuner may pick any kernel of a pass (lines per workgroup, hence LDS
layout: padded, unpadded rows, row-parity or lds_swz swizzles; points per
thread, for the FP64 2048-point lines).  Each launch
is pinned to each of its candidates in turn; the output must match the
default tiles to FP32 rounding, and the default tiles match the oracle
(FP32 <= 2e-6, FP64 <= 1e-12 relative L2)."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu

CASES = [
    # contiguous [8,8,4] (lds_swz), strided [4,8,8] (rows / row-parity), fused strided [8,8]
    ((64, 256, 256), [[8, 8], [4, 8, 8], [8, 8, 4]], 1),
    # the 256-point fused strided tile of the 256^3 plan
    ((256, 64, 256), [[4, 8, 8], [8, 8], [8, 8, 4]], 1),
    # 2048-point contiguous lines (16 points per thread), radix-4 strided pass
    ((1024, 2048), [[4, 4, 4, 4, 4], [4, 4, 4, 4, 4, 2]], 1),
    # FP64 tiles keep the padded layouts
    ((64, 64, 256), [[8, 8], [8, 8], [4, 8, 8]], 0),
    # FP64 2048-point contiguous lines: split re/im exchange, 16- and 8-point phases
    ((64, 2048), [[8, 8], [4, 4, 4, 4, 4, 2]], 0),
    # 1D 2^16 FP64 (the drop-in's radix-2 and radix-4 plans): 256 lines per pass, so the plan
    # starts on the narrow tiles that give the most workgroups
    ((65536,), [[2] * 16], 0),
    ((65536,), [[4] * 8], 0),
    # FP64 strided 2048-point lines (2-, 4- and 8-line tiles, 8 or 16 points per thread, rows
    # folded in LDS): the fused pass on axis 0 (on request), then A^T / conj(A) passes on axis 0
    # with the product on axis 1 (the default)
    ((2048, 1024), [[4, 4, 4, 4, 4, 2], [4, 4, 4, 4, 4]], 0, 0),
    ((2048, 1024), [[4, 4, 4, 4, 4, 2], [4, 4, 4, 4, 4]], 0),
    # ... on an axis whose inner size (2) is smaller than a tile's line count: one tile spans
    # several outer indices
    ((512, 2048, 2), [[8, 8, 8], [4, 4, 4, 4, 4, 2], [2]], 0),
    # ... with 4 lines in all: the 8-line tile has 4 lines past the end (they load line 0 and
    # store nothing)
    ((2048, 4), [[4, 4, 4, 4, 4, 2], [4]], 0),
]


def rnd(seed, shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


@pytest.mark.parametrize("case", CASES)
def test_every_tile_shape_of_every_launch(gpu, oracle, case):
    dims, scheds, prec = case[:3]
    fused_axis = case[3] if len(case) > 3 else -1
    x = rnd(31, dims)
    k = rnd(32, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    dt = np.float32 if prec == 1 else np.float64
    tol_ref = 2e-6 if prec == 1 else 1e-12
    tol_tile = 1e-6 if prec == 1 else 1e-14
    pl = gpu.Plan(dims, schedules=scheds, precision=prec, spectrum_precision=prec, fused_axis=fused_axis)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    y0 = dy.numpy()
    assert rel_l2(y0, y_ref) < tol_ref
    default = [nm for nm, _ in pl.profile(dx, dy, 1)]
    tried = 0
    for i, name0 in enumerate(default):
        cands = pl.candidates(i)
        assert name0 in cands, (i, name0, cands)
        for nm in cands:
            pl.pin(i, nm)
            dy.upload(np.zeros(dims, complex))
            pl.convolve(dx, dy)
            assert rel_l2(dy.numpy(), y0) < tol_tile, (i, nm)
            tried += 1
        pl.pin(i, name0)
    assert tried > len(default)
    assert [nm for nm, _ in pl.profile(dx, dy, 1)] == default


def test_pin_rejects_foreign_kernels(gpu):
    pl = gpu.Plan((64, 256), schedules=[[8, 8], [8, 8, 4]], precision=gpu.PREC_F32)
    pl.set_kernel(rnd(33, (64, 256)))
    with pytest.raises(gpu.PacgError, match="not a tile shape"):
        pl.pin(0, "f64_dif_c_8x8x4_h_l8")
    with pytest.raises(gpu.PacgError, match="no pass kernel"):
        pl.candidates(99)


def test_small_arrays_start_on_tiles_that_fill_the_cus(gpu):
    # 2^16 points in lines of 256: 256 lines per pass. The default tiles (16 and 8 lines per
    # workgroup) would give 16 and 32 workgroups; the plan starts on the narrowest compiled ones.
    pl = gpu.Plan((65536,), schedules=[[2] * 16], precision=gpu.PREC_F64)
    desc = pl.describe()
    assert "f64_dif_s_2x2x2x2x2x2x2x2_h_l4" in desc and "f64_dif_c_2x2x2x2x2x2x2x2_h_l4" in desc, desc
    # the same 256-point pass over 65536 rows (65536 lines) keeps its default 8-line tile,
    # over 256 rows it starts on the 4-line one
    small = gpu.Plan((256, 256), schedules=["batch", [2] * 8], precision=gpu.PREC_F64).describe()
    big = gpu.Plan((65536, 256), schedules=["batch", [2] * 8], precision=gpu.PREC_F64).describe()
    assert "f64_dif_c_2x2x2x2x2x2x2x2_h_l4" in small, small
    assert "f64_dif_c_2x2x2x2x2x2x2x2_h_l8" in big, big


STREAM_CASES = [
    # first launch a contiguous A^T pass (axis 2), FP64 and FP32
    ((16, 32, 64), [[4, 4], [8, 4], [8, 8]], 0),
    ((16, 32, 64), [[4, 4], [8, 4], [8, 8]], 1),
    # a strided A^T pass first (2048-point FP64 lines, the product on axis 1)
    ((2048, 1024), [[4, 4, 4, 4, 4, 2], [4, 4, 4, 4, 4]], 0),
    # 1D: the top pass, strided
    ((65536,), [[2] * 16], 0),
    # a one-phase contiguous first pass (16-point lines: 16-byte loads)
    ((64, 16), [[8, 8], [4, 4]], 1),
]


@pytest.mark.parametrize("case", STREAM_CASES)
def test_streamed_input_matches_default_loads(gpu, oracle, case):
    """The first launch's non-temporal loads of x (the autotuner's other load policy) are a
    cache hint: the arithmetic is the same, though the two loads are separate kernel builds whose
    FP32 contractions (which product fuses into an FMA) may differ, so the two convolutions agree
    to rounding (as two tile shapes do); the default one matches the oracle.  The autotuner then
    keeps one of the two policies and the result."""
    dims, scheds, prec = case
    x = rnd(41, dims)
    k = rnd(42, dims)
    dt = np.float32 if prec == 1 else np.float64
    pl = gpu.Plan(dims, schedules=scheds, precision=prec, spectrum_precision=prec)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    y0 = dy.numpy()
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    assert rel_l2(y0, y_ref) < (2e-6 if prec == 1 else 1e-12)
    assert "x:nt" not in pl.describe()
    pl.stream_input(True)
    assert pl.describe().endswith(" x:nt"), pl.describe()
    dy.upload(np.zeros(dims, complex))
    pl.convolve(dx, dy)
    tol_tile = 1e-6 if prec == 1 else 1e-14
    d = rel_l2(dy.numpy(), y0)
    assert d < tol_tile, d
    pl.stream_input(False)
    assert "x:nt" not in pl.describe()
    pl.autotune(dx, dy, 2)
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y0) < tol_tile
