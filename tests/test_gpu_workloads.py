"""The benchmark workloads at their full sizes.

Parity at these sizes: against the oracle where it finishes in seconds, else
against numpy.fft in complex128 (the oracle matches numpy to 1e-15 at the
sizes test_oracle.py covers), plus size-independent properties."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_dropin_1d_65536_r2(gpu, oracle, synth):
    """1D n = 2^16, radix 2, through the reference-named entry point (host arrays)."""
    x, k = synth.workload_inputs("1d_65536")
    n = x.size
    y_ref, ath_ref = oracle.pa_convolution(x, k, [[2] * 16])
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    hr, hi = np.ascontiguousarray(k.real), np.ascontiguousarray(k.imag)
    gpu.permutation_avoiding_convolution(re, im, n, 2, None, None, hr, hi)
    assert rel_l2(re + 1j * im, y_ref) < 1e-12
    assert rel_l2(hr + 1j * hi, ath_ref) < 1e-12  # h <- A^T h, unscaled, bit-reversed order


def test_2d_2048x2048_radix4_fp64(gpu, oracle, synth):
    """2D 2048^2 with radix-4 stages and one radix-2 on top (2048 = 4^5 * 2), FP64."""
    x, k = synth.workload_inputs("2d_2048x2048")
    sched = [4, 4, 4, 4, 4, 2]
    pl = gpu.Plan(x.shape, schedules=[sched, sched], precision=gpu.PREC_F64)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    y = dy.numpy()
    ref = np.fft.ifft2(np.fft.fft2(x) * np.fft.fft2(k))
    assert rel_l2(y, ref) < 1e-12
    # the same schedule against the oracle on a 64-row slab of the problem
    xs, ks = x[:64], k[:64]
    pl2 = gpu.Plan(xs.shape, schedules=[[4, 4, 4], sched], precision=gpu.PREC_F64)
    pl2.set_kernel(ks)
    d2 = gpu.DeviceComplex.from_numpy(xs)
    pl2.convolve(d2)
    y2_ref, _ = oracle.pa_convolution(xs, ks, [[4, 4, 4], sched])
    assert rel_l2(d2.numpy(), y2_ref) < 1e-12


def test_2d_2048x2048_dropin_r2_matches_reference_algorithm(gpu, oracle, synth):
    """The reference can only run 2048^2 with radix 2 (supports_radix_4(11) is false)."""
    x, k = synth.workload_inputs("2d_2048x2048")
    x, k = x[:256], k[:256]  # 256 x 2048 keeps the oracle fast
    y_ref, ath_ref = oracle.pa_convolution(x, k, [[2] * 8, [2] * 11])
    re, im = np.ascontiguousarray(x.real.ravel()), np.ascontiguousarray(x.imag.ravel())
    hr, hi = np.ascontiguousarray(k.real.ravel()), np.ascontiguousarray(k.imag.ravel())
    gpu.permutation_avoiding_convolution_2d(re, im, 256, 2048, 2, None, None, None, None, hr, hi)
    assert rel_l2((re + 1j * im).reshape(x.shape), y_ref) < 1e-12
    assert rel_l2((hr + 1j * hi).reshape(x.shape), ath_ref) < 1e-12


@pytest.mark.parametrize("prec,sprec,tol,scheds,launches,stream_x", [
    (0, 0, 1e-12, [[8, 8, 4]] * 3, 5, False),
    (0, 0, 1e-12, [[4, 8, 8]] * 3, 5, True),
    (1, 1, 2e-6, [[4, 8, 8]] * 3, 5, False),
    (1, 1, 2e-6, [[8, 8, 4]] * 3, 5, False),
    (1, 1, 2e-6, [[8, 8, 4], [4, 8, 8], [8, 8, 4]], 5, False),
    (1, 0, 2e-6, [[4, 8, 8]] * 3, 5, False),
])
def test_3d_256cubed(gpu, synth, prec, sprec, tol, scheds, launches, stream_x):
    """3D 256^3, radix-8 schedules, Gaussian stencil (sum 1): FP64 and FP32 (spectrum stored in
    FP32 or FP64) against numpy, plus the delta-kernel identity at full size (one pass per axis:
    5 launches).  FP64 arrays and FP64-stored spectra exceed the Infinity Cache together (512 and
    384 MiB), so their fused pass streams the spectrum with non-temporal loads; one FP64 case
    also streams x in the first launch (the load policy the autotuner picks there)."""
    x, k = synth.workload_inputs("3d_256cubed")
    pl = gpu.Plan(x.shape, schedules=scheds, precision=prec, spectrum_precision=sprec)
    if stream_x:
        pl.stream_input(True)
    assert pl.launches_per_convolve() == launches, pl.describe()
    pl.set_kernel(k)
    dt = np.float64 if prec == 0 else np.float32
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    ref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(k))
    assert rel_l2(dy.numpy(), ref) < tol
    delta = np.zeros(x.shape, complex)
    delta[0, 0, 0] = 1.0
    pl.set_kernel(delta)
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), x) < tol


def test_3d_256cubed_shift_and_linearity(gpu, synth):
    """Headline plan (FP32, 4-8-8 per axis) at full size against closed forms: a shifted delta
    kernel rolls x by its offset on every axis, and the convolution is linear in x."""
    x, _ = synth.workload_inputs("3d_256cubed")
    pl = gpu.Plan(x.shape, schedules=[[4, 8, 8]] * 3, precision=1, spectrum_precision=1)
    shift = (17, 200, 3)
    k = np.zeros(x.shape, complex)
    k[shift] = 1.0
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), np.roll(x, shift, axis=(0, 1, 2))) < 2e-6
    g = np.random.default_rng(21)
    x2 = g.standard_normal(x.shape) + 1j * g.standard_normal(x.shape)
    pl.convolve(gpu.DeviceComplex.from_numpy(x2, np.float32), dy)
    y2 = dy.numpy()
    pl.convolve(gpu.DeviceComplex.from_numpy(2.0 * x - x2, np.float32), dy)
    y12 = dy.numpy()
    assert rel_l2(y12, 2.0 * np.roll(x, shift, axis=(0, 1, 2)) - y2) < 2e-6


@pytest.mark.parametrize("prec", [1, 0])
def test_1d_2e28_shifted_delta(gpu, prec):
    """2^28 points, the largest size of the reference's published per-element costs
    (per-element-butterfly-cost-plot.py:17-34), at full size: a kernel that is a delta at `shift`
    rolls x by `shift` (a closed form that fixes every index of the convolution), FP32 and FP64
    plans with their default schedules.  Real and imaginary parts are checked apart, so the host
    never holds a complex copy of the 2^28-point arrays."""
    n = 1 << 28
    shift = 123_456_789
    dt = np.float32 if prec else np.float64
    g = np.random.default_rng(280 + prec)
    xr = g.standard_normal(n, dtype=np.float32).astype(dt, copy=False)
    xi = g.standard_normal(n, dtype=np.float32).astype(dt, copy=False)
    pl = gpu.Plan((n,), precision=prec, spectrum_precision=prec)
    k = np.zeros(n)
    k[shift] = 1.0
    pl.set_kernel(k)
    del k
    dx = gpu.DeviceComplex((n,), dt)
    dx.re.upload(xr)
    dx.im.upload(xi)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    num = den = 0.0
    for part, xp in ((dy.re, xr), (dy.im, xi)):
        yp = part.download().astype(np.float64)
        ref = np.roll(xp, shift).astype(np.float64)
        num += float(np.dot(yp - ref, yp - ref))
        den += float(np.dot(ref, ref))
        del yp, ref
    assert (num / den) ** 0.5 < (2e-6 if prec else 1e-12), ((num / den) ** 0.5, pl.describe())
