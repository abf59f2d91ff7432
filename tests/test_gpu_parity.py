"""GPU parity: the HIP pass kernels against the oracle on the same inputs.

FP64 bar: relative L2 <= 1e-12 against the oracle's FP64 convolution; the digit
reversal / spectrum order is compared exactly (integer bookkeeping).
FP32 bar: relative L2 <= 2e-6 against FP64 (FP32 butterflies, FP64-generated twiddles).
"""
from __future__ import annotations

import numpy as np
import pytest
from conftest import golden_cases, load_golden, rel_l2
from oracle import A, CONJ, FFT, FFT2, IFFT, TRANS, UFFT, UIFFT, pure_schedule

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-6

CASES = [
    # dims, schedules (stage 0 first)
    ((64,), [[2] * 6]),
    ((256,), [[4] * 4]),
    ((512,), [[8] * 3]),
    ((4096,), [[8] * 4]),
    ((1 << 16,), [[2] * 16]),
    ((1 << 16,), [[4] * 8]),
    ((2048,), [[4] * 5 + [2]]),
    ((256,), [[8, 8, 4]]),
    ((256,), [[4, 8, 8]]),
    ((32,), [[2, 2, 8]]),
    ((8,), [[8]]),
    ((2,), [[2]]),
    ((1,), [[]]),
    ((16, 64), [[4, 4], [4, 4, 4]]),
    ((64, 4), [[4] * 3, [4]]),
    ((128, 256), [[2] * 7, [8, 8, 4]]),
    ((8, 16, 32), [[8], [4, 4], [2, 4, 4]]),
    ((16, 16, 16), [[4, 4], [4, 4], [4, 4]]),
    ((32, 32, 64), [[8, 4], [4, 8], [8, 8]]),
    ((1, 16, 8), [[], [4, 4], [8]]),
    ((4, 1, 8), [[4], [], [8]]),
    # FP64: axis 0 in one strided 2048-point pass (1024 lines)
    ((2048, 1024), [[4] * 5 + [2], [4] * 5]),
]


def rnd(seed, shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


@pytest.mark.parametrize("dims,scheds", CASES)
def test_convolve_fp64(gpu, oracle, dims, scheds):
    x = rnd(1, dims)
    k = rnd(2, dims)
    y_ref, s_ref = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds)
    pl.set_kernel(k)
    # spectrum = unscaled A^T k in the schedule's digit-reversed order
    assert rel_l2(pl.spectrum(), s_ref) < TOL64
    assert np.array_equal(pl.rho(), oracle.rho(list(dims), scheds).astype(np.int32))
    dx = gpu.DeviceComplex.from_numpy(x)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < TOL64
    pl.convolve(dx)  # in place
    assert rel_l2(dx.numpy(), y_ref) < TOL64


@pytest.mark.parametrize("dims,scheds", CASES)
def test_operators_fp64(gpu, oracle, dims, scheds):
    x = rnd(3, dims)
    pl = gpu.Plan(dims, schedules=scheds)
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
        pl.execute(op, dx)  # in place
        assert rel_l2(dx.numpy(), ref()) < TOL64, op


@pytest.mark.parametrize("dims,scheds", CASES)
def test_pair_and_fft_based_fp64(gpu, oracle, dims, scheds):
    x = rnd(5, dims)
    h = rnd(6, dims)
    pl = gpu.Plan(dims, schedules=scheds)
    y_ref, ath_ref = oracle.pa_convolution(x, h, scheds)
    dx, dh = gpu.DeviceComplex.from_numpy(x), gpu.DeviceComplex.from_numpy(h)
    pl.convolve_pair(dx, dh)
    assert rel_l2(dx.numpy(), y_ref) < TOL64
    assert rel_l2(dh.numpy(), ath_ref) < TOL64
    yf_ref, fh_ref = oracle.fft_based_convolution(x, h, scheds)
    dx, dh = gpu.DeviceComplex.from_numpy(x), gpu.DeviceComplex.from_numpy(h)
    pl.convolve_fft_based(dx, dh)
    assert rel_l2(dx.numpy(), yf_ref) < TOL64
    assert rel_l2(dh.numpy(), fh_ref) < TOL64


@pytest.mark.parametrize("dims,scheds", CASES)
@pytest.mark.parametrize("prec", [0, 1])
def test_permutation_op(gpu, oracle, dims, scheds, prec):
    """PACG_OP_PERM: x[p] <- x[rho(p)] with the oracle's digit reversal, exactly (data movement only),
    into another array and in place (through the plan's scratch)."""
    x = rnd(11, dims)
    dt = np.float64 if prec == 0 else np.float32
    rho = oracle.rho(list(dims), scheds)
    want = x.astype(np.complex128 if prec == 0 else np.complex64).ravel()[rho].reshape(dims)
    pl = gpu.Plan(dims, schedules=scheds, precision=prec)
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.execute(gpu.OP_PERM, dx, dy)
    assert np.array_equal(dy.numpy(), want)
    pl.execute(gpu.OP_PERM, dx)  # in place
    assert np.array_equal(dx.numpy(), want)
    pl.close()


def test_launch_floor(gpu):
    d_us, g_us = gpu.launch_floor(200)
    assert 0.0 < d_us < 1000.0 and 0.0 < g_us < 1000.0


@pytest.mark.parametrize("dims,scheds", [c for c in CASES if np.prod(c[0]) >= 8])
@pytest.mark.parametrize("sprec", [0, 1])
def test_convolve_fp32(gpu, oracle, dims, scheds, sprec):
    x = rnd(7, dims)
    k = rnd(8, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=gpu.PREC_F32, spectrum_precision=sprec)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x, np.float32)
    pl.convolve(dx)
    assert rel_l2(dx.numpy(), y_ref) < TOL32


@pytest.mark.parametrize("name", golden_cases())
def test_dropin_against_reference_golden(gpu, name):
    g = load_golden(name)
    dims = [int(d) for d in g["dims"]]
    r = int(g["radix"])
    nd = len(dims)
    sfx = {1: "", 2: "_2d", 3: "_3d"}[nd]
    tw = [None] * (2 * nd)

    def split(z):
        return np.ascontiguousarray(z.real.ravel()), np.ascontiguousarray(z.imag.ravel())

    def join(a, b):
        return (a + 1j * b).reshape(dims)

    rho = g["rho"]
    # permutation-avoiding convolution: x <- y, h <- A^T h
    xr, xi = split(g["x"])
    hr, hi = split(g["h"])
    getattr(gpu, "permutation_avoiding_convolution" + sfx)(xr, xi, *dims, r, *tw, hr, hi)
    assert rel_l2(join(xr, xi), g["pa_y"]) < TOL64
    assert rel_l2(join(hr, hi), g["pa_h"]) < TOL64
    # FFT-based convolution: x <- y, h <- F h
    xr, xi = split(g["x"])
    hr, hi = split(g["h"])
    getattr(gpu, "fft_based_convolution" + sfx)(xr, xi, *dims, r, *tw, rho, hr, hi)
    assert rel_l2(join(xr, xi), g["fb_y"]) < TOL64
    assert rel_l2(join(hr, hi), g["fb_h"]) < TOL64
    for nm in ("fft", "fft2", "ifft"):
        xr, xi = split(g["x"])
        getattr(gpu, nm + sfx)(xr, xi, *dims, r, *tw, rho)
        assert rel_l2(join(xr, xi), g[nm]) < TOL64, nm
    for nm in ("ufft", "uifft"):
        xr, xi = split(g["x"])
        getattr(gpu, nm + sfx)(xr, xi, *dims, r, *tw)
        assert rel_l2(join(xr, xi), g[nm]) < TOL64, nm
    if nd == 1 or r in (2, 4):
        for kind, key in (("", "A"), ("conjugate_", "conjA"), ("transposed_", "AT")):
            if nd > 1 and r == 8:
                continue
            xr, xi = split(g["x"])
            getattr(gpu, "apply_%sbutterflies_r%d%s" % (kind, r, sfx))(xr, xi, *dims, *tw)
            assert rel_l2(join(xr, xi), g[key]) < TOL64, kind


def _bitrev(t):
    i = np.arange(1 << t)
    out = np.zeros_like(i)
    for b in range(t):
        out |= ((i >> b) & 1) << (t - 1 - b)
    return out


@pytest.mark.parametrize("t", [12, 17, 18, 20, 22])
def test_dropin_copy_paths(gpu, t):
    """Host arrays of every copy path of the drop-in entry points: up to 2^17 doubles through the
    pinned staging buffer, above that through the pinned ring in 8 MiB chunks (2^18: one partial
    chunk, 2^20: one full chunk, 2^22: four chunks, so ring slots are reused within one array)."""
    n = 1 << t
    g = np.random.default_rng(40 + t)
    x = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
    h = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
    Fh = np.fft.fft(h)
    y_ref = np.fft.ifft(np.fft.fft(x) * Fh)
    rho = _bitrev(t).astype(np.int32)
    for _ in range(2):  # the second call reuses the cached plan, staging buffer and ring
        xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
        hr, hi = np.ascontiguousarray(h.real), np.ascontiguousarray(h.imag)
        gpu.permutation_avoiding_convolution(xr, xi, n, 2, None, None, hr, hi)
        assert rel_l2(xr + 1j * xi, y_ref) < TOL64
        assert rel_l2(hr + 1j * hi, Fh[rho]) < TOL64  # h <- A^T h = P F h
        xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
        hr, hi = np.ascontiguousarray(h.real), np.ascontiguousarray(h.imag)
        gpu.fft_based_convolution(xr, xi, n, 2, None, None, rho, hr, hi)
        assert rel_l2(xr + 1j * xi, y_ref) < TOL64
        assert rel_l2(hr + 1j * hi, Fh) < TOL64  # h <- F h


@pytest.mark.parametrize("dims,r", [((1 << 24,), 4), ((1 << 21,), 8), ((1024, 4, 4), 4), ((4, 4, 2048), 2)])
def test_dropin_large_and_anisotropic(gpu, dims, r):
    """Drop-in convolution at large 1D sizes (2^24 = 4^12 points, 2^21 = 8^7) and on anisotropic 3D
    shapes like those of the reference's *_anisotropic driver programs (two axes of 4 points),
    against numpy.fft."""
    n = int(np.prod(dims))
    g = np.random.default_rng(60 + len(dims) + r)
    x = (g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)).reshape(dims)
    h = (g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)).reshape(dims)
    y_ref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(h))
    xr, xi = np.ascontiguousarray(x.real.ravel()), np.ascontiguousarray(x.imag.ravel())
    hr, hi = np.ascontiguousarray(h.real.ravel()), np.ascontiguousarray(h.imag.ravel())
    sfx = {1: "", 2: "_2d", 3: "_3d"}[len(dims)]
    getattr(gpu, "permutation_avoiding_convolution" + sfx)(xr, xi, *dims, r, *([None] * (2 * len(dims))), hr, hi)
    assert rel_l2((xr + 1j * xi).reshape(dims), y_ref) < TOL64


def test_dropin_rejects_bad_input(gpu):
    x = np.zeros(24)
    y = np.zeros(24)
    with pytest.raises(gpu.PacgError):
        gpu.apply_transposed_butterflies_r4(x, y, 24, None, None)
    with pytest.raises(gpu.PacgError):
        gpu.permutation_avoiding_convolution(x[:16], y[:16], 16, 3, None, None, x[:16].copy(), y[:16].copy())
    assert np.all(x == 0) and np.all(y == 0)


def test_graph_replay_matches_eager(gpu, oracle):
    dims = (64, 64)
    scheds = [[4, 4, 4], [8, 8]]
    x = rnd(9, dims)
    k = rnd(10, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x)
    dy = pl.alloc()
    pl.capture(dx, dy)
    for it in range(3):
        dy.upload(np.zeros(dims, complex))
        pl.replay()
        pl.sync()
        assert rel_l2(dy.numpy(), y_ref) < TOL64, it
        pl.convolve(dx, dy)  # eager work between replays
        pl.sync()


@pytest.mark.parametrize("prec,tol", [(0, TOL64), (1, TOL32)])
def test_autotuned_plan_keeps_parity(gpu, oracle, prec, tol):
    dims = (64, 32, 256)
    scheds = [[8, 8], [4, 8], [8, 8, 4]]
    x = rnd(11, dims)
    k = rnd(12, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=prec)
    pl.set_kernel(k)
    dt = np.float64 if prec == 0 else np.float32
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.autotune(dx, dy, 2)
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < tol
    names = [nm for nm, _ in pl.profile(dx, dy, 2)]
    assert len(names) == pl.launches_per_convolve() == 5


@pytest.mark.parametrize("prec,tol", [(0, TOL64), (1, TOL32)])
def test_unaligned_buffers_take_the_scalar_path(gpu, oracle, prec, tol):
    dims = (16, 256)
    scheds = [[4, 4], [8, 8, 4]]
    x = rnd(13, dims)
    k = rnd(14, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=prec)
    pl.set_kernel(k)
    dt = np.float64 if prec == 0 else np.float32
    dx = gpu.DeviceComplex.shifted(dims, dt, 1)
    dx.upload(x)
    dy = gpu.DeviceComplex.shifted(dims, dt, 3)
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < tol
    for op, kind in ((gpu.OP_TRANS, TRANS), (gpu.OP_CONJ, CONJ), (gpu.OP_FWD, A)):
        pl.execute(op, dx, dy)
        assert rel_l2(dy.numpy(), oracle.butterflies_nd(x, scheds, kind)) < tol


def test_dropin_calls_from_several_threads(gpu):
    """Drop-in calls from four host threads at once (ctypes releases the GIL): the calls serialise
    on the library's mutex and share the pinned ring and copy pool, and every result matches."""
    import threading

    sizes = [1 << 14, 1 << 19, 1 << 20, 1 << 21]
    g = np.random.default_rng(77)
    cases = []
    for n in sizes:
        x = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
        h = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
        cases.append((n, x, h, np.fft.ifft(np.fft.fft(x) * np.fft.fft(h))))
    errors = []

    def work(case, reps):
        n, x, h, y_ref = case
        try:
            for _ in range(reps):
                xr, xi = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
                hr, hi = np.ascontiguousarray(h.real), np.ascontiguousarray(h.imag)
                gpu.permutation_avoiding_convolution(xr, xi, n, 2, None, None, hr, hi)
                e = rel_l2(xr + 1j * xi, y_ref)
                if not e < TOL64:
                    errors.append((n, e))
        except Exception as exc:  # surfaced below
            errors.append((n, repr(exc)))

    threads = [threading.Thread(target=work, args=(c, 3)) for c in cases]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors


def test_strided_2048_point_pass_with_external_twiddle(gpu, oracle):
    """A strided 2048-point FP64 pass that is not the bottom pass of its axis (65536 points = 8 * 4 *
    2048, lower span 32) multiplies by the pass-external twiddle w_K^{j' k}: convolution and the three
    operators against the oracle."""
    dims = (65536, 32)
    scheds = [[8] + [4] * 6 + [2], [4, 8]]
    pl = gpu.Plan(dims, schedules=scheds)
    assert "P=2048 L=32 f64_dif_s_4x4x4x4x4x2" in pl.describe(), pl.describe()
    x = rnd(51, dims)
    k = rnd(52, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl.set_kernel(k)
    dx = gpu.DeviceComplex.from_numpy(x)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < TOL64
    for op, kind in ((gpu.OP_TRANS, TRANS), (gpu.OP_CONJ, CONJ), (gpu.OP_FWD, A)):
        pl.execute(op, dx, dy)
        assert rel_l2(dy.numpy(), oracle.butterflies_nd(x, scheds, kind)) < TOL64, op
