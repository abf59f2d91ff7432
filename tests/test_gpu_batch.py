"""Batched uneven-length convolutions (one persistent launch) and batch-axis
plans, against the oracle pair by pair.  FP32 bar: rel-L2 <= 2e-6 per pair."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL32 = 2e-6


def rnd(seed, n):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)


@pytest.fixture(scope="module")
def small_batch(gpu):
    ms = np.array([14, 8, 9, 13, 10, 11, 12, 8, 14, 12, 9, 10, 11, 13, 8, 8, 8, 10, 12, 14], dtype=np.int32)
    xs = [rnd(100 + i, 1 << int(m)) for i, m in enumerate(ms)]
    ks = [rnd(200 + i, 1 << int(m)) for i, m in enumerate(ms)]
    b = gpu.Batch(ms)
    dx = gpu.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = gpu.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    return b, xs, ks, dx, dk, dy


def check(b, xs, ks, dy, oracle, scheds):
    ys = b.unpack(dy.numpy())
    for i, (x, k, y) in enumerate(zip(xs, ks, ys)):
        m = int(b.log2n[i])
        ref, _ = oracle.pa_convolution(x, k, [scheds[m]])
        assert rel_l2(y, ref) < TOL32, (i, m, rel_l2(y, ref))


@pytest.mark.parametrize("log2n", list(range(8, 15)))
def test_every_variant_of_every_length(gpu, oracle, small_batch, log2n):
    b, xs, ks, dx, dk, dy = small_batch
    for v in range(gpu.Batch.num_variants(log2n)):
        b.set_variant(log2n, v)
        dy.upload(np.zeros(b.total, complex))
        b.convolve(dx, dk, dy)
        b.sync()
        scheds = {m: gpu.Batch.schedule(m, b.variant(m)) for m in range(8, 15)}
        assert int(np.prod(scheds[log2n])) == 1 << log2n
        check(b, xs, ks, dy, oracle, scheds)
    b.set_variant(log2n, 0)


def test_autotune_then_graph_replay(gpu, oracle, small_batch):
    b, xs, ks, dx, dk, dy = small_batch
    b.autotune(dx, dk, dy, 2)
    scheds = {m: gpu.Batch.schedule(m, b.variant(m)) for m in range(8, 15)}
    b.capture(dx, dk, dy)
    for it in range(3):
        dy.upload(np.zeros(b.total, complex))
        b.replay()
        b.sync()
        check(b, xs, ks, dy, oracle, scheds)
        b.convolve(dx, dk, dy)  # eager launch between replays
        b.sync()


def test_identity_and_shift_kernels(gpu):
    ms = np.array([8, 11, 14], dtype=np.int32)
    b = gpu.Batch(ms)
    xs = [rnd(7 + i, 1 << int(m)) for i, m in enumerate(ms)]
    ks = []
    for i, m in enumerate(ms):
        k = np.zeros(1 << int(m), complex)
        k[i] = 1.0  # delta at i: y = roll(x, i)
        ks.append(k)
    dx = gpu.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = gpu.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    b.convolve(dx, dk, dy)
    b.sync()
    for i, (x, y) in enumerate(zip(xs, b.unpack(dy.numpy()))):
        assert np.abs(y - np.roll(x, i)).max() < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, TOL32)])
def test_batch_axis_plan_convolves_rows(gpu, oracle, prec, tol):
    rows, n = 5, 512
    sched = [8, 8, 8]
    x = np.stack([rnd(300 + r, n) for r in range(rows)])
    k = np.stack([rnd(400 + r, n) for r in range(rows)])
    pl = gpu.Plan((rows, n), schedules=["batch", sched], precision=prec)
    pl.set_kernel(k)
    dt = np.float64 if prec == 0 else np.float32
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    pl.convolve(dx)
    y = dx.numpy()
    for r in range(rows):
        ref, _ = oracle.pa_convolution(x[r], k[r], [sched])
        assert rel_l2(y[r], ref) < tol
    for op in (gpu.OP_FFT, gpu.OP_IFFT):
        dx = gpu.DeviceComplex.from_numpy(x, dt)
        pl.execute(op, dx)
        ref = np.fft.fft(x, axis=1) if op == gpu.OP_FFT else np.fft.ifft(x, axis=1)
        assert rel_l2(dx.numpy(), ref) < tol


def test_ragged_items_of_8192_points(gpu, oracle):
    """Items of 8192-point convolutions hold two; an odd count leaves a last item with one."""
    ms = np.array([13, 12, 13, 14, 13, 8], dtype=np.int32)
    xs = [rnd(300 + i, 1 << int(m)) for i, m in enumerate(ms)]
    ks = [rnd(400 + i, 1 << int(m)) for i, m in enumerate(ms)]
    b = gpu.Batch(ms)
    dx = gpu.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = gpu.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    b.convolve(dx, dk, dy)
    b.sync()
    check(b, xs, ks, dy, oracle, {m: gpu.Batch.schedule(m, b.variant(m)) for m in range(8, 15)})


def test_both_dealings_give_identical_results(gpu, synth):
    """Items dealt longest processing time first or in slices of the shape-ordered list: each
    convolution is the same item computed by the same kernel code, only on another workgroup, so
    the outputs are bit-identical (500 pairs of the benchmark's length distribution, every
    variant set the autotuner may choose here); the autotuner keeps one of the two dealings."""
    xs, ks = synth.batched_inputs(4, 500)
    ms = np.array([int(np.log2(len(x))) for x in xs], dtype=np.int32)
    b = gpu.Batch(ms)
    dx = gpu.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = gpu.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    assert b.dealing() == gpu.Batch.DEAL_LPT
    for combo in ([0] * 7, [1, 1, 1, 1, 1, 2, 2]):
        for m, v in zip(range(8, 15), combo):
            b.set_variant(m, v)
        out = []
        for dealing in (gpu.Batch.DEAL_LPT, gpu.Batch.DEAL_SLICES):
            b.set_dealing(dealing)
            assert b.dealing() == dealing
            dy.upload(np.zeros(int(b.total), complex))
            b.convolve(dx, dk, dy)
            b.sync()
            out.append(dy.numpy())
        assert np.array_equal(out[0], out[1]), combo
    ys = b.unpack(out[0])
    for x, k, y in zip(xs[:40], ks[:40], ys[:40]):
        assert rel_l2(y, np.fft.ifft(np.fft.fft(x) * np.fft.fft(k))) < TOL32
    b.autotune(dx, dk, dy, 2)
    assert b.dealing() in (gpu.Batch.DEAL_LPT, gpu.Batch.DEAL_SLICES)
    with pytest.raises(gpu.PacgError):
        b.set_dealing(2)


def test_full_size_batch_of_4096_pairs(gpu, synth):
    """The benchmark batch itself (4096 pairs, lengths 2^U{8..14}, seed 4), autotuned and replayed
    from its graph: every pair of every length against numpy.fft in complex128, per pair and over the
    whole batch."""
    xs, ks = synth.batched_inputs(4, 4096)
    ms = np.array([int(np.log2(len(x))) for x in xs], dtype=np.int32)
    b = gpu.Batch(ms)
    dx = gpu.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = gpu.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    b.autotune(dx, dk, dy, 2)
    b.capture(dx, dk, dy)
    for _ in range(3):
        dy.upload(np.zeros(int(b.total), complex))
        b.replay()
        b.sync()
        ys = b.unpack(dy.numpy())
        num = den = 0.0
        worst = 0.0
        for x, k, y in zip(xs, ks, ys):
            ref = np.fft.ifft(np.fft.fft(x) * np.fft.fft(k))
            d = y - ref
            num += float(np.vdot(d, d).real)
            den += float(np.vdot(ref, ref).real)
            worst = max(worst, rel_l2(y, ref))
        assert worst < TOL32, worst
        assert (num / den) ** 0.5 < 1e-6
