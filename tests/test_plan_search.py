"""Radix-schedule search (pacg_plan_search): its candidate lists on the CPU, and on the GPU that the
schedules it returns make plans that keep parity with the oracle."""
from __future__ import annotations

import numpy as np
import pytest
from conftest import rel_l2


def test_candidates_cover_every_mix(pacg):
    c256 = pacg.schedule_candidates(256)
    for want in ([8, 8, 4], [4, 8, 8], [4, 4, 4, 4], [8, 4, 4, 2], [2, 4, 4, 8]):
        assert want in c256, (want, c256)
    assert len(c256) == 5
    c2048 = pacg.schedule_candidates(2048)
    assert [4, 4, 4, 4, 4, 2] in c2048 and [8, 8, 8, 4] in c2048 and [4, 8, 8, 8] in c2048
    for n in (2, 4, 8, 64, 256, 2048, 1 << 16):
        cands = pacg.schedule_candidates(n)
        assert cands, n
        assert len({tuple(c) for c in cands}) == len(cands), n  # no duplicates
        for c in cands:
            assert int(np.prod(c)) == n and set(c) <= {2, 4, 8} and c.count(2) <= 1, (n, c)
    assert pacg.schedule_candidates(2) == [[2]]
    assert pacg.schedule_candidates(8) == [[8], [4, 2], [2, 4]]


def test_candidates_reject_bad_lengths(pacg):
    for n in (0, 3, 96):
        with pytest.raises(pacg.PacgError):
            pacg.schedule_candidates(n)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,prec", [((64, 512), 0), ((16, 64, 256), 1), ((4096,), 0)])
def test_searched_schedules_keep_parity(gpu, oracle, dims, prec):
    scheds, changed, fused = gpu.search_schedules(dims, precision=prec, spectrum_precision=prec, reps=3)
    assert 0 <= changed <= len(dims)
    assert -1 <= fused < len(dims)
    for n, sc in zip(dims, scheds):
        assert sc is None or int(np.prod(sc)) == n, (n, sc)
    scheds = [sc if sc is not None else gpu.schedule_candidates(n)[0] for n, sc in zip(dims, scheds)]
    g = np.random.default_rng(41)
    x = g.uniform(-1, 1, dims) + 1j * g.uniform(-1, 1, dims)
    k = g.uniform(-1, 1, dims) + 1j * g.uniform(-1, 1, dims)
    y_ref, _ = oracle.pa_convolution(x, k, scheds)
    pl = gpu.Plan(dims, schedules=scheds, precision=prec, spectrum_precision=prec, fused_axis=fused)
    pl.set_kernel(k)
    dt = np.float32 if prec else np.float64
    dx = gpu.DeviceComplex.from_numpy(x, dt)
    dy = pl.alloc()
    pl.convolve(dx, dy)
    assert rel_l2(dy.numpy(), y_ref) < (2e-6 if prec else 1e-12)
