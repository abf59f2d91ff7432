"""The HBM probe (pacg_hbm_probe, the device counterpart of the reference's
memory_bw_test.c:27-67): each mode touches exactly the bytes asked for, and the
timed rate is finite and positive.  The rates themselves are reported by bench.py."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nbytes", [16, 4096 + 48, 64 << 20])
def test_probe_modes_touch_exactly_the_asked_bytes(gpu, nbytes):
    n = (nbytes // 4) + 64  # floats, with a tail the probe must leave alone
    g = np.random.default_rng(7)
    src_h = g.standard_normal(n).astype(np.float32)
    src_h[nbytes // 4:] = 0.0
    src = gpu.DeviceArray(n, np.float32)
    src.upload(src_h)
    dst = gpu.DeviceArray(n, np.float32)
    dst.upload(np.full(n, 7.0, np.float32))
    ms = gpu.hbm_probe(2, dst, src, nbytes, reps=3)  # copy
    out = dst.download()
    k = nbytes // 4
    assert np.array_equal(out[:k], src_h[:k])
    assert np.all(out[k:] == 7.0)
    assert np.isfinite(ms) and ms > 0.0
    ms = gpu.hbm_probe(1, dst, None, nbytes, reps=3)  # write (zeros)
    out = dst.download()
    assert np.all(out[:k] == 0.0) and np.all(out[k:] == 7.0)
    assert ms > 0.0
    zero = gpu.DeviceArray(n, np.float32)
    zero.upload(np.zeros(n, np.float32))
    dst.upload(np.full(n, 7.0, np.float32))
    ms = gpu.hbm_probe(0, dst, zero, nbytes, reps=3)  # read: dst untouched
    assert np.all(dst.download() == 7.0)
    assert ms > 0.0


def test_probe_copy_rate_is_plausible(gpu):
    nbytes = 256 << 20
    src = gpu.DeviceArray(nbytes // 4, np.float32)
    dst = gpu.DeviceArray(nbytes // 4, np.float32)
    src.upload(np.zeros(nbytes // 4, np.float32))
    ms = gpu.hbm_probe(2, dst, src, nbytes, reps=5)
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    assert 500.0 < gbs < 20000.0, gbs
