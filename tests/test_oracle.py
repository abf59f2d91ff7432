"""The oracle pinned before it is trusted (CPU only).

Pins, each chosen so a plausible slip (dropped twiddle, wrong sign, swapped
digit order, transposed operand) fails one of them:
  * golden fixtures written from the reference's own compiled library;
  * the DFT identities A P = F, A^T = P F, conj(A) P = n F^-1 against a
    brute-force O(n^2) DFT (numpy matrix) and numpy.fft;
  * circular-convolution known answers (delta, shifted delta, constant) and an
    O(n^2) brute-force convolution, for pure and mixed-radix schedules;
  * twiddles against mpmath at 40 digits; digit reversal invariants.
"""
from __future__ import annotations

import numpy as np
import pytest
from conftest import golden_cases, load_golden, rel_l2
from oracle import A, CONJ, FFT, FFT2, IFFT, TRANS, UFFT, UIFFT, pure_schedule

MIXED = [[4, 4, 4, 4, 4, 2], [2, 4, 4, 4, 4, 4], [8, 8, 4], [4, 8, 8], [2, 8, 4, 8], [8, 2], [2, 2, 8, 4]]


def brute_dft(x, sign=-1):
    n = x.size
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


def rnd(seed, *shape):
    g = np.random.default_rng(seed)
    return g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)


@pytest.mark.parametrize("name", golden_cases())
def test_oracle_matches_reference_golden(oracle, name):
    g = load_golden(name)
    dims = [int(d) for d in g["dims"]]
    r = int(g["radix"])
    sc = [pure_schedule(n, r) for n in dims]
    x, h = g["x"], g["h"]
    if len(dims) == 1 or r != 8:
        assert rel_l2(oracle.butterflies_nd(x, sc, A), g["A"]) < 1e-14
    assert rel_l2(oracle.butterflies_nd(x, sc, CONJ), g["conjA"]) < 1e-14
    assert rel_l2(oracle.butterflies_nd(x, sc, TRANS), g["AT"]) < 1e-14
    y, ath = oracle.pa_convolution(x, h, sc)
    assert rel_l2(y, g["pa_y"]) < 1e-14 and rel_l2(ath, g["pa_h"]) < 1e-14
    yf, fh = oracle.fft_based_convolution(x, h, sc)
    assert rel_l2(yf, g["fb_y"]) < 1e-14 and rel_l2(fh, g["fb_h"]) < 1e-14
    for nm, w in (("fft", FFT), ("fft2", FFT2), ("ufft", UFFT), ("ifft", IFFT), ("uifft", UIFFT)):
        assert rel_l2(oracle.transform(x, sc, w), g[nm]) < 1e-14, nm
    assert np.array_equal(oracle.rho(dims, sc), g["rho"].astype(np.int64))
    wre, wim = oracle.twiddles_ref_layout(dims[-1], r)
    assert np.array_equal(wre, g["tw_re"]) and np.array_equal(wim, g["tw_im"])


@pytest.mark.parametrize("sched", [[2] * 6, [4] * 3, [8] * 2, [2] * 9, [8] * 3] + MIXED)
def test_dft_identities_brute_force(oracle, sched):
    n = int(np.prod(sched))
    x = rnd(n, n)
    F = brute_dft(x)
    rho = oracle.rho([n], [sched])
    # A^T x = P F x: the DFT in digit-reversed order
    assert np.abs(oracle.butterflies(x, sched, TRANS) - F[rho]).max() < 1e-11 * n
    # A (P x) = F x
    assert np.abs(oracle.butterflies(x[rho], sched, A) - F).max() < 1e-11 * n
    # conj(A) (P x) = n F^-1 x
    assert np.abs(oracle.butterflies(x[rho], sched, CONJ) - brute_dft(x, +1)).max() < 1e-11 * n


@pytest.mark.parametrize("sched", [[2] * 6, [4] * 3, [8] * 2] + MIXED)
def test_digit_reversal_invariants(oracle, sched):
    n = int(np.prod(sched))
    rho = oracle.rho([n], [sched])
    assert rho[0] == 0
    assert np.array_equal(np.sort(rho), np.arange(n))  # a bijection
    if len(set(sched)) == 1:
        assert np.array_equal(rho[rho], np.arange(n))  # pure radix: an involution
        r = sched[0]
        t = len(sched)
        for i in (1, 2, n // 2, n - 1):
            assert rho[i] == oracle.reverse_index_r(i, t, r)


def test_reverse_index_known_values(oracle):
    # base-r digit reversal worked by hand: 6 = 110b -> 011b = 3; (1,2)_4 -> (2,1)_4; (1,2)_8 -> (2,1)_8
    assert oracle.reverse_index_r(6, 3, 2) == 3
    assert oracle.reverse_index_r(1 + 2 * 4, 2, 4) == 2 + 1 * 4
    assert oracle.reverse_index_r(1 + 2 * 8, 2, 8) == 2 + 1 * 8
    assert oracle.reverse_index_r(1, 5, 2) == 16


def test_twiddles_against_mpmath(oracle):
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    for r, n in ((2, 256), (4, 256), (8, 512)):
        wre, wim = oracle.twiddles_ref_layout(n, r)
        off = 0
        ell = 1
        while ell * r <= n:
            k = ell * r
            for m in range(r - 1):
                for j in range(0, ell, max(1, ell // 16)):
                    ang = -2 * mp.pi * (m + 1) * j / k
                    assert abs(wre[off + m * ell + j] - float(mp.cos(ang))) <= 1.5e-15
                    assert abs(wim[off + m * ell + j] - float(mp.sin(ang))) <= 1.5e-15
            off += (r - 1) * ell
            ell *= r
        assert off == n - 1


def _brute_conv(x, h):
    n = x.size
    y = np.zeros(n, complex)
    for m in range(n):
        y[m] = sum(x[j] * h[(m - j) % n] for j in range(n))
    return y


@pytest.mark.parametrize("sched", [[2] * 6, [4] * 3, [8] * 2, [8, 8, 4], [4, 2]])
def test_pa_convolution_known_answers(oracle, sched):
    n = int(np.prod(sched))
    x = rnd(7, n)
    delta = np.zeros(n, complex)
    delta[0] = 1
    y, _ = oracle.pa_convolution(x, delta, [sched])
    assert np.abs(y - x).max() < 1e-13
    for s in (1, 3, n - 1):
        e = np.zeros(n, complex)
        e[s] = 1
        y, _ = oracle.pa_convolution(x, e, [sched])
        assert np.abs(y - np.roll(x, s)).max() < 1e-13
    y, _ = oracle.pa_convolution(x, np.ones(n, complex), [sched])
    assert np.abs(y - x.sum()).max() < 1e-12 * n
    h = rnd(8, n)
    y, ath = oracle.pa_convolution(x, h, [sched])
    assert np.abs(y - _brute_conv(x, h)).max() < 1e-12 * n
    # the side effect on h: unscaled A^T h
    assert np.abs(ath - oracle.butterflies(h, sched, TRANS)).max() == 0


@pytest.mark.parametrize("dims,scheds", [
    ((16, 64), [[4, 4], [2, 4, 8]]),
    ((8, 16, 32), [[8], [4, 4], [2, 4, 4]]),
    ((4, 2, 8), [[4], [2], [8]]),
    ((1, 16), [[], [4, 4]]),
])
def test_nd_against_numpy(oracle, dims, scheds):
    x = rnd(3, *dims)
    h = rnd(4, *dims)
    y, _ = oracle.pa_convolution(x, h, scheds)
    ref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(h))
    assert rel_l2(y, ref) < 1e-14
    yf, _ = oracle.fft_based_convolution(x, h, scheds)
    assert rel_l2(yf, ref) < 1e-14
    assert rel_l2(oracle.transform(x, scheds, FFT), np.fft.fftn(x)) < 1e-14
    assert rel_l2(oracle.transform(x, scheds, FFT2), np.fft.fftn(x)) < 1e-14
    assert rel_l2(oracle.transform(x, scheds, IFFT), np.fft.ifftn(x)) < 1e-14
    rho = oracle.rho(list(dims), scheds)
    assert rel_l2(oracle.butterflies_nd(x, scheds, TRANS).ravel(), np.fft.fftn(x).ravel()[rho]) < 1e-14


def test_max_error_semantics(oracle):
    y = np.array([1 + 1j, 2 + 0j, 0 + 3j])
    ref = np.array([1 + 1j, 2 + 4j, 3 + 3j])
    assert oracle.max_error(y, ref) == 4.0


def test_oracle_planted_mistake_is_caught(oracle):
    """A conjugated twiddle in the oracle's output breaks the DFT identity."""
    n = 64
    x = rnd(5, n)
    good = oracle.butterflies(x, [4, 4, 4], TRANS)
    bad = oracle.butterflies(x, [4, 4, 4], CONJ)  # plausible slip: wrong operator
    F = np.fft.fft(x)
    rho = oracle.rho([n], [[4, 4, 4]])
    assert np.abs(good - F[rho]).max() < 1e-12
    assert np.abs(bad - F[rho]).max() > 1.0
