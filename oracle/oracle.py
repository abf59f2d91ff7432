"""ctypes front-end for the test-only checkers.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

* ``Oracle``  wraps ``liboracle.so`` (pacg_oracle.c, the clean-room restatement).
* ``RefLib``  wraps ``_ref/libref.so`` (the reference's own sources compiled by
  oracle/Makefile), exposing the reference's fft.h entry points with numpy
  arguments.  Absent when the reference was not available at build time.

Schedules are lists of radices, ``sched[q]`` = radix of stage q, stage 0 the
unit-stride one; a pure radix-r schedule is ``[r] * t``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAXT = 64
A, CONJ, TRANS = 0, 1, 2
FFT, FFT2, UFFT, IFFT, UIFFT = 0, 1, 2, 3, 4

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_i64p = C.POINTER(C.c_int64)


def build():
    """Compile liboracle.so (and _ref/libref.so when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _dims_scheds(dims, scheds):
    nd = len(dims)
    d = (C.c_long * 3)(*[int(x) for x in dims] + [0] * (3 - nd))
    s = (C.c_int * (3 * MAXT))()
    ts = (C.c_int * 3)()
    for a, sc in enumerate(scheds):
        ts[a] = len(sc)
        for q, r in enumerate(sc):
            s[a * MAXT + q] = int(r)
    return nd, d, s, ts


def pure_schedule(n: int, r: int):
    t = 0
    m = 1
    while m < n:
        m *= r
        t += 1
    if m != n:
        raise ValueError(f"{n} is not a power of {r}")
    return [r] * t


class Oracle:
    def __init__(self, path: str | None = None):
        path = path or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        L.or_butterflies.argtypes = [_dp, _dp, C.c_long, _ip, C.c_int, C.c_int]
        L.or_butterflies_nd.argtypes = [_dp, _dp, C.c_int, _lp, _ip, _ip, C.c_int]
        L.or_reverse_index_r.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_reverse_index_r.restype = C.c_int
        L.or_rev_index_mixed.argtypes = [C.c_long, C.c_long, _ip, C.c_int]
        L.or_rev_index_mixed.restype = C.c_long
        L.or_perm_nd.argtypes = [C.c_int, _lp, _ip, _ip, _i64p]
        L.or_gather.argtypes = [_dp, _dp, C.c_long, _i64p]
        L.or_scatter.argtypes = [_dp, _dp, C.c_long, _i64p]
        L.or_twiddles_ref_layout.argtypes = [C.c_long, C.c_int, _dp, _dp]
        L.or_twiddles_ref_layout.restype = C.c_long
        L.or_transform_nd.argtypes = [_dp, _dp, C.c_int, _lp, _ip, _ip, C.c_int]
        L.or_pa_convolution_nd.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _lp, _ip, _ip]
        L.or_fft_based_convolution_nd.argtypes = [_dp, _dp, _dp, _dp, C.c_int, _lp, _ip, _ip]
        L.or_max_error.argtypes = [_dp, _dp, _dp, C.c_long]
        L.or_max_error.restype = C.c_double

    # -- butterflies ---------------------------------------------------------
    def butterflies(self, x: np.ndarray, sched, kind: int) -> np.ndarray:
        """A (kind 0), conj(A) (1) or A^T (2) of a complex vector; returns a copy."""
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        s = (C.c_int * max(1, len(sched)))(*sched)
        self.L.or_butterflies(_d(re), _d(im), re.size, s, len(sched), kind)
        return re + 1j * im

    def butterflies_nd(self, x: np.ndarray, scheds, kind: int) -> np.ndarray:
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        nd, d, s, ts = _dims_scheds(x.shape, scheds)
        self.L.or_butterflies_nd(_d(re), _d(im), nd, d, s, ts, kind)
        return re + 1j * im

    # -- digit reversal --------------------------------------------------------
    def reverse_index_r(self, x: int, t: int, r: int) -> int:
        return self.L.or_reverse_index_r(x, t, r)

    def rho(self, dims, scheds) -> np.ndarray:
        n = int(np.prod(dims))
        out = np.empty(n, dtype=np.int64)
        nd, d, s, ts = _dims_scheds(dims, scheds)
        self.L.or_perm_nd(nd, d, s, ts, out.ctypes.data_as(_i64p))
        return out

    def twiddles_ref_layout(self, n: int, r: int):
        wre = np.zeros(max(1, n), dtype=np.float64)
        wim = np.zeros(max(1, n), dtype=np.float64)
        cnt = self.L.or_twiddles_ref_layout(n, r, _d(wre), _d(wim))
        return wre[:cnt], wim[:cnt]

    # -- transforms and convolutions -----------------------------------------
    def transform(self, x: np.ndarray, scheds, which: int) -> np.ndarray:
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        nd, d, s, ts = _dims_scheds(x.shape, scheds)
        self.L.or_transform_nd(_d(re), _d(im), nd, d, s, ts, which)
        return re + 1j * im

    def pa_convolution(self, x: np.ndarray, h: np.ndarray, scheds):
        """Returns (y, A^T h) -- the reference overwrites h with A^T h."""
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        hre = np.ascontiguousarray(h.real, dtype=np.float64).copy()
        him = np.ascontiguousarray(h.imag, dtype=np.float64).copy()
        nd, d, s, ts = _dims_scheds(x.shape, scheds)
        self.L.or_pa_convolution_nd(_d(re), _d(im), _d(hre), _d(him), nd, d, s, ts)
        return re + 1j * im, hre + 1j * him

    def fft_based_convolution(self, x: np.ndarray, h: np.ndarray, scheds):
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        hre = np.ascontiguousarray(h.real, dtype=np.float64).copy()
        him = np.ascontiguousarray(h.imag, dtype=np.float64).copy()
        nd, d, s, ts = _dims_scheds(x.shape, scheds)
        self.L.or_fft_based_convolution_nd(_d(re), _d(im), _d(hre), _d(him), nd, d, s, ts)
        return re + 1j * im, hre + 1j * him

    def max_error(self, y: np.ndarray, ref: np.ndarray) -> float:
        re = np.ascontiguousarray(y.real, dtype=np.float64).ravel()
        im = np.ascontiguousarray(y.imag, dtype=np.float64).ravel()
        flat = np.ascontiguousarray(np.stack([ref.real.ravel(), ref.imag.ravel()], 1).ravel())
        return self.L.or_max_error(_d(re), _d(im), _d(flat), re.size)


class RefLib:
    """The reference library compiled from its own sources (fft.h names)."""

    PATH = os.path.join(HERE, "_ref", "libref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.PATH)

    def __init__(self):
        L = C.CDLL(self.PATH)
        self.L = L
        dpp = C.POINTER(_dp)
        for r in (2, 4, 8):
            f = getattr(L, f"precompute_twiddles_r{r}")
            f.argtypes = [C.c_int, C.POINTER(dpp), C.POINTER(dpp)]
            g = getattr(L, f"precompute_index_reversal_permutation_r{r}")
            g.argtypes = [C.c_int]
            g.restype = _ip
            for kind in ("", "conjugate_", "transposed_"):
                h = getattr(L, f"apply_{kind}butterflies_r{r}")
                h.argtypes = [_dp, _dp, C.c_int, dpp, dpp]
        for r in (2, 4):
            for kind in ("", "conjugate_", "transposed_"):
                getattr(L, f"apply_{kind}butterflies_r{r}_2d").argtypes = [_dp, _dp, C.c_int, C.c_int, dpp, dpp, dpp, dpp]
                getattr(L, f"apply_{kind}butterflies_r{r}_3d").argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, dpp, dpp]
            getattr(L, f"precompute_index_reversal_permutation_r{r}_2d").argtypes = [C.c_int, C.c_int]
            getattr(L, f"precompute_index_reversal_permutation_r{r}_2d").restype = _ip
            getattr(L, f"precompute_index_reversal_permutation_r{r}_3d").argtypes = [C.c_int, C.c_int, C.c_int]
            getattr(L, f"precompute_index_reversal_permutation_r{r}_3d").restype = _ip
        L.free_twiddles.argtypes = [dpp, dpp, C.c_int]
        L.reverse_index_r2.argtypes = [C.c_int, C.c_int]
        L.reverse_index_r4.argtypes = [C.c_int, C.c_int]
        L.reverse_index_r8.argtypes = [C.c_int, C.c_int]
        L.apply_permutation_split.argtypes = [_dp, _dp, C.c_int, _ip]
        L.permutation_avoiding_convolution.argtypes = [_dp, _dp, C.c_int, C.c_int, dpp, dpp, _dp, _dp]
        L.permutation_avoiding_convolution_2d.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, _dp, _dp]
        L.permutation_avoiding_convolution_3d.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, dpp, dpp, _dp, _dp]
        L.fft_based_convolution.argtypes = [_dp, _dp, C.c_int, C.c_int, dpp, dpp, _ip, _dp, _dp]
        L.fft_based_convolution_2d.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, _ip, _dp, _dp]
        L.fft_based_convolution_3d.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, dpp, dpp, _ip, _dp, _dp]
        for nm in ("fft", "fft2", "ifft"):
            getattr(L, nm).argtypes = [_dp, _dp, C.c_int, C.c_int, dpp, dpp, _ip]
            getattr(L, nm + "_2d").argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, _ip]
            getattr(L, nm + "_3d").argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, dpp, dpp, _ip]
        for nm in ("ufft", "uifft"):
            getattr(L, nm).argtypes = [_dp, _dp, C.c_int, C.c_int, dpp, dpp]
            getattr(L, nm + "_2d").argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp]
            getattr(L, nm + "_3d").argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int, dpp, dpp, dpp, dpp, dpp, dpp]

    # twiddle tables are host double** owned by the library; keep them alive
    def twiddles(self, n: int, r: int):
        dpp = C.POINTER(_dp)
        wr, wi = dpp(), dpp()
        getattr(self.L, f"precompute_twiddles_r{r}")(n, C.byref(wr), C.byref(wi))
        return wr, wi

    def twiddle_levels(self, n: int, r: int):
        """Flattened copy of the jagged tables (same layout as Oracle.twiddles_ref_layout)."""
        wr, wi = self.twiddles(n, r)
        t = len(pure_schedule(n, r))
        outr, outi = [], []
        for q in range(t):
            cnt = (r - 1) * r ** q
            outr.append(np.ctypeslib.as_array(wr[q], shape=(cnt,)).copy())
            outi.append(np.ctypeslib.as_array(wi[q], shape=(cnt,)).copy())
        self.L.free_twiddles(wr, wi, t)
        return np.concatenate(outr) if outr else np.zeros(0), np.concatenate(outi) if outi else np.zeros(0)

    def rho(self, dims, r: int) -> np.ndarray:
        ts = [len(pure_schedule(n, r)) for n in dims]
        n = int(np.prod(dims))
        if len(dims) == 1:
            p = getattr(self.L, f"precompute_index_reversal_permutation_r{r}")(dims[0])
        elif len(dims) == 2:
            p = getattr(self.L, f"precompute_index_reversal_permutation_r{r}_2d")(*dims)
        else:
            p = getattr(self.L, f"precompute_index_reversal_permutation_r{r}_3d")(*dims)
        out = np.ctypeslib.as_array(p, shape=(n,)).astype(np.int64)
        C.CDLL(None).free(p)
        del ts
        return out

    def _tws(self, dims, r):
        return [self.twiddles(n, r) for n in dims]

    def _free(self, tws, dims, r):
        for (wr, wi), n in zip(tws, dims):
            self.L.free_twiddles(wr, wi, len(pure_schedule(n, r)))

    def butterflies(self, x: np.ndarray, r: int, kind: int) -> np.ndarray:
        """kind 0 A, 1 conj(A), 2 A^T, through apply_*butterflies_r{r}{,_2d,_3d}."""
        name = {A: "", CONJ: "conjugate_", TRANS: "transposed_"}[kind]
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        dims = list(x.shape)
        tws = self._tws(dims, r)
        flat = [a for pair in tws for a in pair]
        sfx = {1: "", 2: "_2d", 3: "_3d"}[len(dims)]
        getattr(self.L, f"apply_{name}butterflies_r{r}{sfx}")(_d(re), _d(im), *dims, *flat)
        self._free(tws, dims, r)
        return re + 1j * im

    def pa_convolution(self, x: np.ndarray, h: np.ndarray, r: int):
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        hre = np.ascontiguousarray(h.real, dtype=np.float64).copy()
        him = np.ascontiguousarray(h.imag, dtype=np.float64).copy()
        dims = list(x.shape)
        tws = self._tws(dims, r)
        flat = [a for pair in tws for a in pair]
        sfx = {1: "", 2: "_2d", 3: "_3d"}[len(dims)]
        getattr(self.L, f"permutation_avoiding_convolution{sfx}")(_d(re), _d(im), *dims, r, *flat, _d(hre), _d(him))
        self._free(tws, dims, r)
        return re + 1j * im, hre + 1j * him

    def fft_based_convolution(self, x: np.ndarray, h: np.ndarray, r: int):
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        hre = np.ascontiguousarray(h.real, dtype=np.float64).copy()
        him = np.ascontiguousarray(h.imag, dtype=np.float64).copy()
        dims = list(x.shape)
        tws = self._tws(dims, r)
        flat = [a for pair in tws for a in pair]
        rho = self.rho(dims, r).astype(np.int32)
        sfx = {1: "", 2: "_2d", 3: "_3d"}[len(dims)]
        getattr(self.L, f"fft_based_convolution{sfx}")(_d(re), _d(im), *dims, r, *flat,
                                                        rho.ctypes.data_as(_ip), _d(hre), _d(him))
        self._free(tws, dims, r)
        return re + 1j * im, hre + 1j * him

    def transform(self, x: np.ndarray, r: int, name: str) -> np.ndarray:
        """name in fft, fft2, ufft, ifft, uifft."""
        re = np.ascontiguousarray(x.real, dtype=np.float64).copy()
        im = np.ascontiguousarray(x.imag, dtype=np.float64).copy()
        dims = list(x.shape)
        tws = self._tws(dims, r)
        flat = [a for pair in tws for a in pair]
        sfx = {1: "", 2: "_2d", 3: "_3d"}[len(dims)]
        args = [_d(re), _d(im), *dims, r, *flat]
        if name in ("fft", "fft2", "ifft"):
            rho = self.rho(dims, r).astype(np.int32)
            args.append(rho.ctypes.data_as(_ip))
        getattr(self.L, name + sfx)(*args)
        self._free(tws, dims, r)
        return re + 1j * im
