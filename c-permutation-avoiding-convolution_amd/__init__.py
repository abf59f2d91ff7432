"""pacg -- permutation-avoiding FFT convolution on MI355X (gfx950).

Python front-end of ``lib/libpacg.so`` (ctypes).  Two layers:

* the reference's own C entry points (``fft.h``), same names and argument
  meaning, taking numpy float64 arrays that are modified in place exactly as
  the reference modifies them -- ``permutation_avoiding_convolution``,
  ``apply_transposed_butterflies_r4_3d``, ``fft_2d`` ...;
* ``Plan``: the extended device API (``include/pacg.h``): explicit mixed-radix
  schedules, FP64 / FP32 working precision, a cached FP64-computed kernel
  spectrum, device buffers and HIP-graph replay.

The compute always runs in the HIP kernels of libpacg.so; there is no CPU
fallback -- importing this package without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libpacg.so")

PREC_F64, PREC_F32 = 0, 1
OP_TRANS, OP_CONJ, OP_FWD, OP_FFT, OP_FFT2, OP_IFFT, OP_UIFFT, OP_PERM = range(8)
MAX_STAGES = 32

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class PacgError(RuntimeError):
    pass


class _Desc(C.Structure):
    _fields_ = [("nd", C.c_int), ("dims", C.c_int * 3), ("nstages", C.c_int * 3),
                ("radix", (C.c_int * MAX_STAGES) * 3), ("precision", C.c_int),
                ("spectrum_precision", C.c_int), ("fused_axis", C.c_int), ("mfma", C.c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise PacgError("libpacg.so is not built (%s); run __graft_entry__.build() or "
                        "python c-permutation-avoiding-convolution_amd/build.py" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    L.pacg_last_error.restype = C.c_char_p
    L.pacg_version.restype = C.c_char_p
    L.pacg_plan_create.argtypes = [C.POINTER(_vp), C.POINTER(_Desc)]
    L.pacg_plan_destroy.argtypes = [_vp]
    L.pacg_plan_set_stream.argtypes = [_vp, _vp]
    L.pacg_plan_stream.argtypes = [_vp]
    L.pacg_plan_stream.restype = _vp
    L.pacg_plan_size.argtypes = [_vp]
    L.pacg_plan_size.restype = C.c_longlong
    L.pacg_plan_rho.argtypes = [_vp, _ip]
    L.pacg_plan_describe.argtypes = [_vp, C.c_char_p, C.c_int]
    L.pacg_execute.argtypes = [_vp, C.c_int, _vp, _vp, _vp, _vp]
    L.pacg_set_kernel.argtypes = [_vp, _vp, _vp, C.c_int]
    L.pacg_get_spectrum.argtypes = [_vp, _vp, _vp, C.c_int]
    L.pacg_convolve.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pacg_convolve_pair.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pacg_convolve_fft_based.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pacg_convolve_fft_based_noalloc.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pacg_graph_capture_convolve.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.pacg_graph_replay.argtypes = [_vp]
    L.pacg_sync.argtypes = [_vp]
    L.pacg_profile_convolve.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_float),
                                        C.POINTER(C.c_char_p), C.c_int]
    L.pacg_device_sync.argtypes = []
    L.pacg_plan_search.argtypes = [C.POINTER(_Desc), C.c_int]
    L.pacg_schedule_candidates.argtypes = [C.c_int, _ip, C.c_int, C.c_int]
    L.pacg_batch_create.argtypes = [C.POINTER(_vp), C.c_int, _ip]
    L.pacg_batch_destroy.argtypes = [_vp]
    L.pacg_batch_total.argtypes = [_vp]
    L.pacg_batch_total.restype = C.c_longlong
    L.pacg_batch_offset.argtypes = [_vp, C.c_int]
    L.pacg_batch_offset.restype = C.c_longlong
    L.pacg_batch_stream.argtypes = [_vp]
    L.pacg_batch_stream.restype = _vp
    L.pacg_batch_items.argtypes = [_vp]
    for nm in ("pacg_batch_convolve", "pacg_batch_capture"):
        getattr(L, nm).argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.pacg_batch_autotune.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int]
    L.pacg_batch_set_variant.argtypes = [_vp, C.c_int, C.c_int]
    L.pacg_batch_set_dealing.argtypes = [_vp, C.c_int]
    L.pacg_batch_get_dealing.argtypes = [_vp]
    L.pacg_batch_get_variant.argtypes = [_vp, C.c_int]
    L.pacg_batch_num_variants.argtypes = [C.c_int]
    L.pacg_batch_schedule.argtypes = [C.c_int, C.c_int, _ip, C.c_int]
    L.pacg_batch_replay.argtypes = [_vp]
    L.pacg_batch_sync.argtypes = [_vp]
    L.pacg_plan_autotune.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int]
    L.pacg_plan_launch_candidates.argtypes = [_vp, C.c_int, C.POINTER(C.c_char_p), C.c_int]
    L.pacg_plan_pin_launch.argtypes = [_vp, C.c_int, C.c_char_p]
    L.pacg_plan_stream_input.argtypes = [_vp, C.c_int]
    L.pacg_bw_probe.argtypes = [C.c_int, _vp, _vp, C.c_uint64, _vp]
    L.pacg_hbm_probe.argtypes = [C.c_int, _vp, _vp, C.c_uint64, C.c_int, _vp, C.POINTER(C.c_float)]
    L.pacg_launch_floor.argtypes = [C.c_int, _vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.pacg_dropin_release.argtypes = []
    L.pacg_dropin_release.restype = None
    L.pacg_desc_init.argtypes = [C.POINTER(_Desc), C.c_int, _ip, C.c_int, C.c_int]
    # runtime glue (device memory, copies, events)
    L.pacg_dev_malloc.argtypes = [C.POINTER(_vp), C.c_size_t]
    L.pacg_dev_free.argtypes = [_vp]
    L.pacg_memcpy_h2d.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.pacg_memcpy_d2h.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.pacg_memcpy_d2d.argtypes = [_vp, _vp, C.c_size_t, _vp]
    L.pacg_stream_sync.argtypes = [_vp]
    L.pacg_device_count.argtypes = [_ip]
    L.pacg_hip_error_string.argtypes = [C.c_int]
    L.pacg_hip_error_string.restype = C.c_char_p
    L.pacg_event_create.argtypes = [C.POINTER(_vp)]
    L.pacg_event_destroy.argtypes = [_vp]
    L.pacg_event_record.argtypes = [_vp, _vp]
    L.pacg_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), _vp, _vp]
    return L


lib = _load()


def last_error() -> str:
    return lib.pacg_last_error().decode()


def _check(rc: int, what: str):
    if rc != 0:
        raise PacgError("%s failed (%d): %s" % (what, rc, last_error()))


def device_count() -> int:
    n = C.c_int(0)
    rc = lib.pacg_device_count(C.byref(n))
    return n.value if rc == 0 else 0


# ---------------------------------------------------------------------------
# device memory (split complex arrays)
# ---------------------------------------------------------------------------
class DeviceArray:
    """A device buffer of ``n`` real elements of ``dtype`` (float64 / float32)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = _vp()
        _check(lib.pacg_dev_malloc(C.byref(p), max(1, self.nbytes)), "device allocation")
        self.ptr = p.value

    # Plans run on their own non-blocking streams, which the null stream does
    # not order against: host copies wait for the whole device first.
    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        assert a.size == self.n
        _check(lib.pacg_device_sync(), "device sync")
        _check(lib.pacg_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes, stream), "H2D copy")
        _check(lib.pacg_stream_sync(stream), "sync")

    def download(self, stream=None) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib.pacg_device_sync(), "device sync")
        _check(lib.pacg_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream), "D2H copy")
        _check(lib.pacg_stream_sync(stream), "sync")
        return out

    @classmethod
    def view(cls, base: "DeviceArray", offset: int, n: int) -> "DeviceArray":
        """A non-owning window of ``n`` elements starting ``offset`` elements into ``base``."""
        v = cls.__new__(cls)
        v.n, v.dtype = int(n), base.dtype
        v.nbytes = v.n * v.dtype.itemsize
        assert offset + n <= base.n
        v.ptr = base.ptr + int(offset) * base.dtype.itemsize
        v._owner = base
        return v

    def free(self):
        if self.ptr and not hasattr(self, "_owner"):
            lib.pacg_dev_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceComplex:
    """Split complex device array (re, im) with numpy round trips."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        n = int(np.prod(self.shape))
        self.re = DeviceArray(n, dtype)
        self.im = DeviceArray(n, dtype)
        self.dtype = np.dtype(dtype)

    @classmethod
    def shifted(cls, shape, dtype=np.float64, offset: int = 1):
        """Arrays that start ``offset`` elements into their allocation (unaligned base pointers)."""
        d = cls.__new__(cls)
        d.shape = tuple(int(s) for s in shape)
        n = int(np.prod(d.shape))
        d._bases = (DeviceArray(n + offset, dtype), DeviceArray(n + offset, dtype))
        d.re = DeviceArray.view(d._bases[0], offset, n)
        d.im = DeviceArray.view(d._bases[1], offset, n)
        d.dtype = np.dtype(dtype)
        return d

    @classmethod
    def from_numpy(cls, z: np.ndarray, dtype=np.float64):
        d = cls(z.shape, dtype)
        d.upload(z)
        return d

    def upload(self, z: np.ndarray):
        self.re.upload(np.real(z))
        self.im.upload(np.imag(z))

    def numpy(self) -> np.ndarray:
        return (self.re.download().astype(np.float64) + 1j * self.im.download().astype(np.float64)).reshape(self.shape)


# ---------------------------------------------------------------------------
# extended plan API
# ---------------------------------------------------------------------------
def _make_desc(dims, schedules, radix, precision, spectrum_precision, fused_axis, mfma) -> _Desc:
    """pacg_desc for the Plan arguments (schedules per axis, or one radix for all, or defaults)."""
    d = _Desc()
    if radix is not None:
        arr = (C.c_int * 3)(*(list(dims) + [0] * (3 - len(dims))))
        _check(lib.pacg_desc_init(C.byref(d), len(dims), arr, int(radix), int(precision)), "pacg_desc_init")
    else:
        d.nd = len(dims)
        for a, n in enumerate(dims):
            d.dims[a] = n
            sc = (schedules or [None] * len(dims))[a]
            if isinstance(sc, str) and sc == "batch":
                d.nstages[a] = -1  # batch axis: rows convolved independently
            elif sc:
                d.nstages[a] = len(sc)
                for q, r in enumerate(sc):
                    d.radix[a][q] = int(r)
        d.precision = int(precision)
    d.spectrum_precision = int(spectrum_precision)
    d.fused_axis = int(fused_axis)
    d.mfma = 1 if mfma else 0
    return d


def search_schedules(dims, schedules=None, precision=PREC_F64, spectrum_precision=PREC_F64, fused_axis=-1,
                     mfma=False, reps: int = 5):
    """pacg_plan_search: the fastest radix schedule per axis (coordinate descent from `schedules`,
    defaults where None), then, when `fused_axis` is -1, the fastest axis for the fused pass.
    Returns (schedules, number of axes whose schedule changed, fused axis: -1 for the default
    rule)."""
    dims = [int(x) for x in dims]
    d = _make_desc(dims, schedules, None, precision, spectrum_precision, fused_axis, mfma)
    rc = lib.pacg_plan_search(C.byref(d), int(reps))
    if rc < 0:
        _check(-rc, "pacg_plan_search")
    out = []
    for a in range(len(dims)):
        ns = d.nstages[a]
        out.append("batch" if ns < 0 else ([int(d.radix[a][q]) for q in range(ns)] if ns > 0 else None))
    return out, int(rc), int(d.fused_axis)


def schedule_candidates(n: int):
    """pacg_schedule_candidates: the schedules pacg_plan_search tries for a length n."""
    cnt = lib.pacg_schedule_candidates(int(n), None, 0, MAX_STAGES)
    if cnt < 0:
        _check(-cnt, "pacg_schedule_candidates")
    buf = (C.c_int * (max(cnt, 1) * MAX_STAGES))()
    lib.pacg_schedule_candidates(int(n), buf, cnt, MAX_STAGES)
    return [[v for v in buf[i * MAX_STAGES:(i + 1) * MAX_STAGES] if v] for i in range(cnt)]


class Plan:
    """Device plan for one shape / schedule / precision (pacg_plan_create).

    ``schedules``: per axis, a list of radices (stage 0 = unit stride) or None
    for the default; ``radix``: pure radix for every axis (alternative).
    """

    def __init__(self, dims, schedules=None, radix=None, precision=PREC_F64, spectrum_precision=PREC_F64,
                 fused_axis=-1, mfma=False):
        dims = [int(d) for d in dims]
        d = _make_desc(dims, schedules, radix, precision, spectrum_precision, fused_axis, mfma)
        self.dims = dims
        self.precision = int(precision)
        self.dtype = np.float64 if precision == PREC_F64 else np.float32
        p = _vp()
        _check(lib.pacg_plan_create(C.byref(p), C.byref(d)), "pacg_plan_create")
        self.ptr = p.value
        self.n = int(lib.pacg_plan_size(self.ptr))

    def describe(self) -> str:
        buf = C.create_string_buffer(4096)
        lib.pacg_plan_describe(self.ptr, buf, 4096)
        return buf.value.decode()

    def launches_per_convolve(self) -> int:
        buf = C.create_string_buffer(4096)
        return int(lib.pacg_plan_describe(self.ptr, buf, 4096))

    @property
    def stream(self):
        return lib.pacg_plan_stream(self.ptr)

    def rho(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        _check(lib.pacg_plan_rho(self.ptr, out.ctypes.data_as(_ip)), "pacg_plan_rho")
        return out

    def alloc(self) -> DeviceComplex:
        return DeviceComplex(self.dims, self.dtype)

    def execute(self, op: int, x: DeviceComplex, y: DeviceComplex | None = None):
        y = y or x
        _check(lib.pacg_execute(self.ptr, op, x.re.ptr, x.im.ptr, y.re.ptr, y.im.ptr), "pacg_execute")
        return y

    def set_kernel(self, k: np.ndarray):
        kr = np.ascontiguousarray(np.real(k), dtype=np.float64)
        ki = np.ascontiguousarray(np.imag(k), dtype=np.float64)
        _check(lib.pacg_set_kernel(self.ptr, kr.ctypes.data, ki.ctypes.data, 0), "pacg_set_kernel")

    def spectrum(self) -> np.ndarray:
        sr = np.empty(self.n)
        si = np.empty(self.n)
        _check(lib.pacg_get_spectrum(self.ptr, sr.ctypes.data, si.ctypes.data, 0), "pacg_get_spectrum")
        return (sr + 1j * si).reshape(self.dims)

    def convolve(self, x: DeviceComplex, y: DeviceComplex | None = None):
        y = y or x
        _check(lib.pacg_convolve(self.ptr, x.re.ptr, x.im.ptr, y.re.ptr, y.im.ptr), "pacg_convolve")
        return y

    def convolve_pair(self, x: DeviceComplex, h: DeviceComplex):
        _check(lib.pacg_convolve_pair(self.ptr, x.re.ptr, x.im.ptr, h.re.ptr, h.im.ptr), "pacg_convolve_pair")

    def convolve_fft_based(self, x: DeviceComplex, h: DeviceComplex):
        _check(lib.pacg_convolve_fft_based(self.ptr, x.re.ptr, x.im.ptr, h.re.ptr, h.im.ptr),
               "pacg_convolve_fft_based")

    def capture(self, x: DeviceComplex, y: DeviceComplex):
        _check(lib.pacg_graph_capture_convolve(self.ptr, x.re.ptr, x.im.ptr, y.re.ptr, y.im.ptr), "graph capture")

    def replay(self):
        _check(lib.pacg_graph_replay(self.ptr), "graph replay")

    def sync(self):
        _check(lib.pacg_sync(self.ptr), "pacg_sync")

    def autotune(self, x: DeviceComplex, y: DeviceComplex, reps: int = 5) -> int:
        """Pick the fastest compiled tile shape for each launch of convolve (x -> y)."""
        rc = lib.pacg_plan_autotune(self.ptr, x.re.ptr, x.im.ptr, y.re.ptr, y.im.ptr, int(reps))
        if rc < 0:
            raise PacgError("pacg_plan_autotune failed: %s" % last_error())
        return rc

    def candidates(self, launch: int):
        """Names of the compiled kernels (tile shapes) that can run launch `launch` of convolve."""
        names = (C.c_char_p * 64)()
        n = lib.pacg_plan_launch_candidates(self.ptr, int(launch), names, 64)
        if n < 0:
            raise PacgError("pacg_plan_launch_candidates failed: %s" % last_error())
        return [names[i].decode() for i in range(min(n, 64))]

    def pin(self, launch: int, name: str):
        """Run launch `launch` of convolve with the named kernel (one of candidates(launch))."""
        if lib.pacg_plan_pin_launch(self.ptr, int(launch), name.encode()):
            raise PacgError("pacg_plan_pin_launch failed: %s" % last_error())

    def stream_input(self, on: bool = True):
        """Load x non-temporally in the first launch of convolve (the autotuner's other choice)."""
        if lib.pacg_plan_stream_input(self.ptr, int(bool(on))):
            raise PacgError("pacg_plan_stream_input failed: %s" % last_error())

    def profile(self, x: DeviceComplex, y: DeviceComplex, reps: int = 10):
        """Mean duration (ms) of every kernel launch of one convolve, HIP events on the plan stream."""
        buf = (C.c_float * 64)()
        names = (C.c_char_p * 64)()
        nl = lib.pacg_profile_convolve(self.ptr, x.re.ptr, x.im.ptr, y.re.ptr, y.im.ptr, int(reps), buf, names, 64)
        if nl < 0:
            raise PacgError("pacg_profile_convolve failed: %s" % last_error())
        return [(names[i].decode(), float(buf[i])) for i in range(nl)]

    def close(self):
        if getattr(self, "ptr", None):
            lib.pacg_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bw_probe(mode: int, dst: DeviceArray, src: DeviceArray, nbytes: int, stream=None):
    """HBM probe kernel: mode 0 read, 1 write, 2 copy (16-byte accesses), one launch, no timing."""
    _check(lib.pacg_bw_probe(int(mode), dst.ptr, src.ptr, int(nbytes), stream), "bw probe")


def hbm_probe(mode: int, dst, src, nbytes: int, reps: int = 10, stream=None) -> float:
    """pacg_hbm_probe: mean ms of one probe launch (mode 0 read, 1 write, 2 copy of nbytes);
    dst / src are DeviceArray objects (or None where the mode does not use them)."""
    ms = C.c_float(0.0)
    rc = lib.pacg_hbm_probe(int(mode), dst.ptr if dst is not None else None, src.ptr if src is not None else None,
                            C.c_uint64(int(nbytes)), int(reps), stream, C.byref(ms))
    if rc != 0:
        raise PacgError("pacg_hbm_probe: %s" % last_error())
    return ms.value


class Batch:
    """Many FP32 convolutions of lengths 2^8..2^14 in one persistent launch (pacg_batch_*)."""

    def __init__(self, log2_lengths):
        m = np.ascontiguousarray(log2_lengths, dtype=np.int32)
        p = _vp()
        _check(lib.pacg_batch_create(C.byref(p), int(m.size), m.ctypes.data_as(_ip)), "pacg_batch_create")
        self.ptr = p.value
        self.log2n = m.copy()
        self.count = int(m.size)
        self.total = int(lib.pacg_batch_total(self.ptr))
        self.offsets = np.array([lib.pacg_batch_offset(self.ptr, i) for i in range(self.count)], dtype=np.int64)

    @property
    def stream(self):
        return lib.pacg_batch_stream(self.ptr)

    def alloc(self) -> DeviceComplex:
        return DeviceComplex((self.total,), np.float32)

    def pack(self, arrays) -> np.ndarray:
        return np.concatenate([np.asarray(a, dtype=np.complex128).ravel() for a in arrays])

    def unpack(self, flat: np.ndarray):
        return [flat[o:o + (1 << int(m))] for o, m in zip(self.offsets, self.log2n)]

    def convolve(self, x: DeviceComplex, k: DeviceComplex, y: DeviceComplex):
        _check(lib.pacg_batch_convolve(self.ptr, x.re.ptr, x.im.ptr, k.re.ptr, k.im.ptr, y.re.ptr, y.im.ptr),
               "pacg_batch_convolve")

    def autotune(self, x: DeviceComplex, k: DeviceComplex, y: DeviceComplex, reps: int = 3) -> int:
        rc = lib.pacg_batch_autotune(self.ptr, x.re.ptr, x.im.ptr, k.re.ptr, k.im.ptr, y.re.ptr, y.im.ptr, int(reps))
        if rc < 0:
            raise PacgError("pacg_batch_autotune failed: %s" % last_error())
        return rc

    def set_variant(self, log2n: int, v: int):
        _check(lib.pacg_batch_set_variant(self.ptr, int(log2n), int(v)), "pacg_batch_set_variant")

    def variant(self, log2n: int) -> int:
        return int(lib.pacg_batch_get_variant(self.ptr, int(log2n)))

    DEAL_LPT, DEAL_SLICES = 0, 1

    def set_dealing(self, dealing: int):
        """Items to workgroups: longest processing time first (0) or contiguous slices (1)."""
        _check(lib.pacg_batch_set_dealing(self.ptr, int(dealing)), "pacg_batch_set_dealing")

    def dealing(self) -> int:
        return int(lib.pacg_batch_get_dealing(self.ptr))

    @staticmethod
    def num_variants(log2n: int) -> int:
        return int(lib.pacg_batch_num_variants(int(log2n)))

    @staticmethod
    def schedule(log2n: int, v: int):
        buf = (C.c_int * 32)()
        ns = lib.pacg_batch_schedule(int(log2n), int(v), buf, 32)
        if ns < 0:
            raise PacgError(last_error())
        return [buf[i] for i in range(ns)]

    def capture(self, x: DeviceComplex, k: DeviceComplex, y: DeviceComplex):
        _check(lib.pacg_batch_capture(self.ptr, x.re.ptr, x.im.ptr, k.re.ptr, k.im.ptr, y.re.ptr, y.im.ptr),
               "pacg_batch_capture")

    def replay(self):
        _check(lib.pacg_batch_replay(self.ptr), "pacg_batch_replay")

    def sync(self):
        _check(lib.pacg_batch_sync(self.ptr), "pacg_batch_sync")

    def close(self):
        if getattr(self, "ptr", None):
            lib.pacg_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Timer:
    """HIP events on a stream (the stream the kernels run on)."""

    def __init__(self, stream):
        self.stream = stream
        a, b = _vp(), _vp()
        _check(lib.pacg_event_create(C.byref(a)), "event")
        _check(lib.pacg_event_create(C.byref(b)), "event")
        self.a, self.b = a.value, b.value

    def start(self):
        _check(lib.pacg_event_record(self.a, self.stream), "event record")

    def stop(self) -> float:
        _check(lib.pacg_event_record(self.b, self.stream), "event record")
        _check(lib.pacg_stream_sync(self.stream), "sync")
        ms = C.c_float(0)
        _check(lib.pacg_event_elapsed_ms(C.byref(ms), self.a, self.b), "elapsed")
        return ms.value

    def __del__(self):
        try:
            lib.pacg_event_destroy(self.a)
            lib.pacg_event_destroy(self.b)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# the reference's entry points (fft.h), numpy arrays modified in place
# ---------------------------------------------------------------------------
def _f64(a: np.ndarray):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
        raise TypeError("expected a C-contiguous float64 numpy array")
    return a.ctypes.data_as(_dp)


def _rho(rho):
    r = np.ascontiguousarray(rho, dtype=np.int32)
    return r, r.ctypes.data_as(_ip)


def _raise_if_error(name):
    msg = last_error()
    if msg:
        raise PacgError("%s: %s" % (name, msg))


_NULL_TW = None  # the GPU path does not read the caller's twiddle tables


def _bfly(name, nd):
    fn = getattr(lib, name)
    fn.argtypes = [_dp, _dp] + [C.c_int] * nd + [_vp] * (2 * nd)
    fn.restype = None

    def call(re, im, *dims_and_tw):
        dims = [int(d) for d in dims_and_tw[:nd]]
        fn(_f64(re), _f64(im), *dims, *([_NULL_TW] * (2 * nd)))
        _raise_if_error(name)

    call.__name__ = name
    call.__doc__ = "GPU %s (reference fft.h); twiddle arguments are accepted and ignored." % name
    return call


for _r in (2, 4, 8):
    for _k in ("", "conjugate_", "transposed_"):
        _n = "apply_%sbutterflies_r%d" % (_k, _r)
        globals()[_n] = _bfly(_n, 1)
for _r in (2, 4):
    for _k in ("", "conjugate_", "transposed_"):
        for _sfx, _nd in (("_2d", 2), ("_3d", 3)):
            _n = "apply_%sbutterflies_r%d%s" % (_k, _r, _sfx)
            globals()[_n] = _bfly(_n, _nd)


def _transform(name, nd, with_rho):
    fn = getattr(lib, name)
    fn.argtypes = [_dp, _dp] + [C.c_int] * (nd + 1) + [_vp] * (2 * nd) + ([_ip] if with_rho else [])
    fn.restype = None

    def call(re, im, *args):
        dims = [int(d) for d in args[:nd]]
        r = int(args[nd])
        rest = list(args[nd + 1:])
        extra = []
        keep = None
        if with_rho:
            keep, p = _rho(rest[-1])
            extra = [p]
        fn(_f64(re), _f64(im), *dims, r, *([_NULL_TW] * (2 * nd)), *extra)
        del keep
        _raise_if_error(name)

    call.__name__ = name
    return call


for _sfx, _nd in (("", 1), ("_2d", 2), ("_3d", 3)):
    for _nm, _wr in (("fft", True), ("fft2", True), ("ifft", True), ("ufft", False), ("uifft", False)):
        globals()[_nm + _sfx] = _transform(_nm + _sfx, _nd, _wr)


def _conv(name, nd, with_rho):
    fn = getattr(lib, name)
    fn.argtypes = [_dp, _dp] + [C.c_int] * (nd + 1) + [_vp] * (2 * nd) + ([_ip] if with_rho else []) + [_dp, _dp]
    fn.restype = None

    def call(re, im, *args):
        dims = [int(d) for d in args[:nd]]
        r = int(args[nd])
        rest = list(args[nd + 1:])
        h_re, h_im = rest[-2], rest[-1]
        extra = []
        keep = None
        if with_rho:
            keep, p = _rho(rest[-3])
            extra = [p]
        fn(_f64(re), _f64(im), *dims, r, *([_NULL_TW] * (2 * nd)), *extra, _f64(h_re), _f64(h_im))
        del keep
        _raise_if_error(name)

    call.__name__ = name
    return call


for _sfx, _nd in (("", 1), ("_2d", 2), ("_3d", 3)):
    globals()["permutation_avoiding_convolution" + _sfx] = _conv("permutation_avoiding_convolution" + _sfx, _nd, False)
    globals()["fft_based_convolution" + _sfx] = _conv("fft_based_convolution" + _sfx, _nd, True)


def launch_floor(reps: int = 1000, stream=None):
    """(direct, graph) microseconds per launch of an empty kernel: back-to-back launches, and
    replays of a HIP graph holding one launch."""
    d, g = C.c_float(0), C.c_float(0)
    rc = lib.pacg_launch_floor(int(reps), stream, C.byref(d), C.byref(g))
    if rc != 0:
        raise PacgError("pacg_launch_floor failed (HIP error %d)" % rc)
    return float(d.value), float(g.value)


def release():
    """Free the device plans cached by the reference-named entry points."""
    lib.pacg_dropin_release()
