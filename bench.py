"""Benchmark: permutation-avoiding 3D convolution, 256^3 complex, on one MI355X.

Workload (one "step" = one convolution y = (1/n) conj(A)[(A^T k) .* (A^T x)]):
  3D 256^3 complex, radix-8 schedule per axis (stages 4, 8, 8), bulk butterflies
  in FP32 (five launches: one axis per pass, the middle one fused with the product), twiddles generated in FP64 on the GPU and rounded, kernel spectrum
  A^T k computed in FP64 on the GPU; x re/im i.i.d. N(0,1) and k a separable
  Gaussian stencil (sigma = 3 cells), seed 2.  Inputs are resident in HBM
  when the timed region starts; the spectrum is computed once, before it.

Reported beside the headline FP32 number: the plain FP64 GPU path on the
same workload (the speed-up denominator), relative L2 / max error of the
FP32 result against an FP64 reference (numpy.fft in complex128), the
roofline of the dominant kernel (HIP events around every launch on the
plan's stream) and of the whole convolution, and the reference's own C code
timed on this host's cores (oracle/_ref, built from the reference sources).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "c-permutation-avoiding-convolution_amd")
HBM_PEAK_GBS = 8000.0
N1 = 256
SCHED = [4, 8, 8]
SCHEDS32 = [SCHED] * 3
SCHEDS64 = [SCHED] * 3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_pacg():
    spec = importlib.util.spec_from_file_location("pacg", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pacg"] = mod
    spec.loader.exec_module(mod)
    return mod


def time_convs(pacg, pl, dx, dy, steps, warmup):
    for _ in range(warmup):
        pl.convolve(dx, dy)
    pl.sync()
    t = pacg.Timer(pl.stream)
    t.start()
    for _ in range(steps):
        pl.convolve(dx, dy)
    ms = t.stop()
    pl.sync()
    return ms / steps


def time_graph(pacg, pl, dx, dy, steps, warmup):
    """ms per convolution replayed from a HIP graph captured once (pacg_graph_capture_convolve):
    the same launches without the per-launch host submission."""
    pl.capture(dx, dy)
    for _ in range(warmup):
        pl.replay()
    pl.sync()
    t = pacg.Timer(pl.stream)
    t.start()
    for _ in range(steps):
        pl.replay()
    ms = t.stop()
    pl.sync()
    return ms / steps


def step_distribution(pacg, pl, dx, dy, steps):
    """Each of `steps` convolutions timed on its own (HIP events around it, a host sync between):
    median, standard deviation, min and max in ms."""
    t = pacg.Timer(pl.stream)
    ms = []
    for _ in range(steps):
        t.start()
        pl.convolve(dx, dy)
        ms.append(t.stop())
    v = np.array(ms)
    return {"median_ms": float(np.median(v)), "std_ms": float(v.std()), "min_ms": float(v.min()),
            "max_ms": float(v.max()), "steps": steps}


def profile_launches(pacg, pl, dx, dy, reps):
    return pl.profile(dx, dy, reps)


def alg_bytes_launch(name, n, e, s):
    """Algorithmic HBM bytes of one pass launch: read + write the array, plus the spectrum if fused."""
    return 2 * n * e + (n * s if "_fused_" in name else 0)


def conv_alg_bytes(launches, n, e, s):
    """Algorithmic HBM bytes of one convolution run as `launches` read-modify-write launches of the
    array, one of which also reads the spectrum."""
    return launches * 2 * n * e + n * s


def measured_traffic(variant):
    """HBM bytes per launch of `variant` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs,
    each counter scaled by what it reads for a copy kernel of known size (profiles/traffic_3d256_fp32.json),
    with where the figure comes from: (bytes or None, source).  The counters are not read in this run."""
    rel = os.path.join("profiles", "traffic_3d256_fp32.json")
    path = os.path.join(ROOT, rel)
    if not os.path.exists(path):
        return None, "no %s" % rel
    with open(path) as f:
        d = json.load(f)
    entry = d.get("kernels", {}).get(variant)
    if not entry:
        return None, "%s has no entry for %s (the dominant kernel of this run); regenerate it" % (rel, variant)
    return float(entry["traffic_bytes"]), "%s (%s; not measured in this run)" % (rel, "; ".join(d.get("commands", [])))


def host_cpu_info():
    """CPU model (/proc/cpuinfo), cores of the host and of this process, and the compile flags of the
    CPU reference library (oracle/Makefile REF_CFLAGS)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    flags = None
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("REF_CFLAGS"):
                    flags = line.split("=", 1)[1].strip()
    except OSError:
        pass
    return {"cpu_model": model, "host_cores": os.cpu_count(), "process_cores": len(os.sched_getaffinity(0)),
            "compile_flags": flags}


def _load_synth():
    """synth.py by path (numpy only): the CPU child process loads no GPU library."""
    spec = importlib.util.spec_from_file_location("pacg_synth_cpu", os.path.join(PKG_DIR, "synth.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _cpu_baseline_child(reps, core, q):
    """Pinned to one core (the reference's protocol: taskset -c 0, bench1d_forward_fft.sh:109), the
    reference's permutation_avoiding_convolution_3d (r = 4: 256 = 4^4) on the headline's inputs:
    one warm-up and `reps` timed calls, h re-copied before each, untimed (bench1d_convolution.c:238-260)."""
    os.sched_setaffinity(0, {core})
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, RefLib  # checker / baseline only

    x, k = _load_synth().workload_inputs("3d_256cubed")
    kind = "reference" if RefLib.available() else "port"
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    hre0 = np.ascontiguousarray(k.real)
    him0 = np.ascontiguousarray(k.imag)
    times = []
    if kind == "reference":
        ref = RefLib()
        _dp = C.POINTER(C.c_double)
        tws = ref._tws([N1] * 3, 4)
        flat = [a for pair in tws for a in pair]
        f = ref.L.permutation_avoiding_convolution_3d
        for it in range(reps + 1):
            xr, xi = re.copy(), im.copy()
            hr, hi = hre0.copy(), him0.copy()  # the call overwrites h (convolution.c:45-53)
            t0 = time.perf_counter()
            f(xr.ctypes.data_as(_dp), xi.ctypes.data_as(_dp), N1, N1, N1, 4, *flat,
              hr.ctypes.data_as(_dp), hi.ctypes.data_as(_dp))
            dt = time.perf_counter() - t0
            if it > 0:
                times.append(dt)
        ref._free(tws, [N1] * 3, 4)
    else:
        orc = Oracle()
        for it in range(reps + 1):
            t0 = time.perf_counter()
            orc.pa_convolution(x, k, [[4] * 4] * 3)
            dt = time.perf_counter() - t0
            if it > 0:
                times.append(dt)
    q.put((kind, times, sorted(os.sched_getaffinity(0))))


def cpu_baseline(reps):
    """The reference's 256^3 convolution on one pinned host core, in a child process (see
    _cpu_baseline_child), with the host's CPU model, core count and the library's compile flags."""
    import multiprocessing as mp

    core = min(os.sched_getaffinity(0))
    ctx = mp.get_context("spawn")
    q = ctx.Queue()
    p_ = ctx.Process(target=_cpu_baseline_child, args=(reps, core, q))
    p_.start()
    kind, times, pinned = q.get(timeout=1800)
    p_.join(timeout=60)
    mean = float(np.mean(times))
    return {"value": 1.0 / mean, "unit": "conv/s", "cores": 1, "kind": kind, "pinned_cores": pinned,
            "warmup_calls": 1, "timed_calls": len(times), "mean_s": mean, "std_s": float(np.std(times)),
            "sample": "256^3 complex FP64 PA convolution (h transformed in the call: 3 transforms), radix 4, "
                      "one core pinned with sched_setaffinity, 1 warm-up + %d timed calls (mean %.3f s +/- %.3f s), "
                      "h re-copied before each call, untimed" % (len(times), mean, float(np.std(times))),
            "n": N1 ** 3, **host_cpu_info()}


def cpu_baseline_workloads(synth, gpu):
    """The reference's own entry points on one host core for the other workloads, at the radix the
    reference runs them with (2 for 2^16, 2048^2 and the batched lengths, 4 for 4096 x 256): one
    warm-up and one or more timed calls each, h re-copied before every call.  `gpu` maps a workload
    to the GPU's ms for the same work (None when it was not measured)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib  # checker / baseline only

    if not RefLib.available():
        return None
    ref = RefLib()
    _dp = C.POINTER(C.c_double)
    L = ref.L

    def arrays(x, k):
        return [np.ascontiguousarray(a, dtype=np.float64) for a in (x.real, x.imag, k.real, k.imag)]

    def timed(call, reps):
        ts = []
        for it in range(reps + 1):
            t = call()
            if it:
                ts.append(t)
        return float(np.mean(ts))

    def one(f, dims, r, x, k):
        re0, im0, hr0, hi0 = arrays(x.ravel(), k.ravel())
        tws = ref._tws(dims, r)
        flat = [a for pair in tws for a in pair]

        def call():
            re, im, hr, hi = re0.copy(), im0.copy(), hr0.copy(), hi0.copy()
            t0 = time.perf_counter()
            if len(dims) == 1:
                f(re.ctypes.data_as(_dp), im.ctypes.data_as(_dp), dims[0], r, *flat, hr.ctypes.data_as(_dp),
                  hi.ctypes.data_as(_dp))
            else:
                f(re.ctypes.data_as(_dp), im.ctypes.data_as(_dp), *dims, r, *flat, hr.ctypes.data_as(_dp),
                  hi.ctypes.data_as(_dp))
            return time.perf_counter() - t0
        return call, tws

    out = {}
    for wl, f, r, reps in (("1d_65536", L.permutation_avoiding_convolution, 2, 10),
                           ("2d_2048x2048", L.permutation_avoiding_convolution_2d, 2, 1),
                           ("2d_4096x256", L.permutation_avoiding_convolution_2d, 4, 3)):
        x, k = synth.workload_inputs(wl)
        dims = list(x.shape)
        call, tws = one(f, dims, r, x, k)
        sec = timed(call, reps)
        ref._free(tws, dims, r)
        out[wl] = {"s_per_conv": sec, "radix": r, "timed_calls": reps}
    # batched: every pair through the 1D entry point, twiddles per length prepared beforehand
    xs, ks = synth.batched_inputs(4, 4096)
    lens = sorted({len(x) for x in xs})
    tw = {n: ref._tws([n], 2) for n in lens}
    data = [arrays(x, k) for x, k in zip(xs, ks)]

    def batch_call():
        work = [[a.copy() for a in d] for d in data]
        t0 = time.perf_counter()
        for (re, im, hr, hi) in work:
            n = re.size
            flat = [a for pair in tw[n] for a in pair]
            L.permutation_avoiding_convolution(re.ctypes.data_as(_dp), im.ctypes.data_as(_dp), n, 2, *flat,
                                               hr.ctypes.data_as(_dp), hi.ctypes.data_as(_dp))
        return time.perf_counter() - t0
    sec = timed(batch_call, 1)
    for n in lens:
        ref._free(tw[n], [n], 2)
    out["batched_4096"] = {"s_per_batch": sec, "radix": 2, "timed_calls": 1}
    for wl, e in out.items():
        g = gpu.get(wl)
        sec = e.get("s_per_conv", e.get("s_per_batch"))
        e["gpu_ms"] = g
        e["gpu_speedup"] = (sec * 1e3 / g) if g else None
    return {"cores": 1, "kind": "reference", "unit": "s per call", "workloads": out,
            "note": "FP64 reference entry points on one host core; the GPU figures are the FP64 plans "
                    "(1D, 2048^2), the FP32 VALU plan (4096 x 256) and the FP32 batched launch"}


def _cpu_throughput_worker(seed, barrier, out_q):
    """One single-threaded process of the CPU throughput figure: the reference's
    permutation_avoiding_convolution_3d at 256^3, radix 4, a warm-up call, then one call
    timed from a barrier shared by all processes.  Inputs are its own random arrays
    (the arithmetic does not depend on the values)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib  # checker / baseline only

    g = np.random.default_rng(seed)
    shape = (N1 * N1 * N1,)
    re, im, hre0, him0 = (g.standard_normal(shape) for _ in range(4))
    ref = RefLib()
    _dp = C.POINTER(C.c_double)
    tws = ref._tws([N1] * 3, 4)
    flat = [a for pair in tws for a in pair]
    f = ref.L.permutation_avoiding_convolution_3d
    dt = 0.0
    for it in range(2):
        xr, xi, hr, hi = re.copy(), im.copy(), hre0.copy(), him0.copy()
        if it == 1:
            barrier.wait()
        t0 = time.perf_counter()
        f(xr.ctypes.data_as(_dp), xi.ctypes.data_as(_dp), N1, N1, N1, 4, *flat, hr.ctypes.data_as(_dp),
          hi.ctypes.data_as(_dp))
        dt = time.perf_counter() - t0
    ref._free(tws, [N1] * 3, 4)
    out_q.put(dt)


def cpu_throughput(procs):
    """`procs` independent single-threaded processes of the reference convolution at once (child
    processes started fresh, none touches the GPU): aggregate conv/s = procs / slowest call."""
    import multiprocessing as mp

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import RefLib  # checker / baseline only

    if not RefLib.available():
        return None
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(procs)
    q = ctx.Queue()
    ps = [ctx.Process(target=_cpu_throughput_worker, args=(100 + i, barrier, q)) for i in range(procs)]
    for p_ in ps:
        p_.start()
    dts = [q.get(timeout=600) for _ in ps]
    for p_ in ps:
        p_.join(timeout=60)
    slow = max(dts)
    return {"value": procs / slow, "unit": "conv/s", "processes": procs, "kind": "reference",
            "sample": "256^3 complex FP64 PA convolution, radix 4, %d single-threaded processes at once, one timed "
                      "call each after a warm-up (slowest %.3f s, fastest %.3f s)" % (procs, slow, min(dts))}


def bench_fp64_workloads(pacg, synth, steps, warmup):
    """1D 2^16 (radix 2) and 2D 2048^2 (radix 4 with one radix-2 stage), FP64: device-resident
    plan rate, and the drop-in entry point's rate from host arrays (PCIe included)."""
    out = {}
    for wl, sched_of, name in (("1d_65536", lambda n: [2] * 16, "fp64_1d_65536_r2"),
                               ("2d_2048x2048", lambda n: [4, 4, 4, 4, 4, 2], "fp64_2d_2048x2048_r4")):
        x, k = synth.workload_inputs(wl)
        n = x.size
        d = x.ndim
        scheds = [sched_of(s) for s in x.shape]
        pl = pacg.Plan(x.shape, schedules=scheds, precision=pacg.PREC_F64)
        pl.set_kernel(k)
        dx = pacg.DeviceComplex.from_numpy(x)
        dy = pl.alloc()
        pl.autotune(dx, dy, 3)
        ms = time_convs(pacg, pl, dx, dy, steps, warmup)
        ms_graph = time_graph(pacg, pl, dx, dy, steps, warmup)
        prof = profile_launches(pacg, pl, dx, dy, max(5, steps // 2))
        b_alg = (2 * d - 1) * 2 * n * 16 + n * 16
        y = dy.numpy()
        ref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(k))
        entry = {"ms_per_conv": ms, "conv_per_s": 1000.0 / ms, "launches": pl.launches_per_convolve(),
                 "alg_bytes": b_alg, "frac": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                 "gflops": 2 * 5 * n * np.log2(n) / (ms * 1e-3) / 1e9,
                 "rel_l2_vs_numpy": float(np.linalg.norm(y - ref) / np.linalg.norm(ref)),
                 "max_abs_vs_numpy": float(np.abs(y - ref).max()), "plan": pl.describe(),
                 "kernels": [{"kernel": nm, "ms": t} for nm, t in prof],
                 "graph_replay_ms_per_conv": ms_graph}
        del pl, dx, dy
        # drop-in entry point, host arrays in and out (the reference's radix: 2 for both)
        re0, im0 = np.ascontiguousarray(x.real.ravel()), np.ascontiguousarray(x.imag.ravel())
        hr0, hi0 = np.ascontiguousarray(k.real.ravel()), np.ascontiguousarray(k.imag.ravel())
        reps = 10
        times = []
        for it in range(reps + 1):
            re, im, hr, hi = re0.copy(), im0.copy(), hr0.copy(), hi0.copy()
            t0 = time.perf_counter()
            if d == 1:
                pacg.permutation_avoiding_convolution(re, im, n, 2, None, None, hr, hi)
            else:
                pacg.permutation_avoiding_convolution_2d(re, im, x.shape[0], x.shape[1], 2, None, None, None, None,
                                                         hr, hi)
            if it:
                times.append(time.perf_counter() - t0)
        # median: a call now and then takes several times longer in the runtime's pageable copies
        entry["dropin_host_arrays_ms"] = 1000.0 * float(np.median(times))
        entry["dropin_host_arrays_ms_mean_min_max"] = [1000.0 * float(np.mean(times)), 1000.0 * float(np.min(times)),
                                                       1000.0 * float(np.max(times))]
        out[name] = entry
    pacg.release()
    return out


def bench_mfma(pacg, synth, steps, warmup):
    """4096 x 256 complex, radix 8-8-8-8 / 8-8-4, radix-8 stages on split-FP16 MFMA (seed 3),
    beside the same plan on the VALU butterflies."""
    x, k = synth.workload_inputs("2d_4096x256")
    n = x.size
    scheds = [[8, 8, 8, 8], [8, 8, 4]]
    ref = np.fft.ifft2(np.fft.fft2(x) * np.fft.fft2(k))
    out = {"workload": "2D 4096x256 complex FP32, U(-1,1) seed 3, schedules 8-8-8-8 / 8-8-4"}
    for name, mf in (("mfma", True), ("valu", False)):
        pl = pacg.Plan(x.shape, schedules=scheds, precision=pacg.PREC_F32, spectrum_precision=pacg.PREC_F32,
                       mfma=mf)
        pl.set_kernel(k)
        dx = pacg.DeviceComplex.from_numpy(x, np.float32)
        dy = pl.alloc()
        pl.autotune(dx, dy, 3)
        ms = time_convs(pacg, pl, dx, dy, steps, warmup)
        prof = profile_launches(pacg, pl, dx, dy, max(5, steps // 4))
        y = dy.numpy()
        b_alg = 3 * 2 * n * 8 + n * 8
        out[name] = {"ms_per_conv": ms, "conv_per_s": 1000.0 / ms, "launches": pl.launches_per_convolve(),
                     "frac_of_B_alg": b_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "gflops": 2 * 5 * n * np.log2(n) / (ms * 1e-3) / 1e9,
                     "rel_l2": float(np.linalg.norm(y - ref) / np.linalg.norm(ref)),
                     "max_abs": float(np.abs(y - ref).max()), "plan": pl.describe(),
                     "kernels": [{"kernel": nm, "ms": t} for nm, t in prof]}
        del pl, dx, dy
    return out


def bench_hbm_probe(pacg):
    """pacg_hbm_probe on 512 MiB buffers: the rates a plain streaming kernel reaches (read, write, copy;
    bytes moved per second, a copy counting its reads and its writes), beside the 8 TB/s peak."""
    nb = 512 << 20
    src = pacg.DeviceArray(nb // 4, np.float32)
    dst = pacg.DeviceArray(nb // 4, np.float32)
    src.upload(np.zeros(nb // 4, np.float32))
    out = {"bytes": nb}
    for name, mode, moved in (("read", 0, nb), ("write", 1, nb), ("copy", 2, 2 * nb)):
        ms = pacg.hbm_probe(mode, dst, src, nb, reps=10)
        out[name + "_GBs"] = moved / (ms * 1e-3) / 1e9
    del src, dst
    return out


# The reference's published per-call seconds (one host core, FP64), from its plot scripts:
# butterflies (ufft in the forward_fft programs / uifft in the backward_fft programs),
# per-element-butterfly-cost-plot.py:17-34, and the digit-reversal permutation timed in the same two
# programs, per-element-permutation-cost-plot.py:17-34.  Keyed by (dims, radix) -> {n: (fwd, bwd)}.
REF_PUBLISHED = {
    "butterflies": {
        ("1d", 2): {2**20: (.0170, .0140), 2**22: (.0780, .0650), 2**24: (.3320, .5065), 2**26: (1.5445, 2.5100),
                    2**28: (6.205, 11.1795)},
        ("1d", 4): {2**20: (.0135, .011), 2**22: (.0535, .0550), 2**24: (.249, .330), 2**26: (1.0815, 1.563),
                    2**28: (4.5890, 6.9505)},
        ("2d", 2): {2**20: (.029, .0355), 2**22: (.2000, .1420), 2**24: (1.0385, 1.1605), 2**26: (4.591, 5.829),
                    2**28: (20.0615, 23.7825)},
        ("2d", 4): {2**20: (.0280, .0340), 2**24: (1.0295, 1.1345), 2**28: (19.8865, 23.5950)},
        ("3d", 2): {2**21: (.0054, .0795), 2**24: (0.7220, 0.8175), 2**27: (6.528, 10.0875)},
        ("3d", 4): {2**24: (0.7035, 0.7785)},
    },
    "permutation": {
        ("1d", 2): {2**20: (.011, .016), 2**22: (.0740, .0930), 2**24: (.3745, .8435), 2**26: (1.687, 4.1915),
                    2**28: (5.5190, 17.3465)},
        ("1d", 4): {2**20: (.011, .0135), 2**22: (.068, .090), 2**24: (.3485, .8315), 2**26: (1.6500, 4.1445),
                    2**28: (5.5645, 17.3075)},
        ("2d", 2): {2**20: (.0015, .002), 2**22: (.006, .007), 2**24: (.0650, .0700), 2**26: (.2960, .3405),
                    2**28: (1.250, 1.4735)},
        ("2d", 4): {2**20: (.001, .0015), 2**24: (.0625, .0650), 2**28: (1.2300, 1.3915)},
        ("3d", 2): {2**21: (.0020, .0020), 2**24: (.0215, .0280), 2**27: (.184, .270)},
        ("3d", 4): {2**24: (.0255, .0295)},
    },
}


def _timed_calls(pacg, stream, reset, call, reps):
    """Mean ms of `reps` calls, each timed alone with HIP events on `stream` after an untimed `reset`
    (the reference re-copies its inputs outside the timed region, bench1d_convolution.c:238-260)."""
    t = pacg.Timer(stream)
    reset()
    call()  # warm-up
    ms = []
    for _ in range(reps):
        reset()
        t.start()
        call()
        ms.append(t.stop())
    return float(np.mean(ms)), float(np.std(ms))


def bench_schedule_search(pacg, synth, steps, warmup):
    """Radix-schedule search (pacg_plan_search) from each configuration's own schedules: the schedules it
    keeps and both plans' rates (tiles autotuned, HIP events, device-resident, FP64 spectrum for FP64)."""
    out = {}
    cases = (("3d_256cubed_f32", "3d_256cubed", SCHEDS32, pacg.PREC_F32, pacg.PREC_F32),
             ("3d_256cubed_f64", "3d_256cubed", SCHEDS64, pacg.PREC_F64, pacg.PREC_F64),
             ("2d_2048x2048_f64", "2d_2048x2048", [[4, 4, 4, 4, 4, 2]] * 2, pacg.PREC_F64, pacg.PREC_F64),
             ("1d_65536_f64", "1d_65536", [[2] * 16], pacg.PREC_F64, pacg.PREC_F64))
    for name, wl, start, prec, sp in cases:
        x, k = synth.workload_inputs(wl)
        t0 = time.time()
        found, changed, fused = pacg.search_schedules(x.shape, schedules=start, precision=prec,
                                                      spectrum_precision=sp, reps=5)
        search_s = time.time() - t0
        dt = np.float32 if prec == pacg.PREC_F32 else np.float64
        dx = pacg.DeviceComplex.from_numpy(x, dt)
        res = {}
        for tag, sc, fa in (("start", start, -1), ("found", found, fused)):
            pl = pacg.Plan(x.shape, schedules=sc, precision=prec, spectrum_precision=sp, fused_axis=fa)
            pl.set_kernel(k)
            dy = pl.alloc()
            pl.autotune(dx, dy, 3)
            res[tag] = {"schedules": sc, "fused_axis": fa, "ms_per_conv": time_convs(pacg, pl, dx, dy, steps, warmup),
                        "plan": pl.describe()}
            del pl, dy
        out[name] = {"axes_changed": changed, "search_s": search_s, **res,
                     "speedup": res["start"]["ms_per_conv"] / res["found"]["ms_per_conv"]}
        del dx
    return out


def _errs(y, ref):
    return {"rel_l2": float(np.linalg.norm(y - ref) / np.linalg.norm(ref)), "max_abs": float(np.abs(y - ref).max())}


def bench_reference_semantics(pacg, x, k, steps):
    """Device-resident convolutions with the reference's own semantics at 256^3: every call transforms
    h as well (h <- A^T h, convolution.c:45-53), so three transforms per call, as in the CPU baseline
    (pacg_convolve_pair), in FP32 and FP64; x and h reset from device copies before each call,
    untimed.  Also the FFT-based convolution (convolution.c:12-29: three ordered transforms, three
    digit-reversal permutations and the product) on the same inputs in FP64."""
    n = x.size
    ref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(k))
    out = {}
    for prec, e, name in ((pacg.PREC_F32, 8, "fp32"), (pacg.PREC_F64, 16, "fp64")):
        dt = np.float32 if prec == pacg.PREC_F32 else np.float64
        pl = pacg.Plan([N1] * 3, schedules=SCHEDS32, precision=prec)
        x0, h0 = pacg.DeviceComplex.from_numpy(x, dt), pacg.DeviceComplex.from_numpy(k, dt)
        dx, dh = pacg.DeviceComplex.from_numpy(x, dt), pacg.DeviceComplex.from_numpy(k, dt)
        nb = n * np.dtype(dt).itemsize

        def reset():
            for dst, src in ((dx, x0), (dh, h0)):
                pacg.lib.pacg_memcpy_d2d(dst.re.ptr, src.re.ptr, nb, pl.stream)
                pacg.lib.pacg_memcpy_d2d(dst.im.ptr, src.im.ptr, nb, pl.stream)

        ms, sd = _timed_calls(pacg, pl.stream, reset, lambda: pl.convolve_pair(dx, dh), max(5, steps // 2))
        y = dx.numpy()
        b = 5 * 2 * n * e + n * e + 3 * 2 * n * e  # x: 5 round trips + spectrum read; h: 3 round trips
        out["pa_" + name] = {"ms_per_conv": ms, "std_ms": sd, "conv_per_s": 1000.0 / ms, "alg_bytes": b,
                             "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "gflops": 3 * 5 * n * np.log2(n) / (ms * 1e-3) / 1e9, **_errs(y, ref)}
        if prec == pacg.PREC_F64:
            ms_f, sd_f = _timed_calls(pacg, pl.stream, reset, lambda: pl.convolve_fft_based(dx, dh),
                                      max(5, steps // 2))
            y = dx.numpy()
            out["fft_based_fp64"] = {"ms_per_conv": ms_f, "std_ms": sd_f, "conv_per_s": 1000.0 / ms_f,
                                     "gflops": 3 * 5 * n * np.log2(n) / (ms_f * 1e-3) / 1e9, **_errs(y, ref),
                                     "fft_based_over_pa": ms_f / ms}
        del pl, x0, h0, dx, dh
    return out


def bench_fft_based_1d(pacg, steps, log2n=24):
    """FFT-based against permutation-avoiding convolution (both with the reference's semantics:
    h transformed in the call) for one 1D FP64 array of 2^24 points, default radix-4 schedule,
    U(-1,1) inputs: what skipping the three permutations is worth on the device."""
    n = 1 << log2n
    g = np.random.default_rng(5)
    x = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
    k = g.uniform(-1, 1, n) + 1j * g.uniform(-1, 1, n)
    ref = np.fft.ifft(np.fft.fft(x) * np.fft.fft(k))
    pl = pacg.Plan([n], precision=pacg.PREC_F64)
    x0, h0 = pacg.DeviceComplex.from_numpy(x), pacg.DeviceComplex.from_numpy(k)
    dx, dh = pacg.DeviceComplex.from_numpy(x), pacg.DeviceComplex.from_numpy(k)
    nb = n * 8

    def reset():
        for dst, src in ((dx, x0), (dh, h0)):
            pacg.lib.pacg_memcpy_d2d(dst.re.ptr, src.re.ptr, nb, pl.stream)
            pacg.lib.pacg_memcpy_d2d(dst.im.ptr, src.im.ptr, nb, pl.stream)

    out = {"n": n, "plan": pl.describe()}
    for name, fn in (("pa", pl.convolve_pair), ("fft_based", pl.convolve_fft_based)):
        ms, sd = _timed_calls(pacg, pl.stream, reset, lambda: fn(dx, dh), max(5, steps // 2))
        out[name] = {"ms_per_conv": ms, "std_ms": sd, "gflops": 3 * 5 * n * log2n / (ms * 1e-3) / 1e9,
                     **_errs(dx.numpy(), ref)}
    out["fft_based_over_pa"] = out["fft_based"]["ms_per_conv"] / out["pa"]["ms_per_conv"]
    del pl, x0, h0, dx, dh
    return out


def bench_per_element(pacg, reps=3):
    """Device counterparts of the reference's only published numbers: FP64 per-call and per-element
    times of ufft (A), uifft (conj(A) / n) and the digit-reversal permutation (gather into another
    array) at the sizes of its plot scripts, pure radix 2 and 4 as there, device-resident, HIP
    events; the reference's seconds beside them (REF_PUBLISHED, one host core)."""
    cases = []
    for r in (2, 4):
        for t in (20, 22, 24, 26, 28):
            cases.append(("1d", r, [1 << t]))
        for t in ((10, 12, 14) if r == 4 else (10, 11, 12, 13, 14)):
            cases.append(("2d", r, [1 << t] * 2))
    cases += [("3d", 2, [1 << 7] * 3), ("3d", 2, [1 << 8] * 3), ("3d", 2, [1 << 9] * 3), ("3d", 4, [1 << 8] * 3)]
    zeros = None
    out = []
    for dims_name, r, dims in cases:
        n = int(np.prod(dims))
        row = {"dims": dims, "radix": r, "n": n}
        try:
            pl = pacg.Plan(dims, radix=r, precision=pacg.PREC_F64)
            if zeros is None or zeros.size < n:
                zeros = np.zeros(n)
            dx = pacg.DeviceComplex(tuple(dims), np.float64)
            dy = pacg.DeviceComplex(tuple(dims), np.float64)
            for d in (dx, dy):
                d.re.upload(zeros[:n])
                d.im.upload(zeros[:n])
            for name, op in (("ufft", pacg.OP_FWD), ("uifft", pacg.OP_UIFFT), ("permutation", pacg.OP_PERM)):
                ms, _ = _timed_calls(pacg, pl.stream, lambda: None, lambda: pl.execute(op, dx, dy), reps)
                row[name + "_s"] = ms * 1e-3
                row[name + "_ns_per_element"] = ms * 1e6 / n
            row["launches"] = pl.launches_per_convolve()
            del pl, dx, dy
        except Exception as exc:  # a shape without compiled kernels is reported, not fatal
            row["error"] = str(exc)[:200]
        pub_b = REF_PUBLISHED["butterflies"].get((dims_name, r), {}).get(n)
        pub_p = REF_PUBLISHED["permutation"].get((dims_name, r), {}).get(n)
        if pub_b:
            row["reference_ufft_s"], row["reference_uifft_s"] = pub_b
            if "ufft_s" in row:
                row["speedup_ufft"] = pub_b[0] / row["ufft_s"]
        if pub_p:
            row["reference_permutation_s"] = pub_p
        out.append(row)
    return {"note": "FP64, device-resident, mean of %d calls after a warm-up; reference_* are the reference's "
                    "published one-core seconds (forward_fft / backward_fft programs)" % reps, "rows": out}


def bench_batched(pacg, synth, steps, warmup):
    """4096 FP32 pairs of lengths 2^U{8..14} (seed 4) in one persistent, graph-captured launch,
    against per-length plain FP64 launches (one batch-axis plan per length)."""
    xs, ks = synth.batched_inputs(4, 4096)
    ms = np.array([int(np.log2(len(x))) for x in xs], dtype=np.int32)
    b = pacg.Batch(ms)
    dx = pacg.DeviceComplex.from_numpy(b.pack(xs), np.float32)
    dk = pacg.DeviceComplex.from_numpy(b.pack(ks), np.float32)
    dy = b.alloc()
    changed = b.autotune(dx, dk, dy, 3)
    b.capture(dx, dk, dy)
    for _ in range(warmup):
        b.replay()
    b.sync()
    t = pacg.Timer(b.stream)
    t.start()
    for _ in range(steps):
        b.replay()
    ms_batch = t.stop() / steps
    t.start()
    for _ in range(steps):
        b.convolve(dx, dk, dy)
    ms_direct = t.stop() / steps
    ys = b.unpack(dy.numpy())
    # per-length plain FP64 launches
    ms64 = 0.0
    num = den = 0.0
    maxerr = 0.0
    for m in range(8, 15):
        idx = np.nonzero(ms == m)[0]
        if idx.size == 0:
            continue
        X = np.stack([xs[i] for i in idx])
        K = np.stack([ks[i] for i in idx])
        pl = pacg.Plan(X.shape, schedules=["batch", None], precision=pacg.PREC_F64)
        pl.set_kernel(K)
        ddx = pacg.DeviceComplex.from_numpy(X)
        ddy = pl.alloc()
        ms64 += time_convs(pacg, pl, ddx, ddy, max(3, steps // 4), 1)
        Y = ddy.numpy()
        for j, i in enumerate(idx):
            d = ys[i] - Y[j]
            num += float(np.vdot(d, d).real)
            den += float(np.vdot(Y[j], Y[j]).real)
            maxerr = max(maxerr, float(np.abs(d).max()))
        del pl, ddx, ddy
    total = int(b.total)
    b_alg = 24 * total
    flops = float(sum(3 * 5 * (1 << int(m)) * int(m) for m in ms))
    return {"workload": "4096 FP32 pairs, lengths 2^U{8..14} (seed 4), one persistent hipGraph-captured launch",
            "points": total, "ms_per_batch": ms_batch, "conv_per_s": 4096 * 1000.0 / ms_batch,
            "alg_bytes": b_alg, "achieved_GBs": b_alg / (ms_batch * 1e-3) / 1e9,
            "frac": b_alg / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "gflops": flops / (ms_batch * 1e-3) / 1e9,
            "variants": {int(m): pacg.Batch.schedule(int(m), b.variant(int(m))) for m in range(8, 15)},
            "autotuned_lengths": changed,
            "dealing": ("longest processing time first", "slices of the shape-ordered items")[b.dealing()],
            "direct_launch_ms_per_batch": ms_direct,
            "direct_launch_frac": b_alg / (ms_direct * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "fp64_per_length_ms": ms64, "speedup_vs_fp64_per_length": ms64 / ms_batch,
            "rel_l2_vs_fp64": (num / den) ** 0.5 if den > 0 else 0.0, "max_abs_vs_fp64": maxerr}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-reps", type=int, default=10,
                    help="timed calls of the CPU baseline after one warm-up (the reference's default -r 10, "
                         "bench1d_convolution.c:74)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes of the CPU throughput figure (0: skip); the box's CPU share is 16 per GPU")
    ap.add_argument("--spectrum", choices=["f32", "f64"], default="f32",
                    help="storage of the FP64-computed spectrum on the headline FP32 path")
    ap.add_argument("--no-tune", action="store_true", help="skip the tile autotuner")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-length workload")
    ap.add_argument("--no-mfma", action="store_true", help="skip the 4096x256 MFMA workload")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the reference-semantics, FFT-based and per-element (ufft / uifft / permutation) sweeps")
    ap.add_argument("--profile-fp64-2d", action="store_true",
                    help="only run the 2D 2048^2 FP64 plan (radix 4, one radix-2 stage per axis), autotuned, "
                         "`steps` times (for rocprofv3)")
    ap.add_argument("--profile-mfma", action="store_true",
                    help="only run the 4096x256 convolution `steps` times with the MFMA plan, then with the VALU "
                         "plan (for rocprofv3 counters)")
    ap.add_argument("--profile-fp64-1d", action="store_true",
                    help="only run the 1D 2^16 FP64 plan (radix 2), autotuned, `steps` times back to back, then "
                         "`steps` empty-kernel launches (for rocprofv3)")
    ap.add_argument("--profile-fp64-3d", action="store_true",
                    help="only run the 256^3 plans whose fused pass streams an FP64 spectrum (FP64 plain, then FP32 "
                         "with the spectrum stored in FP64), autotuned, `steps` times each (for rocprofv3)")
    ap.add_argument("--profile-batched", action="store_true",
                    help="only run the batched workload (graph replay) `steps` times (for rocprofv3)")
    ap.add_argument("--profile", action="store_true",
                    help="only run the FP32 convolution `steps` times after a copy kernel of 2 x 256 MiB "
                         "(counter calibration), for rocprofv3")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus %d)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)

    pacg = load_pacg()
    import pacg.synth as synth

    if a.profile_batched:
        xs, ks = synth.batched_inputs(4, 4096)
        ms_ = np.array([int(np.log2(len(x))) for x in xs], dtype=np.int32)
        b = pacg.Batch(ms_)
        dx = pacg.DeviceComplex.from_numpy(b.pack(xs), np.float32)
        dk = pacg.DeviceComplex.from_numpy(b.pack(ks), np.float32)
        dy = b.alloc()
        b.autotune(dx, dk, dy, 3)  # the schedules the bench line reports
        b.capture(dx, dk, dy)
        for _ in range(a.steps):
            b.replay()
        b.sync()
        print(json.dumps({"profile_batched": True, "steps": a.steps, "points": int(b.total),
                          "dealing": ("longest processing time first", "slices of the shape-ordered items")[b.dealing()],
                          "variants": {m: pacg.Batch.schedule(m, b.variant(m)) for m in range(8, 15)}}))
        return

    if a.profile_fp64_2d:
        x2, k2 = synth.workload_inputs("2d_2048x2048")
        pl = pacg.Plan(x2.shape, schedules=[[4, 4, 4, 4, 4, 2]] * 2, precision=pacg.PREC_F64)
        pl.set_kernel(k2)
        dx = pacg.DeviceComplex.from_numpy(x2)
        dy = pl.alloc()
        pl.autotune(dx, dy, 3)
        pl.sync()
        for _ in range(a.steps):
            pl.convolve(dx, dy)
        pl.sync()
        print(json.dumps({"profile_fp64_2d": True, "steps": a.steps, "plan": pl.describe()}))
        return

    if a.profile_fp64_1d:
        x1, k1 = synth.workload_inputs("1d_65536")
        pl = pacg.Plan(x1.shape, schedules=[[2] * 16], precision=pacg.PREC_F64)
        pl.set_kernel(k1)
        dx = pacg.DeviceComplex.from_numpy(x1)
        dy = pl.alloc()
        pl.autotune(dx, dy, 3)
        pl.sync()
        t = pacg.Timer(pl.stream)
        t.start()
        for _ in range(a.steps):
            pl.convolve(dx, dy)
        ms = t.stop() / a.steps
        floor = pacg.launch_floor(a.steps)
        print(json.dumps({"profile_fp64_1d": True, "steps": a.steps, "plan": pl.describe(), "ms_per_conv": ms,
                          "launch_floor": floor}))
        return

    if a.profile_fp64_3d:
        x3, k3 = synth.workload_inputs("3d_256cubed")
        plans = {}
        for name, prec in (("fp64", pacg.PREC_F64), ("fp32_spectrum_f64", pacg.PREC_F32)):
            pl = pacg.Plan(x3.shape, schedules=SCHEDS64, precision=prec, spectrum_precision=pacg.PREC_F64)
            pl.set_kernel(k3)
            dx = pacg.DeviceComplex.from_numpy(x3, np.float64 if prec == pacg.PREC_F64 else np.float32)
            dy = pl.alloc()
            pl.autotune(dx, dy, 3)  # the tiles the bench line reports
            pl.sync()
            for _ in range(a.steps):
                pl.convolve(dx, dy)
            pl.sync()
            plans[name] = pl.describe()
            del pl, dx, dy
        print(json.dumps({"profile_fp64_3d": True, "steps": a.steps, "plans": plans}))
        return

    if a.profile_mfma:
        x4, k4 = synth.workload_inputs("2d_4096x256")
        dx = pacg.DeviceComplex.from_numpy(x4, np.float32)
        plans = {}
        for name, mf in (("mfma", True), ("valu", False)):  # the MFMA plan, then the VALU plan beside it
            pl = pacg.Plan(x4.shape, schedules=[[8, 8, 8, 8], [8, 8, 4]], precision=pacg.PREC_F32,
                           spectrum_precision=pacg.PREC_F32, mfma=mf)
            pl.set_kernel(k4)
            dy = pl.alloc()
            pl.autotune(dx, dy, 3)  # the tiles the bench line reports
            pl.sync()
            for _ in range(a.steps):
                pl.convolve(dx, dy)
            pl.sync()
            plans[name] = pl.describe()
            del pl, dy
        print(json.dumps({"profile_mfma": True, "steps": a.steps, "plans": plans}))
        return

    t0 = time.time()
    x, k = synth.workload_inputs("3d_256cubed")
    n = x.size
    log("inputs ready (%.1f s)" % (time.time() - t0))
    scheds = SCHEDS32
    sprec = pacg.PREC_F32 if a.spectrum == "f32" else pacg.PREC_F64
    s_bytes = 8 if a.spectrum == "f32" else 16

    # ---- FP32 path (headline) -------------------------------------------------
    pl32 = pacg.Plan([N1] * 3, schedules=scheds, precision=pacg.PREC_F32, spectrum_precision=sprec)
    pl32.set_kernel(k)
    dx32 = pacg.DeviceComplex.from_numpy(x, np.float32)
    dy32 = pl32.alloc()
    tuned32 = 0 if a.no_tune else pl32.autotune(dx32, dy32, 5)
    if a.profile:
        nb = 256 << 20
        src, dst = pacg.DeviceArray(nb // 4, np.float32), pacg.DeviceArray(nb // 4, np.float32)
        src.upload(np.zeros(nb // 4, np.float32))
        for _ in range(3):
            pacg.bw_probe(2, dst, src, nb, pl32.stream)
        for _ in range(a.steps):
            pl32.convolve(dx32, dy32)
        pl32.sync()
        log("profile mode: 3 copies of %d bytes, %d convolutions" % (nb, a.steps))
        print(json.dumps({"profile": True, "copy_bytes": nb, "steps": a.steps, "plan": pl32.describe()}))
        return
    ms32 = time_convs(pacg, pl32, dx32, dy32, a.steps, a.warmup)
    log("fp32: %.3f ms/conv  plan: %s" % (ms32, pl32.describe()))
    dist32 = step_distribution(pacg, pl32, dx32, dy32, a.steps)
    ms32_graph = time_graph(pacg, pl32, dx32, dy32, a.steps, a.warmup)
    prof32 = profile_launches(pacg, pl32, dx32, dy32, max(5, a.steps // 2))
    y32 = dy32.numpy()
    nl32 = len(prof32)
    # in place (y overwrites x, the reference entry points' semantics): x and y share one
    # 128 MiB array, so array + spectrum fit the 256 MiB Infinity Cache
    dxi = pacg.DeviceComplex.from_numpy(x, np.float32)
    ms32_ip = time_convs(pacg, pl32, dxi, dxi, a.steps, a.warmup)
    del dxi

    # ---- FP32 path with the other spectrum storage --------------------------------
    alt = "f64" if a.spectrum == "f32" else "f32"
    pla = pacg.Plan([N1] * 3, schedules=scheds, precision=pacg.PREC_F32,
                    spectrum_precision=pacg.PREC_F64 if alt == "f64" else pacg.PREC_F32)
    pla.set_kernel(k)
    if not a.no_tune:
        pla.autotune(dx32, dy32, 5)
    ms_alt = time_convs(pacg, pla, dx32, dy32, a.steps, a.warmup)
    ya = dy32.numpy()
    del pla

    # ---- plain FP64 path (denominator) ------------------------------------------
    pl64 = pacg.Plan([N1] * 3, schedules=SCHEDS64, precision=pacg.PREC_F64)
    pl64.set_kernel(k)
    dx64 = pacg.DeviceComplex.from_numpy(x, np.float64)
    dy64 = pl64.alloc()
    tuned64 = 0 if a.no_tune else pl64.autotune(dx64, dy64, 3)
    ms64 = time_convs(pacg, pl64, dx64, dy64, max(5, a.steps // 2), a.warmup)
    prof64 = profile_launches(pacg, pl64, dx64, dy64, max(5, a.steps // 4))
    nl64 = len(prof64)
    y64 = dy64.numpy()
    log("fp64: %.3f ms/conv" % ms64)
    del dx64, dy64, pl64

    # ---- errors vs an FP64 reference ---------------------------------------------
    yref = np.fft.ifftn(np.fft.fftn(x) * np.fft.fftn(k))
    nref = np.linalg.norm(yref)
    err32 = float(np.linalg.norm(y32 - yref) / nref)
    max32 = float(np.abs(y32 - yref).max())
    err64 = float(np.linalg.norm(y64 - yref) / nref)
    err_alt = float(np.linalg.norm(ya - yref) / nref)
    log("rel-L2 fp32 %.3e  fp64 %.3e" % (err32, err64))

    # ---- roofline ------------------------------------------------------------
    dom = max(prof32, key=lambda p: p[1])
    dom_bytes = alg_bytes_launch(dom[0], n, 8, s_bytes)
    achieved = dom_bytes / (dom[1] * 1e-3) / 1e9
    # round trips of the array actually made (2d-1 = 5, one axis per launch) plus one spectrum read
    b_alg = conv_alg_bytes(nl32, n, 8, s_bytes)
    b_alg64 = conv_alg_bytes(nl64, n, 16, 16)
    conv_gbs = b_alg / (ms32 * 1e-3) / 1e9
    flops = 2 * 5 * n * np.log2(n)  # two transforms (forward x, inverse), 5 N log2 N each
    traffic, traffic_source = measured_traffic(dom[0])

    out = {
        "metric": "conv/s (3D 256^3 FP32 permutation-avoiding convolution)",
        "value": 1000.0 / ms32,
        "unit": "conv/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms32,
        "step_distribution": dist32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64: x ~ N(0,1), k = separable Gaussian sigma 3)",
        "config": {"workload": "3D 256x256x256 complex PA convolution, radix-8 schedule 4-8-8 per axis, "
                               "FP32 butterflies, spectrum A^T k computed in FP64 on the GPU, stored %s" % a.spectrum,
                   "n": n, "schedules": SCHEDS32, "schedules_fp64": SCHEDS64, "launches_per_conv": len(prof32)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_source,
                     "kernel": dom[0], "kernel_ms": dom[1], "alg_bytes_per_launch": dom_bytes},
        "roofline_conv": {"alg_bytes": b_alg, "launches": nl32, "achieved_GBs": conv_gbs,
                          "frac": conv_gbs / HBM_PEAK_GBS},
        "fp32_graph_replay": {"ms_per_conv": ms32_graph, "conv_per_s": 1000.0 / ms32_graph,
                              "note": "the same 5 launches replayed from one captured HIP graph"},
        "fp32_in_place": {"ms_per_conv": ms32_ip, "conv_per_s": 1000.0 / ms32_ip,
                          "note": "y overwrites x each step (reference semantics); array + spectrum fit the "
                                  "Infinity Cache"},
        "gflops": flops / (ms32 * 1e-3) / 1e9,
        "fp64_plain": {"ms_per_conv": ms64, "conv_per_s": 1000.0 / ms64, "launches": nl64, "alg_bytes": b_alg64,
                       "frac": b_alg64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, "rel_l2": err64},
        "speedup_vs_fp64": ms64 / ms32,
        "fp32_spectrum_%s" % alt: {"ms_per_conv": ms_alt, "conv_per_s": 1000.0 / ms_alt,
                                   "alg_bytes": conv_alg_bytes(nl32, n, 8, 16 if alt == "f64" else 8),
                                   "frac": conv_alg_bytes(nl32, n, 8, 16 if alt == "f64" else 8)
                                   / (ms_alt * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "speedup_vs_fp64": ms64 / ms_alt, "rel_l2": err_alt},
        "autotuned_launches": {"fp32": tuned32, "fp64": tuned64},
        "error_vs_fp64": {"rel_l2": err32, "max_abs": max32, "reference": "numpy.fft complex128"},
        "launches_fp32": [{"kernel": nm, "ms": t} for nm, t in prof32],
        "launches_fp64": [{"kernel": nm, "ms": t} for nm, t in prof64],
    }
    log("fp64 1D 2^16 and 2D 2048^2 ...")
    out.update(bench_fp64_workloads(pacg, synth, a.steps, a.warmup))
    if not a.no_mfma:
        log("mfma workload ...")
        out["mfma_4096x256"] = bench_mfma(pacg, synth, a.steps, a.warmup)
    if not a.no_batched:
        log("batched workload ...")
        out["batched"] = bench_batched(pacg, synth, a.steps, a.warmup)
        log("batched: %.3f ms per batch of 4096" % out["batched"]["ms_per_batch"])
    if not a.no_extra:
        log("reference semantics and FFT-based convolution at 256^3 ...")
        out["reference_semantics_256cubed"] = bench_reference_semantics(pacg, x, k, a.steps)
        log("FFT-based vs PA, 1D 2^24 ...")
        out["fft_based_vs_pa_1d_2e24"] = bench_fft_based_1d(pacg, a.steps)
        log("per-element ufft / uifft / permutation ...")
        out["per_element"] = bench_per_element(pacg)
        log("radix-schedule search ...")
        out["schedule_search"] = bench_schedule_search(pacg, synth, a.steps, a.warmup)
    d_us, g_us = pacg.launch_floor(2000)
    out["launch_floor"] = {"empty_kernel_direct_us": d_us, "empty_graph_replay_us": g_us,
                           "note": "back-to-back launches of an empty kernel, and replays of a HIP graph holding "
                                   "one; the 1D 2^16 line is bounded by these, not by its kernels"}
    log("hbm probe ...")
    probe = bench_hbm_probe(pacg)
    probe["fused_kernel_over_copy"] = out["roofline"]["achieved"] / probe["copy_GBs"]
    out["hbm_probe"] = probe
    if not a.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(a.cpu_reps)
        if a.cpu_procs > 0:
            log("cpu throughput, %d processes ..." % a.cpu_procs)
            out["cpu_baseline_throughput"] = cpu_throughput(min(a.cpu_procs, os.cpu_count() or 1))
        log("cpu baseline, other workloads ...")
        gpu = {"1d_65536": out.get("fp64_1d_65536_r2", {}).get("ms_per_conv"),
               "2d_2048x2048": out.get("fp64_2d_2048x2048_r4", {}).get("ms_per_conv"),
               "2d_4096x256": out.get("mfma_4096x256", {}).get("valu", {}).get("ms_per_conv"),
               "batched_4096": out.get("batched", {}).get("ms_per_batch")}
        out["cpu_baseline_workloads"] = cpu_baseline_workloads(synth, gpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
