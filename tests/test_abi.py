"""The C ABI on the CPU: the library loads, exports every symbol the headers
declare, and its host-side bookkeeping (digit reversal, twiddle tables, swap
permutation, misc helpers, argument validation) matches the reference's
golden fixtures bit for bit.  No GPU compute is called here."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest
from conftest import ROOT, golden_cases, load_golden

HEADERS = [os.path.join(ROOT, "include", h) for h in ("pacg.h", "pacg_fft.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


@pytest.fixture(scope="module")
def L(pacg):
    return pacg.lib


def test_headers_declare_the_reference_interface():
    names = declared_functions(HEADERS[1])
    for must in ("permutation_avoiding_convolution", "permutation_avoiding_convolution_2d",
                 "permutation_avoiding_convolution_3d", "fft_based_convolution_3d", "apply_transposed_butterflies_r8",
                 "apply_conjugate_butterflies_r4_3d", "fft", "ifft_2d", "uifft_3d", "precompute_twiddles_r8",
                 "precompute_index_reversal_permutation_r4_3d", "apply_permutation_split", "get_max_error"):
        assert must in names
    assert len(names) == 66  # every prototype of the reference fft.h:17-231


@pytest.mark.parametrize("hdr", HEADERS)
def test_every_declared_symbol_is_exported(L, hdr):
    missing = []
    for name in declared_functions(hdr):
        try:
            getattr(L, name)
        except AttributeError:
            missing.append(name)
    assert not missing, missing


def _rho_from_lib(L, dims, r):
    nd = len(dims)
    if nd == 1:
        f = getattr(L, "precompute_index_reversal_permutation_r%d" % r)
    else:
        f = getattr(L, "precompute_index_reversal_permutation_r%d_%dd" % (r, nd))
    f.restype = C.POINTER(C.c_int)
    p = f(*[C.c_int(d) for d in dims])
    n = int(np.prod(dims))
    out = np.ctypeslib.as_array(p, shape=(n,)).copy()
    C.CDLL(None).free(p)
    return out


@pytest.mark.parametrize("name", golden_cases())
def test_host_bookkeeping_bit_exact(L, name):
    g = load_golden(name)
    dims = [int(d) for d in g["dims"]]
    r = int(g["radix"])
    if len(dims) == 1 or r in (2, 4):
        assert np.array_equal(_rho_from_lib(L, dims, r), g["rho"])
    # twiddle tables in the reference's jagged layout
    dpp = C.POINTER(C.POINTER(C.c_double))
    wr, wi = dpp(), dpp()
    getattr(L, "precompute_twiddles_r%d" % r)(C.c_int(dims[-1]), C.byref(wr), C.byref(wi))
    t = 0
    m = 1
    while m < dims[-1]:
        m *= r
        t += 1
    re_, im_ = [], []
    for q in range(t):
        cnt = (r - 1) * r ** q
        re_.append(np.ctypeslib.as_array(wr[q], shape=(cnt,)).copy())
        im_.append(np.ctypeslib.as_array(wi[q], shape=(cnt,)).copy())
    L.free_twiddles(wr, wi, C.c_int(t))
    if t:
        assert np.array_equal(np.concatenate(re_), g["tw_re"])
        assert np.array_equal(np.concatenate(im_), g["tw_im"])
    # the swap permutation (index_reversals.c:85-94) on host arrays
    x = g["x"].ravel()
    re = np.ascontiguousarray(x.real)
    im = np.ascontiguousarray(x.imag)
    rho = np.ascontiguousarray(g["rho"], dtype=np.int32)
    L.apply_permutation_split(re.ctypes.data_as(C.POINTER(C.c_double)), im.ctypes.data_as(C.POINTER(C.c_double)),
                              C.c_int(x.size), rho.ctypes.data_as(C.POINTER(C.c_int)))
    assert np.array_equal(re + 1j * im, x[g["rho"]])


def test_reverse_index_known_values(L):
    assert L.reverse_index_r2(6, 3) == 3
    assert L.reverse_index_r4(9, 2) == 6
    assert L.reverse_index_r8(17, 2) == 10
    assert L.reverse_index_r2(1, 5) == 16


def test_misc_helpers(L):
    L.supports_radix_4.restype = C.c_bool
    L.supports_radix_8.restype = C.c_bool
    assert L.supports_radix_4(12) and not L.supports_radix_4(11)
    assert L.supports_radix_8(9) and not L.supports_radix_8(8)
    L.get_max_error.restype = C.c_double
    re = np.array([1.0, 2.0, 0.0])
    im = np.array([1.0, 0.0, 3.0])
    flat = np.array([1.0, 1.0, 2.0, 4.0, 3.0, 3.0])
    dp = C.POINTER(C.c_double)
    assert L.get_max_error(re.ctypes.data_as(dp), im.ctypes.data_as(dp), flat.ctypes.data_as(dp), 3) == 4.0


def test_plan_validation_errors_need_no_gpu(pacg):
    with pytest.raises(pacg.PacgError, match="power of two"):
        pacg.Plan([24])
    with pytest.raises(pacg.PacgError, match="radix"):
        pacg.Plan([64], schedules=[[4, 16]])
    with pytest.raises(pacg.PacgError, match="schedule product"):
        pacg.Plan([64], schedules=[[4, 4]])
    with pytest.raises(pacg.PacgError, match="2\\^31"):
        pacg.Plan([2048, 1024, 1024])
    with pytest.raises(pacg.PacgError):
        pacg.Plan([64], radix=3)


def test_hbm_probe_validation_needs_no_gpu(pacg):
    with pytest.raises(pacg.PacgError, match="mode 3"):
        pacg.hbm_probe(3, None, None, 1 << 20)
    with pytest.raises(pacg.PacgError, match="multiple of 16"):
        pacg.hbm_probe(2, None, None, 24)
    with pytest.raises(pacg.PacgError, match="multiple of 16"):
        pacg.hbm_probe(0, None, None, 0)
    with pytest.raises(pacg.PacgError, match="bad argument"):
        pacg.hbm_probe(2, None, None, 1 << 20)  # copy needs both buffers
    with pytest.raises(pacg.PacgError, match="bad argument"):
        pacg.hbm_probe(1, None, None, 1 << 20, reps=0)


def test_dropin_validation_errors_need_no_gpu(pacg):
    x = np.zeros(24)
    y = np.zeros(24)
    with pytest.raises(pacg.PacgError, match="power of 4"):
        pacg.apply_transposed_butterflies_r4(x, y, 24, None, None)
    z = np.zeros(16)
    with pytest.raises(pacg.PacgError, match="radix 3"):
        pacg.permutation_avoiding_convolution(z, z.copy(), 16, 3, None, None, z.copy(), z.copy())
    bad_rho = np.arange(16, dtype=np.int32)
    bad_rho[3] = 99
    with pytest.raises(pacg.PacgError, match="outside"):
        pacg.fft(z, z.copy(), 16, 2, None, None, bad_rho)
    assert np.all(x == 0) and np.all(z == 0)


def test_product_package_does_not_touch_the_oracle():
    """The shipped package never imports or loads anything under oracle/."""
    pkg = os.path.join(ROOT, "c-permutation-avoiding-convolution_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".c", ".h", ".hip", ".hpp")):
                text = open(os.path.join(dp, f)).read()
                assert "oracle" not in text.lower() or f == "synth.py", f


def test_batch_validation_needs_no_gpu(pacg):
    with pytest.raises(pacg.PacgError, match="outside"):
        pacg.Batch([8, 7, 12])
    with pytest.raises(pacg.PacgError, match="outside"):
        pacg.Batch([15])
    for m in range(8, 15):
        assert pacg.Batch.num_variants(m) >= 1
        for v in range(pacg.Batch.num_variants(m)):
            assert int(np.prod(pacg.Batch.schedule(m, v))) == 1 << m


# ---------------------------------------------------------------------------
# ctypes prototypes.  A function called through ctypes without argtypes gets
# every Python int converted to a 32-bit C int: a 64-bit plan or device
# pointer passed that way arrives truncated, and the library faults on the
# host when it dereferences it.  Every C function the package or bench.py
# calls must carry argtypes (and a restype for non-int results) whose sizes
# match the prototype in include/ or csrc/pacg_internal.h.
# ---------------------------------------------------------------------------
_PROTO_HEADERS = HEADERS + [os.path.join(ROOT, "c-permutation-avoiding-convolution_amd", "csrc", "pacg_internal.h")]
_INT_SIZES = {"int": 4, "unsigned": 4, "uint32_t": 4, "int32_t": 4, "uint64_t": 8, "int64_t": 8, "size_t": 8,
              "long long": 8, "bool": 1, "float": 4, "double": 8}


def _c_type(decl):
    """('ptr' | 'void' | base type name) of one declaration without its parameter name."""
    if "*" in decl:
        return "ptr"
    words = [w for w in decl.replace("const", " ").replace("restrict", " ").split()]
    return " ".join(words)


def prototypes():
    out = {}
    for path in _PROTO_HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        for ret, name, args in re.findall(r"([A-Za-z_][\w \*]*?)\b([a-z_][a-z0-9_]*)\s*\(([^;{}()]*)\)\s*;", text):
            args = args.strip()
            params = []
            if args and args != "void":
                for a in args.split(","):
                    a = a.strip()
                    # drop the parameter name (the last identifier after the type)
                    m = re.match(r"(.*?[\*\s])([A-Za-z_]\w*)$", a)
                    params.append(_c_type(m.group(1) if m else a))
            out[name] = (_c_type(ret.strip()), params)
    return out


def _called_names():
    import ast
    names = set()
    for rel in ("c-permutation-avoiding-convolution_amd/__init__.py", "bench.py"):
        tree = ast.parse(open(os.path.join(ROOT, rel)).read())
        for node in ast.walk(tree):
            if isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name) and node.value.id in ("lib", "L"):
                if node.attr.startswith("pacg_"):
                    names.add(node.attr)
    return names


def _size_ok(ctype, want):
    if want == "ptr":
        return ctype is C.c_void_p or ctype is C.c_char_p or (isinstance(ctype, type) and issubclass(ctype, C._Pointer))
    if want in ("float", "double"):
        return ctype in ((C.c_float,) if want == "float" else (C.c_double,))
    return C.sizeof(ctype) == _INT_SIZES[want] and ctype not in (C.c_float, C.c_double, C.c_void_p)


def binding_problems(protos, names, getfn):
    bad = []
    for name in sorted(names):
        assert name in protos, "no prototype for %s" % name
        ret, params = protos[name]
        fn = getfn(name)
        if params and fn.argtypes is None:
            bad.append("%s: no argtypes" % name)
            continue
        at = list(fn.argtypes or [])
        if len(at) != len(params):
            bad.append("%s: %d argtypes for %d parameters" % (name, len(at), len(params)))
            continue
        for i, (ct, want) in enumerate(zip(at, params)):
            if not _size_ok(ct, want):
                bad.append("%s: parameter %d is %s, binding %s" % (name, i, want, ct))
        if ret == "ptr" and not _size_ok(fn.restype, "ptr"):
            bad.append("%s: pointer result without a pointer restype" % name)
        elif ret in _INT_SIZES and _INT_SIZES[ret] == 8 and (fn.restype is None or C.sizeof(fn.restype) != 8):
            bad.append("%s: 64-bit result with restype %s" % (name, fn.restype))
    return bad


def test_ctypes_bindings_match_the_c_prototypes(pacg):
    protos = prototypes()
    called = _called_names()
    wrapped = [n for n in protos if n in vars(pacg) and callable(vars(pacg)[n])]
    assert {"pacg_plan_create", "pacg_profile_convolve", "pacg_batch_convolve"} <= called
    assert "permutation_avoiding_convolution_3d" in wrapped and len(wrapped) >= 40
    bad = binding_problems(protos, called | set(wrapped), lambda n: getattr(pacg.lib, n))
    assert not bad, bad


def test_prototype_check_catches_a_missing_binding(pacg):
    """A fresh handle of the library has no prototypes: the check reports the
    pointer-truncating calls it would make, and a wrong-width binding."""
    protos = prototypes()
    assert protos["pacg_profile_convolve"][1][0] == "ptr" and len(protos["pacg_profile_convolve"][1]) == 9
    fresh = C.CDLL(pacg.LIB_PATH)
    bad = binding_problems(protos, {"pacg_profile_convolve", "pacg_plan_size"}, lambda n: getattr(fresh, n))
    assert any(b.startswith("pacg_profile_convolve: no argtypes") for b in bad)
    assert any(b.startswith("pacg_plan_size: no argtypes") for b in bad)
    fresh.pacg_plan_size.argtypes = [C.c_int]  # a pointer parameter bound as int
    fresh.pacg_plan_size.restype = C.c_longlong
    bad = binding_problems(protos, {"pacg_plan_size"}, lambda n: getattr(fresh, n))
    assert bad == ["pacg_plan_size: parameter 0 is ptr, binding <class 'ctypes.c_int'>"]
