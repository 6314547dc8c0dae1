"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/, enum values match the reference's headers, the index
maps are bit-exact, and the argument checks that run before any device work
return the reference's error codes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.test_oracle import INDEX_CASES, ref_index_map
from tests.workloads import random_index_set, sphere_triplets, uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for sub in ("spfft", "spfft_amd"):
        d = os.path.join(INCLUDE, sub)
        for f in os.listdir(d):
            if f.endswith(".h"):
                text = open(os.path.join(d, f)).read()
                names.update(re.findall(r"SPFFT_EXPORT\s+\w+\s+(\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) == 76  # 36 double + 36 float + index-map, profiling and no-alloc hooks
    out = subprocess.run(["nm", "-D", "--defined-only", sp.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = names - exported
    assert not missing, missing


def test_enum_values_match_reference_headers():
    # include/spfft/types.h:67-117 and include/spfft/errors.h:33-126 of the reference
    assert (sp.SPFFT_PU_HOST, sp.SPFFT_PU_GPU) == (1, 2)
    assert (sp.SPFFT_TRANS_C2C, sp.SPFFT_TRANS_R2C) == (0, 1)
    assert (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING) == (0, 1)
    assert (sp.SPFFT_EXEC_SYNCHRONOUS, sp.SPFFT_EXEC_ASYNCHRONOUS) == (0, 1)
    assert sp.SPFFT_INDEX_TRIPLETS == 0
    assert sp.SPFFT_INVALID_INDICES_ERROR == 7 and sp.SPFFT_GPU_FFT_ERROR == 22
    # the C header agrees with the Python table
    src = '#include "spfft/spfft.h"\n#include <stdio.h>\nint main(void){printf("%d %d %d %d %d",' \
          ' SPFFT_PU_GPU, SPFFT_TRANS_R2C, SPFFT_FULL_SCALING, SPFFT_EXEC_ASYNCHRONOUS,' \
          ' SPFFT_GPU_FFT_ERROR);return 0;}\n'
    exe = os.path.join(ROOT, "tests", "examples", "bin", "enum_probe")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    subprocess.run(["gcc", "-x", "c", "-I", INCLUDE, "-o", exe, "-"], input=src, text=True,
                   check=True)
    assert subprocess.run([exe], capture_output=True, text=True).stdout == "2 1 1 1 22"


def lib_index_map(r2c, dims, t):
    t = np.ascontiguousarray(np.asarray(t, np.int32).reshape(-1))
    n = t.size // 3
    slots = np.zeros(max(n, 1), np.int32)
    keys = np.zeros(max(n, 1), np.int32)
    ns, zz = ctypes.c_int(), ctypes.c_int()
    ip = ctypes.POINTER(ctypes.c_int)
    f = sp.lib.spfft_amd_convert_index_triplets
    f.argtypes = [ctypes.c_int] * 5 + [ip, ip, ip, ip, ip]
    err = f(int(r2c), *dims, n, t.ctypes.data_as(ip), slots.ctypes.data_as(ip),
            keys.ctypes.data_as(ip), ctypes.byref(ns), ctypes.byref(zz))
    if err:
        return err
    return slots[:n], keys[: ns.value], zz.value


@pytest.mark.parametrize("r2c,dims,centered", INDEX_CASES)
def test_library_index_map_bit_exact(r2c, dims, centered):
    t = random_index_set(7, dims, r2c, centered=centered)
    t = t[np.argsort(uniform(99, len(t)))]
    slots, keys, zz = lib_index_map(r2c, dims, t)
    want = ref_index_map(r2c, dims, t)
    np.testing.assert_array_equal(slots, want[0])
    np.testing.assert_array_equal(keys, want[1])
    o = orc.index_map(r2c, dims, t)
    np.testing.assert_array_equal(slots, o[0])
    assert zz == (0 if len(keys) and keys[0] == 0 else len(keys))


@pytest.mark.parametrize("radius,half", [(12, False), (40, False), (40, True), (30, False)])
def test_library_index_map_bit_exact_spheres(radius, half):
    n = 32 if radius == 12 else (96 if radius == 30 else 128)
    t = sphere_triplets(radius, half=half)
    slots, keys, _ = lib_index_map(half, (n, n, n), t)
    o = orc.index_map(half, (n, n, n), t)
    np.testing.assert_array_equal(slots, o[0])
    np.testing.assert_array_equal(keys, o[1])


def test_library_index_map_errors_match():
    dims = (4, 4, 4)
    for r2c, t in [(False, [[4, 0, 0]]), (False, [[-1, 0, 0], [3, 0, 0]]), (True, [[3, 0, 0]]),
                   (True, [[-1, 0, 0]]), (False, [[-1, -2, 2]])]:
        assert lib_index_map(r2c, dims, t) == 7
    assert lib_index_map(False, (1, 1, 1), [[0, 0, 0], [0, 0, 0]]) == 5


def _create_independent(dims, t, n=None, transform_type=0, pu=sp.SPFFT_PU_GPU, null=False):
    api = sp._API["double"]
    h = ctypes.c_void_p()
    arr = np.ascontiguousarray(np.asarray(t, np.int32).reshape(-1))
    ptr = None if null else arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    n = len(arr) // 3 if n is None else n
    return api["transform_create_independent"](ctypes.byref(h), -1, pu, transform_type, *dims, n,
                                                0, ptr)


def test_error_codes_before_device_work():
    dims = (4, 4, 4)
    assert _create_independent(dims, [[4, 0, 0]]) == sp.SPFFT_INVALID_INDICES_ERROR
    assert _create_independent((-1, 4, 4), [[0, 0, 0]]) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create_independent(dims, [[0, 0, 0]], null=True) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create_independent((1, 1, 1), [[0, 0, 0], [0, 0, 0]]) == \
        sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create_independent(dims, [[0, 0, 0]], n=-1) == sp.SPFFT_INVALID_PARAMETER_ERROR


def test_null_handles():
    api = sp._API["double"]
    v = ctypes.c_int()
    assert api["transform_dim_x"](None, ctypes.byref(v)) == sp.SPFFT_INVALID_HANDLE_ERROR
    assert api["transform_destroy"](None) == sp.SPFFT_INVALID_HANDLE_ERROR
    assert api["grid_destroy"](None) == sp.SPFFT_INVALID_HANDLE_ERROR
    assert api["transform_backward_ptr"](None, None, None) == sp.SPFFT_INVALID_HANDLE_ERROR
    fapi = sp._API["float"]
    assert fapi["grid_max_dim_x"](None, ctypes.byref(v)) == sp.SPFFT_INVALID_HANDLE_ERROR


def test_examples_compile():
    """The C and C++ usage examples build against include/ and link the library."""
    out = os.path.join(ROOT, "tests", "examples", "bin")
    os.makedirs(out, exist_ok=True)
    libdir = os.path.dirname(sp.LIB_PATH)
    ex = os.path.join(ROOT, "tests", "examples")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", INCLUDE, os.path.join(ex, "example.c"),
                    "-o", os.path.join(out, "example_c"), "-L", libdir, "-lspfft", "-lm"],
                   check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-I", INCLUDE, os.path.join(ex, "example.cpp"),
                    "-o", os.path.join(out, "example_cpp"), "-L", libdir, "-lspfft"], check=True)


def _null_handle_args(name, argtypes):
    """Arguments for `name` with a null handle first and valid-looking
    storage for every other parameter."""
    keep = []
    args = [None]
    for t in argtypes[1:]:
        if t is ctypes.c_int:
            args.append(0)
        elif t is ctypes.c_void_p:
            args.append(None)
        else:  # pointer parameters: room for one value
            buf = (ctypes.c_longlong * 4)()
            keep.append(buf)
            args.append(ctypes.cast(buf, t))
    return args, keep


@pytest.mark.parametrize("precision", ["double", "float"])
def test_every_entry_point_rejects_a_null_handle(precision):
    """All 36 functions per precision (reference transform.cpp:178-562,
    grid.cpp:88-290, multi_transform.cpp:84-147): a null grid or transform
    handle gives SPFFT_INVALID_HANDLE_ERROR (the create functions take an
    output handle and a grid instead). A null transform inside a
    multi_transform array is reported as SPFFT_INVALID_PARAMETER_ERROR (the
    reference dereferences it); an empty batch succeeds."""
    api = sp._API[precision]
    assert len(api) == 36
    seen = 0
    for name, f in api.items():
        if name.startswith("multi_transform"):
            n = 1
            handles = (ctypes.c_void_p * 1)(None)
            rest = [(ctypes.c_void_p * 1)(None) if t is not ctypes.POINTER(ctypes.c_int)
                    else (ctypes.c_int * 1)(0) for t in f.argtypes[2:]]
            assert f(n, handles, *rest) == sp.SPFFT_INVALID_PARAMETER_ERROR, name
            assert f(0, handles, *rest) == sp.SPFFT_SUCCESS, name
        elif name == "grid_create":
            continue
        elif name == "transform_create_independent":
            continue
        elif name == "transform_create":
            out = ctypes.c_void_p()
            idx = (ctypes.c_int * 3)(0, 0, 0)
            assert f(ctypes.byref(out), None, sp.SPFFT_PU_HOST, 0, 4, 4, 4, 4, 1, 0, idx) == \
                sp.SPFFT_INVALID_HANDLE_ERROR
        else:
            args, keep = _null_handle_args(name, f.argtypes)
            assert f(*args) == sp.SPFFT_INVALID_HANDLE_ERROR, name
        seen += 1
    assert seen == 34


@pytest.mark.parametrize("precision", ["double", "float"])
def test_grid_create_argument_errors(precision):
    """Grid dimension and processing-unit checks run before any allocation
    (reference grid_internal.cpp:49-67)."""
    api = sp._API[precision]
    h = ctypes.c_void_p()
    bad = [(0, 4, 4, 4, sp.SPFFT_PU_GPU), (4, -1, 4, 4, sp.SPFFT_PU_GPU),
           (4, 4, 0, 4, sp.SPFFT_PU_HOST), (4, 4, 4, -1, sp.SPFFT_PU_HOST),
           (4, 4, 4, 4, 0), (4, 4, 4, 4, 4)]
    for dims in bad:
        assert api["grid_create"](ctypes.byref(h), *dims, -1) == \
            sp.SPFFT_INVALID_PARAMETER_ERROR, dims


def test_noalloc_entry_point_argument_errors():
    """The allocation-free batched entry point (include/spfft_amd/no_alloc.h)
    rejects empty, oversized and null-handle batches and host transforms
    before any device work."""
    f = sp.lib.spfft_amd_multi_transform_noalloc
    vp = ctypes.c_void_p
    one = (vp * 1)(None)
    sc = (ctypes.c_int * 1)(0)
    assert f(0, one, one, one, 0, sc) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert f(1, one, one, one, 0, sc) == sp.SPFFT_INVALID_PARAMETER_ERROR
    many = (vp * 1025)()
    assert f(1025, many, many, many, 1, (ctypes.c_int * 1025)()) == \
        sp.SPFFT_INVALID_PARAMETER_ERROR
    t = sphere_triplets(3)
    host = sp.Transform(1, sp.SPFFT_PU_HOST, sp.SPFFT_TRANS_C2C, 8, 8, 8, len(t),
                        sp.SPFFT_INDEX_TRIPLETS, t)
    buf = np.zeros(2 * 8 ** 3)
    with pytest.raises(sp.SpfftError) as e:
        sp.multi_transform_noalloc([host], [buf], [buf], forward=False)
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR
