"""SPFFT_PU_HOST execution (the library's CPU path) against the oracle, on the
CPU. Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5 against the FP64 oracle;
repeated calls and thread counts bit for bit. The matrix follows the
reference's local test (tests/local_tests/test_local_transform.cpp:95-111):
sizes {1, 2, 11, 12, 13} per dimension plus shapes with 100, centred and
non-centred, C2C backward/forward and R2C forward+backward, with every error
path of transform and grid creation that needs no GPU."""
import ctypes
import glob
import os
import subprocess

import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform

FP64_TOL = 1e-12
FP32_TOL = 1e-5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SMALL = [1, 2, 11, 12, 13]
MATRIX = [(a, b, c) for a in SMALL for b in SMALL for c in SMALL] + [
    (100, 11, 12), (13, 100, 2), (1, 1, 100), (100, 100, 1), (12, 100, 100), (100, 100, 100)]


def host_transform(dims, t, r2c, precision="double", threads=-1):
    cls = sp.Transform if precision == "double" else sp.TransformFloat
    return cls(threads, sp.SPFFT_PU_HOST, sp.SPFFT_TRANS_R2C if r2c else sp.SPFFT_TRANS_C2C,
               *dims, len(t), sp.SPFFT_INDEX_TRIPLETS, t)


def rdt(precision):
    return np.float64 if precision == "double" else np.float32


def run_backward(tr, values, dims, r2c, precision="double"):
    nx, ny, nz = dims
    inp = np.ascontiguousarray(np.asarray(values, np.complex128).view(np.float64)
                               .astype(rdt(precision)))
    out = np.zeros(nx * ny * nz * (1 if r2c else 2), rdt(precision))
    tr.backward(inp, out)
    o = out.astype(np.float64)
    return o.reshape(nz, ny, nx) if r2c else o.view(np.complex128).reshape(nz, ny, nx)


def run_forward(tr, space, r2c, scaling, nv, precision="double"):
    s = np.asarray(space)
    flat = (s.astype(np.float64) if r2c else s.astype(np.complex128).view(np.float64)).ravel()
    out = np.zeros(2 * nv, rdt(precision))
    tr.forward(np.ascontiguousarray(flat.astype(rdt(precision))), out, scaling)
    return out.astype(np.float64).view(np.complex128)


@pytest.mark.parametrize("centered", [False, True])
@pytest.mark.parametrize("dims", MATRIX, ids=lambda d: "x".join(map(str, d)))
def test_host_matrix_c2c(dims, centered):
    t = random_index_set(42, dims, False, centered=centered)
    if len(t) == 0:
        t = np.zeros((1, 3), np.int32)
    v = complex_values(43, len(t))
    o = orc.OraclePlan(False, dims, t)
    tr = host_transform(dims, t, False)
    got = run_backward(tr, v, dims, False)
    assert orc.rel_l2(got, o.backward(v)) <= FP64_TOL
    np.testing.assert_array_equal(run_backward(tr, v, dims, False), got)  # re-zeroed buffers
    nx, ny, nz = dims
    s = complex_values(44, nx * ny * nz).reshape(nz, ny, nx)
    for scaling in (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING):
        f = run_forward(tr, s, False, scaling, len(t))
        assert orc.rel_l2(f, o.forward(s, scaling == sp.SPFFT_FULL_SCALING)) <= FP64_TOL


@pytest.mark.parametrize("centered", [False, True])
@pytest.mark.parametrize("dims", MATRIX, ids=lambda d: "x".join(map(str, d)))
def test_host_matrix_r2c(dims, centered):
    t = random_index_set(45, dims, True, centered=centered)
    if len(t) == 0:
        t = np.zeros((1, 3), np.int32)
    nx, ny, nz = dims
    o = orc.OraclePlan(True, dims, t)
    tr = host_transform(dims, t, True)
    s = uniform(46, nx * ny * nz).reshape(nz, ny, nx)
    f = run_forward(tr, s, True, sp.SPFFT_NO_SCALING, len(t))
    fo = o.forward(s, False)
    assert orc.rel_l2(f, fo) <= FP64_TOL
    assert orc.rel_l2(run_backward(tr, fo, dims, True), o.backward(fo)) <= FP64_TOL
    # non-hermitian half-space input: the symmetry fills decide the result
    v = complex_values(47, len(t))
    assert orc.rel_l2(run_backward(tr, v, dims, True), o.backward(v)) <= FP64_TOL


@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", [32, 96])
def test_host_float_against_fp64_oracle(n, r2c):
    dims = (n, n, n)
    t = sphere_triplets(0.3 * n, half=r2c)
    v = complex_values(48, len(t))
    o = orc.OraclePlan(r2c, dims, t)
    tr = host_transform(dims, t, r2c, "float")
    want = o.backward(v)
    assert orc.rel_l2(run_backward(tr, v, dims, r2c, "float"), want) <= FP32_TOL
    f = run_forward(tr, want, r2c, sp.SPFFT_FULL_SCALING, len(t), "float")
    assert orc.rel_l2(f, o.forward(want, True)) <= FP32_TOL


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_host_golden_fixtures(path):
    g = np.load(path)
    dims = tuple(int(x) for x in g["dims"])
    r2c = bool(g["r2c"])
    tr = host_transform(dims, g["triplets"], r2c)
    b = run_backward(tr, g["values"], dims, r2c)
    flat = b.ravel()
    if "backward" in g:
        assert orc.rel_l2(flat, g["backward"]) <= FP64_TOL
    else:
        assert orc.rel_l2(flat[g["backward_index"]], g["backward_sample"]) <= FP64_TOL
    f = run_forward(tr, b, r2c, sp.SPFFT_FULL_SCALING, len(g["triplets"]))
    assert orc.rel_l2(f, g["forward_scaled"]) <= FP64_TOL


def test_host_thread_counts_agree_bit_for_bit():
    dims = (24, 20, 18)
    t = random_index_set(49, dims, False, centered=True)
    v = complex_values(50, len(t))
    results = [run_backward(host_transform(dims, t, False, threads=k), v, dims, False)
               for k in (1, 3, 8)]
    for r in results[1:]:
        np.testing.assert_array_equal(r, results[0])


def test_host_repeated_triplets_last_wins():
    dims = (12, 12, 12)
    t = sphere_triplets(4)
    t = np.concatenate([t, t[::5]])
    t = t[np.argsort(uniform(51, len(t)), kind="stable")]
    v = complex_values(52, len(t))
    o = orc.OraclePlan(False, dims, t)
    assert orc.rel_l2(run_backward(host_transform(dims, t, False), v, dims, False),
                      o.backward(v)) <= FP64_TOL


def test_host_space_domain_and_locations():
    """backward into the internal space domain, forward from it; a host
    transform has no GPU space domain (reference transform_internal.cpp:341-356)."""
    dims = (16, 12, 10)
    t = random_index_set(53, dims, False)
    v = complex_values(54, len(t))
    tr = host_transform(dims, t, False)
    tr.backward(np.ascontiguousarray(v.view(np.float64)), sp.SPFFT_PU_HOST)
    space = tr.space_domain_data(sp.SPFFT_PU_HOST)
    assert space
    got = np.ctypeslib.as_array(ctypes.cast(space, ctypes.POINTER(ctypes.c_double)),
                                shape=(2 * 16 * 12 * 10,)).copy()
    want = orc.OraclePlan(False, dims, t).backward(v)
    assert orc.rel_l2(got.view(np.complex128).reshape(10, 12, 16), want) <= FP64_TOL
    out = np.zeros(2 * len(t))
    tr.forward(sp.SPFFT_PU_HOST, out, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.view(np.complex128), v) <= FP64_TOL
    for call in (lambda: tr.space_domain_data(sp.SPFFT_PU_GPU),
                 lambda: tr.forward(sp.SPFFT_PU_GPU, out, sp.SPFFT_NO_SCALING),
                 lambda: tr.backward(np.ascontiguousarray(v.view(np.float64)), sp.SPFFT_PU_GPU)):
        with pytest.raises(sp.SpfftError) as e:
            call()
        assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR


def test_host_clone_and_grid_transforms():
    dims = (12, 11, 13)
    t = random_index_set(55, dims, False, centered=True)
    v = complex_values(56, len(t))
    g = sp.Grid(12, 11, 13, 11 * 12, sp.SPFFT_PU_HOST, 2)
    assert g.processing_unit() == sp.SPFFT_PU_HOST and g.num_threads() == 2
    tr = g.create_transform(sp.SPFFT_PU_HOST, sp.SPFFT_TRANS_C2C, *dims, 13, len(t),
                            sp.SPFFT_INDEX_TRIPLETS, t)
    c = tr.clone()
    want = orc.OraclePlan(False, dims, t).backward(v)
    assert orc.rel_l2(run_backward(tr, v, dims, False), want) <= FP64_TOL
    assert orc.rel_l2(run_backward(c, v, dims, False), want) <= FP64_TOL
    assert c.space_domain_data(sp.SPFFT_PU_HOST) != tr.space_domain_data(sp.SPFFT_PU_HOST)
    # a GPU transform needs a grid that includes SPFFT_PU_GPU
    with pytest.raises(sp.SpfftError) as e:
        g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, 13, len(t),
                           sp.SPFFT_INDEX_TRIPLETS, t)
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR


def test_host_multi_transform():
    """multi_transform over host transforms runs them one by one; the
    reference's known answer (values (i, i) -> i N after backward and unscaled
    forward, tests/mpi_tests/test_multi_transform.cpp:52-89)."""
    dims = (12, 13, 11)
    t = random_index_set(57, dims, False, centered=True)
    base = host_transform(dims, t, False)
    trs = [base, base.clone(), base.clone()]
    n = len(t)
    N = 12 * 13 * 11
    ins = [np.ascontiguousarray(((np.arange(n) + k) * (1 + 1j)).view(np.float64)) for k in range(3)]
    spaces = [np.zeros(2 * N) for _ in range(3)]
    outs = [np.zeros(2 * n) for _ in range(3)]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_NO_SCALING] * 3)
    for k in range(3):
        want = (np.arange(n) + k) * (1 + 1j) * N
        got = outs[k].view(np.complex128)
        assert np.max(np.abs(got - want)) <= 1e-9 * np.max(np.abs(want))


def _create(api, dims, t, pu=sp.SPFFT_PU_HOST, ttype=0, n=None, fmt=0, null=False, grid=None,
            zlen=None):
    h = ctypes.c_void_p()
    arr = np.ascontiguousarray(np.asarray(t, np.int32).reshape(-1))
    ptr = None if null else arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    n = len(arr) // 3 if n is None else n
    if grid is None:
        return api["transform_create_independent"](ctypes.byref(h), -1, pu, ttype, *dims, n,
                                                    fmt, ptr)
    return api["transform_create"](ctypes.byref(h), grid, pu, ttype, *dims,
                                   dims[2] if zlen is None else zlen, n, fmt, ptr)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_transform_creation_error_contract(precision):
    """Creation errors of transform_create / transform_create_independent
    (reference transform_internal.cpp:52-77, indices.hpp:120-186,
    parameters.cpp:143-180), on host grids."""
    api = sp._API[precision]
    ok = [[0, 0, 0], [1, 2, 3]]
    assert _create(api, (4, 4, 4), [[4, 0, 0]]) == sp.SPFFT_INVALID_INDICES_ERROR
    assert _create(api, (4, 4, 4), [[-1, 0, 0], [3, 0, 0]]) == sp.SPFFT_INVALID_INDICES_ERROR
    assert _create(api, (4, 4, 4), [[-1, 0, 0]], ttype=1) == sp.SPFFT_INVALID_INDICES_ERROR
    assert _create(api, (4, 4, 4), [[3, 0, 0]], ttype=1) == sp.SPFFT_INVALID_INDICES_ERROR
    assert _create(api, (-1, 4, 4), ok) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create(api, (4, 4, 4), ok, null=True) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create(api, (4, 4, 4), ok, n=-1) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create(api, (1, 1, 1), [[0, 0, 0], [0, 0, 0]]) == sp.SPFFT_INVALID_PARAMETER_ERROR
    assert _create(api, (4, 4, 4), ok, pu=0) == sp.SPFFT_INVALID_PARAMETER_ERROR
    # only index triplets exist (InternalError reports SPFFT_FFTW_ERROR in the reference)
    assert _create(api, (4, 4, 4), ok, fmt=1) == sp.SPFFT_FFTW_ERROR
    g = ctypes.c_void_p()
    assert api["grid_create"](ctypes.byref(g), 8, 8, 8, 4, sp.SPFFT_PU_HOST, 1) == sp.SPFFT_SUCCESS
    try:
        few = [[0, 0, 0], [0, 1, 0], [1, 0, 0], [1, 1, 0], [2, 2, 0]]  # 5 columns > 4
        assert _create(api, (8, 8, 8), few, grid=g) == sp.SPFFT_INVALID_PARAMETER_ERROR
        assert _create(api, (9, 8, 8), ok, grid=g) == sp.SPFFT_INVALID_PARAMETER_ERROR
        assert _create(api, (8, 8, 8), ok, grid=g, zlen=-1) == sp.SPFFT_INVALID_PARAMETER_ERROR
        assert _create(api, (8, 8, 8), ok, grid=g, pu=sp.SPFFT_PU_GPU) == \
            sp.SPFFT_INVALID_PARAMETER_ERROR
        # a transform runs on exactly one processing unit
        assert _create(api, (8, 8, 8), ok, grid=g, pu=sp.SPFFT_PU_HOST | sp.SPFFT_PU_GPU) == \
            sp.SPFFT_INVALID_PARAMETER_ERROR
        h = ctypes.c_void_p()
        arr = np.ascontiguousarray(np.asarray(ok, np.int32).reshape(-1))
        assert api["transform_create"](ctypes.byref(h), g, sp.SPFFT_PU_HOST, 0, 8, 8, 8, 8, 2, 0,
                                       arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == \
            sp.SPFFT_SUCCESS
        v = ctypes.c_int()
        assert api["transform_num_local_elements"](h, ctypes.byref(v)) == sp.SPFFT_SUCCESS
        assert v.value == 8  # the reference's C function reports the z length
        loc = ctypes.c_void_p()
        assert api["transform_get_space_domain"](h, sp.SPFFT_PU_GPU, ctypes.byref(loc)) == \
            sp.SPFFT_INVALID_PARAMETER_ERROR
        assert api["transform_destroy"](h) == sp.SPFFT_SUCCESS
    finally:
        assert api["grid_destroy"](g) == sp.SPFFT_SUCCESS


def test_grid_deep_copy_cpp():
    """spfft::Grid's copy constructor and assignment make deep copies
    (reference src/spfft/grid.cpp:48-53): transforms of a copy run on their own
    buffers. Built against include/ and the library; host grids, so it runs
    without a GPU."""
    src = os.path.join(ROOT, "tests", "examples", "grid_copy.cpp")
    exe = os.path.join(ROOT, "tests", "examples", "bin", "grid_copy")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    libdir = os.path.dirname(sp.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src, "-o",
                    exe, "-L", libdir, "-lspfft"], check=True)
    r = subprocess.run([exe], env=dict(os.environ, LD_LIBRARY_PATH=libdir), capture_output=True,
                       text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
