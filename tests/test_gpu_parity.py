"""GPU parity: the HIP path through the C ABI against the pinned CPU oracle.

Bars (written per test):
* FP64: relative L2 <= 1e-12 against the oracle (SPFFT_PU_HOST restatement);
  data movement (backward twice, host vs device buffers) bit-exact.
* FP32: relative L2 <= 1e-5 against the FP64 oracle.
Matrix of the reference's local test (tests/local_tests/test_local_transform.cpp:95-111):
sizes {1, 2, 11, 12, 13, 100} per dimension, centred and non-centred,
C2C backward, C2C forward, R2C forward+backward.
"""
import glob
import os

import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform, uniform01

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def make(dims, t, r2c, precision="double"):
    cls = sp.Transform if precision == "double" else sp.TransformFloat
    return cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_R2C if r2c else sp.SPFFT_TRANS_C2C, *dims,
               len(t), sp.SPFFT_INDEX_TRIPLETS, t)


def rdt(precision):
    return np.float64 if precision == "double" else np.float32


def gpu_backward(tr, values, dims, r2c, precision="double"):
    nx, ny, nz = dims
    v = dev(np.asarray(values, np.complex128).view(np.float64).astype(rdt(precision)))
    n = nx * ny * nz * (1 if r2c else 2)
    out = torch.empty(n, dtype=v.dtype, device="cuda")
    tr.backward(v.data_ptr(), out.data_ptr())
    o = out.cpu().numpy().astype(np.float64)
    return o.reshape(nz, ny, nx) if r2c else o.view(np.complex128).reshape(nz, ny, nx)


def gpu_forward(tr, space, r2c, scaling, precision="double"):
    s = np.asarray(space)
    flat = (s.astype(np.float64) if r2c else s.astype(np.complex128).view(np.float64)).ravel()
    d = dev(flat.astype(rdt(precision)))
    out = torch.empty(2 * tr_num_values(tr), dtype=d.dtype, device="cuda")
    tr.forward(d.data_ptr(), out.data_ptr(), scaling)
    return out.cpu().numpy().astype(np.float64).view(np.complex128)


_NV = {}


def tr_num_values(tr):
    return _NV[id(tr)]


def build(dims, t, r2c, precision="double"):
    tr = make(dims, t, r2c, precision)
    _NV[id(tr)] = len(t)
    return tr


SIZES = [1, 2, 11, 12, 13, 100]
MATRIX = [(a, b, c) for a in SIZES for b in SIZES for c in SIZES]


@pytest.mark.parametrize("centered", [False, True])
@pytest.mark.parametrize("dims", MATRIX, ids=lambda d: "x".join(map(str, d)))
def test_reference_matrix_c2c(dims, centered):
    t = random_index_set(42, dims, False, centered=centered)
    if len(t) == 0:
        t = np.zeros((1, 3), np.int32)
    v = complex_values(43, len(t))
    o = orc.OraclePlan(False, dims, t)
    tr = build(dims, t, False)
    got = gpu_backward(tr, v, dims, False)
    want = o.backward(v)
    assert orc.rel_l2(got, want) <= FP64_TOL
    again = gpu_backward(tr, v, dims, False)  # buffers re-zeroed on every call
    np.testing.assert_array_equal(again, got)
    nx, ny, nz = dims
    s = complex_values(44, nx * ny * nz).reshape(nz, ny, nx)
    for scaling in (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING):
        f = gpu_forward(tr, s, False, scaling)
        assert orc.rel_l2(f, o.forward(s, scaling == sp.SPFFT_FULL_SCALING)) <= FP64_TOL


@pytest.mark.parametrize("centered", [False, True])
@pytest.mark.parametrize("dims", MATRIX, ids=lambda d: "x".join(map(str, d)))
def test_reference_matrix_r2c(dims, centered):
    t = random_index_set(45, dims, True, centered=centered)
    if len(t) == 0:
        t = np.zeros((1, 3), np.int32)
    nx, ny, nz = dims
    o = orc.OraclePlan(True, dims, t)
    tr = build(dims, t, True)
    # the reference's R2C test: forward of a real field, then backward
    s = uniform(46, nx * ny * nz).reshape(nz, ny, nx)
    f = gpu_forward(tr, s, True, sp.SPFFT_NO_SCALING)
    fo = o.forward(s, False)
    assert orc.rel_l2(f, fo) <= FP64_TOL
    b = gpu_backward(tr, fo, dims, True)
    assert orc.rel_l2(b, o.backward(fo)) <= FP64_TOL
    # arbitrary (non-hermitian) half-space input: the symmetry fill decides
    v = complex_values(47, len(t))
    assert orc.rel_l2(gpu_backward(tr, v, dims, True), o.backward(v)) <= FP64_TOL


SPHERES = [
    ("sphere_32_r12_c2c", (32, 32, 32), 12, False),
    ("sphere_128_r40_c2c", (128, 128, 128), 40, False),
    ("sphere_128_r40_r2c", (128, 128, 128), 40, True),
    ("sphere_96_r30_c2c", (96, 96, 96), 30, False),
    ("sphere_192_r60_c2c", (192, 192, 192), 60, False),
]


@pytest.mark.parametrize("name,dims,radius,r2c", SPHERES, ids=[s[0] for s in SPHERES])
@pytest.mark.parametrize("precision", ["double", "float"])
def test_sphere_configs(name, dims, radius, r2c, precision):
    tol = FP64_TOL if precision == "double" else FP32_TOL
    t = sphere_triplets(radius, half=r2c)
    v = complex_values(0x5EED0000 + len(t) % 97, len(t))
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c, precision)
    want = o.backward(v)
    got = gpu_backward(tr, v, dims, r2c, precision)
    err_b = orc.rel_l2(got, want)
    f = gpu_forward(tr, want, r2c, sp.SPFFT_FULL_SCALING, precision)
    err_f = orc.rel_l2(f, o.forward(want, True))
    print(f"{name} {precision}: backward rel L2 {err_b:.3e}, forward rel L2 {err_f:.3e}")
    assert err_b <= tol and err_f <= tol


def test_host_pointers_match_device_pointers():
    dims = (12, 11, 13)
    t = random_index_set(61, dims, False, centered=True)
    v = complex_values(62, len(t))
    tr = build(dims, t, False)
    d = gpu_backward(tr, v, dims, False)
    host_out = np.zeros(2 * 12 * 11 * 13)
    tr.backward(np.ascontiguousarray(v.view(np.float64)), host_out)
    np.testing.assert_array_equal(host_out.view(np.complex128).reshape(13, 11, 12), d)
    fwd_host = np.zeros(2 * len(t))
    tr.forward(np.ascontiguousarray(d.view(np.float64).ravel()), fwd_host, sp.SPFFT_FULL_SCALING)
    fwd_dev = gpu_forward(tr, d, False, sp.SPFFT_FULL_SCALING)
    np.testing.assert_array_equal(fwd_host.view(np.complex128), fwd_dev)


def test_space_domain_buffers():
    dims = (16, 12, 10)
    t = random_index_set(63, dims, False)
    v = complex_values(64, len(t))
    tr = build(dims, t, False)
    d_in = dev(v.view(np.float64))
    tr.backward(d_in.data_ptr(), sp.SPFFT_PU_GPU)
    out = torch.empty(2 * len(t), dtype=torch.float64, device="cuda")
    tr.forward(sp.SPFFT_PU_GPU, out.data_ptr(), sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.cpu().numpy().view(np.complex128), v) <= FP64_TOL
    # host mirror
    host = np.zeros(2 * len(t))
    tr.backward(np.ascontiguousarray(v.view(np.float64)), sp.SPFFT_PU_HOST)
    tr.forward(sp.SPFFT_PU_HOST, host, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(host.view(np.complex128), v) <= FP64_TOL


def test_async_mode_orders_default_stream():
    dims = (32, 32, 32)
    t = sphere_triplets(12)
    v = complex_values(65, len(t))
    tr = build(dims, t, False)
    tr.set_execution_mode(sp.SPFFT_EXEC_ASYNCHRONOUS)
    assert tr.execution_mode() == sp.SPFFT_EXEC_ASYNCHRONOUS
    d_in = dev(v.view(np.float64))
    out = torch.empty(2 * 32 ** 3, dtype=torch.float64, device="cuda")
    tr.backward(d_in.data_ptr(), out.data_ptr())
    res = out.cpu().numpy().view(np.complex128).reshape(32, 32, 32)  # default-stream copy
    assert orc.rel_l2(res, orc.OraclePlan(False, dims, t).backward(v)) <= FP64_TOL


def test_multi_transform_known_answer():
    # reference tests/mpi_tests/test_multi_transform.cpp:52-89: values (i, i),
    # backward then unscaled forward gives i * N.
    dims = (12, 13, 11)
    t = random_index_set(66, dims, False, centered=True)
    base = build(dims, t, False)
    trs = [base, base.clone(), base.clone()]
    n = len(t)
    N = 12 * 13 * 11
    ins, outs, spaces = [], [], []
    for k in range(3):
        vals = (np.arange(n) + k) * (1 + 1j)
        ins.append(dev(vals.view(np.float64)))
        spaces.append(torch.empty(2 * N, dtype=torch.float64, device="cuda"))
        outs.append(torch.empty(2 * n, dtype=torch.float64, device="cuda"))
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_NO_SCALING] * 3)
    for k in range(3):
        got = outs[k].cpu().numpy().view(np.complex128)
        want = (np.arange(n) + k) * (1 + 1j) * N
        assert np.max(np.abs(got - want)) <= 1e-8 * max(1.0, np.max(np.abs(want)))
    # locations form
    sp.multi_transform_backward(trs, ins, [sp.SPFFT_PU_GPU] * 3)
    sp.multi_transform_forward(trs, [sp.SPFFT_PU_GPU] * 3, outs, [sp.SPFFT_FULL_SCALING] * 3)
    for k in range(3):
        got = outs[k].cpu().numpy().view(np.complex128)
        assert orc.rel_l2(got, (np.arange(n) + k) * (1 + 1j)) <= FP64_TOL


def test_multi_transform_rejects_shared_grid():
    dims = (8, 8, 8)
    t = random_index_set(67, dims, False)
    g = sp.Grid(8, 8, 8, 64, sp.SPFFT_PU_GPU, -1)
    a = g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 8, 8, 8, 8, len(t),
                           sp.SPFFT_INDEX_TRIPLETS, t)
    b = g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 8, 8, 8, 8, len(t),
                           sp.SPFFT_INDEX_TRIPLETS, t)
    x = torch.zeros(2 * len(t), dtype=torch.float64, device="cuda")
    with pytest.raises(sp.SpfftError) as e:
        sp.multi_transform_backward([a, b], [x, x], [sp.SPFFT_PU_GPU] * 2)
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR


def test_grid_limits_and_accessors():
    g = sp.Grid(16, 12, 10, 50, sp.SPFFT_PU_GPU, 3)
    assert (g.max_dim_x(), g.max_dim_y(), g.max_dim_z()) == (16, 12, 10)
    assert g.max_num_local_z_columns() == 50 and g.num_threads() == 3
    assert g.processing_unit() == sp.SPFFT_PU_GPU
    t = random_index_set(68, (16, 12, 10), False, stick_fraction=0.2)
    tr = g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 16, 12, 10, 10, len(t),
                            sp.SPFFT_INDEX_TRIPLETS, t)
    assert (tr.dim_x(), tr.dim_y(), tr.dim_z(), tr.local_z_length()) == (16, 12, 10, 10)
    assert tr.global_size() == 1920 and tr.num_global_elements() == len(t)
    assert tr.local_slice_size() == 1920 and tr.local_z_offset() == 0
    assert tr.num_local_elements() == 10  # reports the z length, as the reference does
    with pytest.raises(sp.SpfftError) as e:  # larger than the grid
        g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 17, 12, 10, 10, 1,
                           sp.SPFFT_INDEX_TRIPLETS, [[0, 0, 0]])
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR
    many = random_index_set(69, (16, 12, 10), False, stick_fraction=1.0)
    with pytest.raises(sp.SpfftError) as e:  # more columns than the grid allows
        g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 16, 12, 10, 10, len(many),
                           sp.SPFFT_INDEX_TRIPLETS, many)
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR


def test_examples_run():
    import os
    import subprocess
    from tests.test_abi import test_examples_compile
    test_examples_compile()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.dirname(sp.LIB_PATH))
    for exe in ("example_c", "example_cpp"):
        r = subprocess.run([os.path.join(root, "tests", "examples", "bin", exe)], env=env, capture_output=True,
                           text=True, timeout=60)
        print(r.stdout, r.stderr)
        assert r.returncode == 0, exe


FAST_LENGTHS = [32, 64, 96, 128, 192]


@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", FAST_LENGTHS)
@pytest.mark.parametrize("fill", [(0.7, 0.7, False), (0.15, 0.5, True), (1.0, 1.0, True)],
                         ids=["dense70", "sparse_centred", "full"])
def test_fixed_length_path_random_sets(n, fill, precision, r2c):
    """Random index sets on the fixed-length kernels (every column pattern,
    including fully dense planes; a dense 192^3 C2C set has more columns than
    the fixed-length LDS tile holds and runs on the generic kernels). For R2C
    the values are arbitrary on the half space, so the stick and plane
    symmetry fills decide the result."""
    sf, ff, centered = fill
    dims = (n, n, n)
    t = random_index_set(70 + n, dims, r2c, stick_fraction=sf, fill_fraction=ff,
                         centered=centered)
    t = t[np.argsort(uniform(71, len(t)))]  # arbitrary value order
    v = complex_values(72, len(t))
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c, precision)
    tol = FP64_TOL if precision == "double" else FP32_TOL
    want = o.backward(v)
    assert orc.rel_l2(gpu_backward(tr, v, dims, r2c, precision), want) <= tol
    f = gpu_forward(tr, want, r2c, sp.SPFFT_FULL_SCALING, precision)
    assert orc.rel_l2(f, o.forward(want, True)) <= tol


@pytest.mark.parametrize("plane", ["1", "0"], ids=["whole_plane", "class_split"])
@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", [32, 64, 96, 128])
def test_batched_multi_transform_uneven_sets(n, precision, plane, monkeypatch):
    """k-point style batch: every transform has its own shifted sphere (uneven
    sets), one fused launch chain per direction; each result matches the
    oracle and the single-transform call. Run with the whole-plane kernels
    (where the shape qualifies) and with the class-split kernels."""
    monkeypatch.setenv("SPFFT_AMD_PLANE", plane)
    dims = (n, n, n)
    rng_k = uniform(80 + n, 3 * 6).reshape(6, 3) * 0.5
    radii = 0.22 * n + 0.08 * n * uniform01(81 + n, 6)
    trs, sets, vals = [], [], []
    for k in range(6):
        t = sphere_triplets(float(radii[k]), kshift=tuple(rng_k[k]))
        trs.append(build(dims, t, False, precision))
        sets.append(t)
        vals.append(complex_values(90 + k, len(t)))
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [dev(v.view(np.float64).astype(rd)) for v in vals]
    spaces = [torch.empty(2 * n ** 3, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
    tol = FP64_TOL if precision == "double" else FP32_TOL
    for k, t in enumerate(sets):
        o = orc.OraclePlan(False, dims, t)
        want = o.backward(vals[k])
        got = spaces[k].cpu().numpy().astype(np.float64).view(np.complex128).reshape(n, n, n)
        assert orc.rel_l2(got, want) <= tol, k
        single = gpu_backward(trs[k], vals[k], dims, False, precision)
        assert orc.rel_l2(got, single) <= (1e-14 if precision == "double" else 1e-6)
        f = outs[k].cpu().numpy().astype(np.float64).view(np.complex128)
        assert orc.rel_l2(f, o.forward(want, True)) <= tol, k


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", [32, 96, 128])
def test_fixed_length_r2c_symmetry_fill_branches(n, precision):
    """R2C sets holding both halves of the x = 0 plane and the whole (0,0)
    stick, with a third of the values exactly zero, so every branch of the
    symmetry fills is taken (lower wins, upper fills a zero lower, both zero,
    self-paired Nyquist entries)."""
    dims = (n, n, n)
    nxf = n // 2 + 1
    x, y, z = np.meshgrid(np.arange(nxf), np.arange(n), np.arange(n), indexing="ij")
    keep = (uniform01(200 + n, nxf * n * n).reshape(nxf, n, n) < 0.4) | ((x == 0) & (y < 3))
    t = np.ascontiguousarray(np.stack([x[keep], y[keep], z[keep]], axis=1).astype(np.int32))
    v = complex_values(201 + n, len(t))
    v[uniform01(202 + n, len(t)) < 0.33] = 0.0
    o = orc.OraclePlan(True, dims, t)
    tr = build(dims, t, True, precision)
    tol = FP64_TOL if precision == "double" else FP32_TOL
    assert orc.rel_l2(gpu_backward(tr, v, dims, True, precision), o.backward(v)) <= tol


@pytest.mark.parametrize("precision", ["double", "float"])
def test_batched_multi_transform_r2c(precision):
    """A batch of R2C half-sphere transforms with shifted centres on the fused
    launch chain."""
    n = 96
    dims = (n, n, n)
    trs, sets, vals = [], [], []
    for k in range(4):
        t = sphere_triplets(20.0 + 3 * k, half=True)
        trs.append(build(dims, t, True, precision))
        sets.append(t)
        vals.append(complex_values(300 + k, len(t)))
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [dev(v.view(np.float64).astype(rd)) for v in vals]
    spaces = [torch.empty(n ** 3, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
    tol = FP64_TOL if precision == "double" else FP32_TOL
    for k, t in enumerate(sets):
        o = orc.OraclePlan(True, dims, t)
        want = o.backward(vals[k])
        got = spaces[k].cpu().numpy().astype(np.float64).reshape(n, n, n)
        assert orc.rel_l2(got, want) <= tol, k
        f = outs[k].cpu().numpy().astype(np.float64).view(np.complex128)
        assert orc.rel_l2(f, o.forward(want, True)) <= tol, k


@pytest.mark.parametrize("precision", ["double", "float"])
def test_batched_graph_replays(precision, monkeypatch):
    """Repeated multi_transform calls on the same transforms and buffers
    replay a captured HIP graph. Between replays the inputs are rewritten
    with eager work and the default stream is synchronised; every replay must
    equal the direct-launch result bit for bit and match the oracle."""
    n = 64
    dims = (n, n, n)
    sets = [sphere_triplets(r, kshift=(0.1 * k, -0.2, 0.3)) for k, r in enumerate((18.0, 20.0, 21.5))]
    trs = [build(dims, t, False, precision) for t in sets]
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    spaces = [torch.empty(2 * n ** 3, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    tol = FP64_TOL if precision == "double" else FP32_TOL
    for rep in range(4):
        vals = [complex_values(500 + 10 * rep + k, len(t)) for k, t in enumerate(sets)]
        for k, v in enumerate(vals):
            ins[k].copy_(dev(v.view(np.float64).astype(rd)))
        torch.cuda.synchronize()
        results = []
        for graphs in ("1", "0"):
            monkeypatch.setenv("SPFFT_AMD_GRAPHS", graphs)
            sp.multi_transform_backward(trs, ins, spaces)
            sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
            torch.cuda.synchronize()
            results.append(([s.cpu().numpy().copy() for s in spaces], [o.cpu().numpy().copy() for o in outs]))
        for k in range(len(trs)):
            np.testing.assert_array_equal(results[0][0][k], results[1][0][k])
            np.testing.assert_array_equal(results[0][1][k], results[1][1][k])
        for k, t in enumerate(sets):
            o = orc.OraclePlan(False, dims, t)
            want = o.backward(vals[k])
            got = results[0][0][k].astype(np.float64).view(np.complex128).reshape(n, n, n)
            assert orc.rel_l2(got, want) <= tol, (rep, k)
            f = results[0][1][k].astype(np.float64).view(np.complex128)
            assert orc.rel_l2(f, o.forward(want, True)) <= tol, (rep, k)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_batched_whole_plane_many_planes_per_workgroup(precision, monkeypatch):
    """16 uneven 64^3 transforms (1024 planes, more than the FP32 full-tile
    kernels' resident grid of one launch, so their workgroups loop over
    several planes of different transforms and prefetch across them). Two
    rounds with new inputs, with and without graph replay; every result
    matches the oracle and the replays match direct launches bit for bit."""
    monkeypatch.setenv("SPFFT_AMD_PLANE", "1")
    n = 64
    dims = (n, n, n)
    shifts = uniform(700, 16 * 3).reshape(16, 3) * 0.5
    radii = 12.0 + 9.0 * uniform01(701, 16)
    sets = [sphere_triplets(float(radii[k]), kshift=tuple(shifts[k])) for k in range(16)]
    trs = [build(dims, t, False, precision) for t in sets]
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    spaces = [torch.empty(2 * n ** 3, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    tol = FP64_TOL if precision == "double" else FP32_TOL
    for rep in range(2):
        vals = [complex_values(720 + 20 * rep + k, len(t)) for k, t in enumerate(sets)]
        for k, v in enumerate(vals):
            ins[k].copy_(dev(v.view(np.float64).astype(rd)))
        torch.cuda.synchronize()
        results = []
        for graphs in ("1", "0"):
            monkeypatch.setenv("SPFFT_AMD_GRAPHS", graphs)
            sp.multi_transform_backward(trs, ins, spaces)
            sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
            torch.cuda.synchronize()
            results.append(([s.cpu().numpy().copy() for s in spaces], [o.cpu().numpy().copy() for o in outs]))
        for k, t in enumerate(sets):
            np.testing.assert_array_equal(results[0][0][k], results[1][0][k])
            np.testing.assert_array_equal(results[0][1][k], results[1][1][k])
            o = orc.OraclePlan(False, dims, t)
            want = o.backward(vals[k])
            got = results[0][0][k].astype(np.float64).view(np.complex128).reshape(n, n, n)
            assert orc.rel_l2(got, want) <= tol, (rep, k)
            f = results[0][1][k].astype(np.float64).view(np.complex128)
            assert orc.rel_l2(f, o.forward(want, True)) <= tol, (rep, k)


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures(path, precision):
    """The HIP library against the pocketfft fixtures (tests/golden/make_fixtures.py)."""
    g = np.load(path)
    dims = tuple(int(x) for x in g["dims"])
    r2c = bool(g["r2c"])
    tol = FP64_TOL if precision == "double" else FP32_TOL
    tr = build(dims, g["triplets"], r2c, precision)
    b = gpu_backward(tr, g["values"], dims, r2c, precision)
    flat = b.ravel()
    if "backward" in g:
        assert orc.rel_l2(flat, g["backward"]) <= tol
    else:
        assert orc.rel_l2(flat[g["backward_index"]], g["backward_sample"]) <= tol
    f = gpu_forward(tr, b, r2c, sp.SPFFT_FULL_SCALING, precision)
    assert orc.rel_l2(f, g["forward_scaled"]) <= tol


def test_batched_kpoints_fp32_error_per_kpoint():
    """16 k-points on a 96^3 grid, each its own sphere |G + k|^2 <= r^2 with
    k ~ U[-1/2, 1/2)^3 and r ~ U[22, 30] (uneven sets), 2 bands each, values
    N(0,1) exp(-|G|^2/900), FP32 through one multi_transform call per
    direction. The error of every k-point against the FP64 oracle is printed
    and must stay <= 1e-5 (relative L2, backward and forward)."""
    from tests.workloads import normal
    n = 96
    dims = (n, n, n)
    ks = uniform(0x5EED0004, 48).reshape(16, 3) * 0.5
    rs = 22.0 + 8.0 * uniform01(0x5EED0005, 16)
    trs, sets, vals = [], [], []
    for ki in range(16):
        t = sphere_triplets(float(rs[ki]), kshift=tuple(ks[ki]))
        w = np.exp(-np.sum(t.astype(np.float64) ** 2, axis=1) / 900.0)
        for b in range(2):
            seed = 0x7000 + 10 * ki + b
            trs.append(build(dims, t, False, "float"))
            sets.append(t)
            vals.append((normal(seed, len(t)) + 1j * normal(seed + 7919, len(t))) * w)
    ins = [dev(v.view(np.float64).astype(np.float32)) for v in vals]
    spaces = [torch.empty(2 * n ** 3, dtype=torch.float32, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=torch.float32, device="cuda") for t in sets]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
    worst = 0.0
    for ki in range(16):
        errs = []
        for b in range(2):
            k = 2 * ki + b
            o = orc.OraclePlan(False, dims, sets[k])
            want = o.backward(vals[k])
            got = spaces[k].cpu().numpy().astype(np.float64).view(np.complex128).reshape(n, n, n)
            f = outs[k].cpu().numpy().astype(np.float64).view(np.complex128)
            errs += [orc.rel_l2(got, want), orc.rel_l2(f, o.forward(want, True))]
        print(f"k-point {ki:2d}: k = ({ks[ki][0]:+.3f}, {ks[ki][1]:+.3f}, {ks[ki][2]:+.3f}), "
              f"r = {rs[ki]:.2f}, {len(sets[2 * ki])} values: FP32 backward rel L2 "
              f"{max(errs[0::2]):.2e}, forward {max(errs[1::2]):.2e}")
        worst = max(worst, max(errs))
    assert worst <= FP32_TOL
