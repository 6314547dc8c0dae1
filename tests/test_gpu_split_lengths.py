"""Lengths served by the z kernels and the split plane stage only (40, 48, 72,
80, 144, 160, 256; on mixed radix schedules 60, 100, 108, 120, 200, 216, 240;
with lines whose lane count does not divide the wave, so that waves have
spare lanes, 50, 54, 90, 150, 162 and 180; with radix-7 passes 56, 84, 98, 112,
126, 168, 196, 224 and 252: no class-split kernels), chosen
without forcing, against the FP64 oracle: FP64
<= 1e-12, FP32 <= 1e-5 relative L2. Cubic grids with
spheres and random column patterns, C2C and R2C, grids mixing these lengths
with each other and with the class-split lengths, uneven batches (one fused
chain per direction), host pointers, and a check that the fixed-length
kernels are the ones that ran."""
import ctypes

import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.test_gpu_parity import build, gpu_backward, gpu_forward, rdt
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {"double": 1e-12, "float": 1e-5}
LENGTHS = [40, 48, 50, 54, 56, 60, 72, 80, 84, 90, 98, 100, 108, 112, 120, 126, 144, 150, 160,
           162, 168, 180, 196, 200, 216, 224, 240, 252, 256]
MIXED = [(72, 80, 48), (160, 144, 256), (256, 48, 96), (80, 128, 160), (192, 72, 144),
         (48, 256, 32), (120, 100, 60), (216, 108, 120), (100, 216, 64), (90, 180, 96),
         (180, 60, 90), (240, 50, 162), (54, 200, 150), (40, 54, 240), (112, 84, 126),
         (196, 56, 98), (224, 252, 168)]


def check_single(dims, t, r2c, precision, seed):
    nx, ny, nz = dims
    print(f"split-only length: {dims} {'R2C' if r2c else 'C2C'} {precision}, {len(t)} values",
          flush=True)
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c, precision)
    v = complex_values(seed, len(t))
    want = o.backward(v)
    assert orc.rel_l2(gpu_backward(tr, v, dims, r2c, precision), want) <= TOL[precision]
    s = uniform(seed + 1, nx * ny * nz).reshape(nz, ny, nx) if r2c else \
        complex_values(seed + 1, nx * ny * nz).reshape(nz, ny, nx)
    for scaling in (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING):
        f = gpu_forward(tr, s, r2c, scaling, precision)
        assert orc.rel_l2(f, o.forward(s, scaling == sp.SPFFT_FULL_SCALING)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", LENGTHS)
def test_split_length_sphere(n, r2c, precision):
    dims = (n, n, n)
    t = sphere_triplets(0.3 * n, half=r2c)
    check_single(dims, t, r2c, precision, 0x5EED0D00 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", LENGTHS)
def test_split_length_random_columns(n, r2c, precision):
    """Random stick pattern (half the columns, 30% of each filled), so rows and
    columns with no sticks, partial sticks and the zero column all occur."""
    dims = (n, n, n)
    t = random_index_set(0x5EED0D20 + n, dims, r2c, stick_fraction=0.5, fill_fraction=0.3,
                         centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0D30 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("dims", MIXED, ids=lambda d: "x".join(map(str, d)))
def test_split_length_mixed_grids(dims, r2c, precision):
    t = random_index_set(0x5EED0D40 + sum(dims), dims, r2c, stick_fraction=0.4,
                         fill_fraction=0.5, centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0D50 + dims[0])


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("dims", [(144, 144, 144), (80, 72, 160), (120, 108, 100),
                                  (90, 180, 90)],
                         ids=lambda d: "x".join(map(str, d)))
def test_split_length_uneven_batch(dims, r2c, precision):
    nx, ny, nz = dims
    sets = [random_index_set(0x5EED0D60 + k, dims, r2c, stick_fraction=f, fill_fraction=0.25)
            for k, f in enumerate((0.1, 0.6, 0.95))]
    trs = [build(dims, t, r2c, precision) for t in sets]
    vals = [complex_values(0x5EED0D70 + k, len(t)) for k, t in enumerate(sets)]
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [torch.from_numpy(v.view(np.float64).astype(rd)).cuda() for v in vals]
    size = nx * ny * nz * (1 if r2c else 2)
    spaces = [torch.empty(size, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
    torch.cuda.synchronize()
    for k, t in enumerate(sets):
        o = orc.OraclePlan(r2c, dims, t)
        want = o.backward(vals[k])
        got = spaces[k].cpu().numpy().astype(np.float64)
        got = got.reshape(nz, ny, nx) if r2c else got.view(np.complex128).reshape(nz, ny, nx)
        assert orc.rel_l2(got, want) <= TOL[precision], k
        f = outs[k].cpu().numpy().astype(np.float64).view(np.complex128)
        assert orc.rel_l2(f, o.forward(got, True)) <= TOL[precision], k


@pytest.mark.parametrize("n,r2c", [(160, False), (90, True), (216, False)],
                         ids=["160-c2c", "90-r2c", "216-c2c"])
def test_split_length_host_pointers(n, r2c):
    """Host buffers: values staged in the work buffer, planes in chunks."""
    dims = (n, n, n)
    nx, ny, nz = dims
    t = sphere_triplets(0.3 * n, half=r2c)
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c)
    v = complex_values(0x5EED0D80 + n, len(t))
    space = np.empty(nx * ny * nz * (1 if r2c else 2))
    tr.backward(v.view(np.float64).ctypes.data, space.ctypes.data)
    want = o.backward(v)
    got = space.reshape(nz, ny, nx) if r2c else space.view(np.complex128).reshape(nz, ny, nx)
    assert orc.rel_l2(got, want) <= 1e-12
    out = np.empty(2 * len(t))
    tr.forward(space.ctypes.data, out.ctypes.data, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.view(np.complex128), o.forward(got, True)) <= 1e-12


def kernels_run(fn):
    """Names of the kernels that fn launched (per-kernel event timing)."""
    lib = sp.lib
    lib.spfft_amd_profiling_enable(1)
    try:
        fn()
        torch.cuda.synchronize()
        names = ctypes.create_string_buffer(64 * 64)
        counts = (ctypes.c_longlong * 64)()
        total = (ctypes.c_double * 64)()
        n = lib.spfft_amd_profiling_report(64, names, counts, total)
    finally:
        lib.spfft_amd_profiling_enable(0)
    return {names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode() for k in range(n)}


@pytest.mark.parametrize("n", LENGTHS)
def test_split_length_uses_fixed_length_kernels(n):
    dims = (n, n, n)
    t = sphere_triplets(0.25 * n)
    tr = build(dims, t, False)
    v = complex_values(0x5EED0D90, len(t))
    ran = kernels_run(lambda: gpu_backward(tr, v, dims, False))
    assert {"k_z_backward_fast", "k_ysplit_backward", "k_xsplit_backward"} <= ran, ran
