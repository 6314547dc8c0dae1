"""The split plane stage (y and x kernels around a compact column buffer;
SPFFT_AMD_PLANE=2 forces it) against the FP64 oracle, at every fixed length,
C2C and R2C, FP64 (<= 1e-12 relative L2) and FP32 (<= 1e-5): random column
patterns, spheres, R2C sets that take every branch of the x = 0 column fill,
uneven batches and host pointers (planes in chunks)."""
import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.test_gpu_parity import build, gpu_backward, gpu_forward, rdt
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform, uniform01

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {"double": 1e-12, "float": 1e-5}
LENGTHS = [32, 64, 96, 128, 192]


@pytest.fixture(autouse=True)
def split_stage(monkeypatch):
    monkeypatch.setenv("SPFFT_AMD_PLANE", "2")


def check_single(dims, t, r2c, precision, seed):
    nx, ny, nz = dims
    print(f"split stage: {dims} {'R2C' if r2c else 'C2C'} {precision}, {len(t)} values",
          flush=True)
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c, precision)
    v = complex_values(seed, len(t))
    want = o.backward(v)
    got = gpu_backward(tr, v, dims, r2c, precision)
    assert orc.rel_l2(got, want) <= TOL[precision]
    # forward of a field: real for R2C, arbitrary complex for C2C
    s = uniform(seed + 1, nx * ny * nz).reshape(nz, ny, nx) if r2c else \
        complex_values(seed + 1, nx * ny * nz).reshape(nz, ny, nx)
    for scaling in (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING):
        f = gpu_forward(tr, s, r2c, scaling, precision)
        assert orc.rel_l2(f, o.forward(s, scaling == sp.SPFFT_FULL_SCALING)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", LENGTHS)
def test_split_stage_random_sets(n, r2c, precision):
    dims = (n, n, n)
    t = random_index_set(0x5EED0B00 + n, dims, r2c, stick_fraction=0.5, fill_fraction=0.3,
                         centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0B10 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", [96, 128, 192])
def test_split_stage_spheres(n, r2c, precision):
    dims = (n, n, 64)
    t = sphere_triplets(0.3 * n, half=r2c)
    t = t[np.abs(t[:, 2]) < 32]
    check_single(dims, np.ascontiguousarray(t), r2c, precision, 0x5EED0B20 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", [32, 96, 192])
def test_split_stage_r2c_fill_branches(n, precision):
    """Both halves of the x = 0 plane present, a third of the values zero:
    every branch of the column fill is taken."""
    dims = (n, n, 32)
    nxf = n // 2 + 1
    x, y, z = np.meshgrid(np.arange(nxf), np.arange(n), np.arange(32), indexing="ij")
    keep = (uniform01(0x5EED0B30 + n, nxf * n * 32).reshape(nxf, n, 32) < 0.3) | \
        ((x == 0) & (y < 3))
    t = np.ascontiguousarray(np.stack([x[keep], y[keep], z[keep]], axis=1).astype(np.int32))
    v = complex_values(0x5EED0B31 + n, len(t))
    v[uniform01(0x5EED0B32 + n, len(t)) < 0.33] = 0.0
    o = orc.OraclePlan(True, dims, t)
    tr = build(dims, t, True, precision)
    assert orc.rel_l2(gpu_backward(tr, v, dims, True, precision), o.backward(v)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", [64, 96, 192])
def test_split_stage_uneven_batch(n, r2c, precision):
    """Three transforms with different column counts in one fused chain (the
    batch grid is sized for the widest set)."""
    dims = (n, n, 32)
    sets = [random_index_set(0x5EED0B40 + k, dims, r2c, stick_fraction=f, fill_fraction=0.3)
            for k, f in enumerate((0.2, 0.6, 0.9))]
    print(f"split stage batch: {dims} {'R2C' if r2c else 'C2C'} {precision}", flush=True)
    trs = [build(dims, t, r2c, precision) for t in sets]
    vals = [complex_values(0x5EED0B50 + k, len(t)) for k, t in enumerate(sets)]
    dt = torch.float64 if precision == "double" else torch.float32
    rd = rdt(precision)
    ins = [torch.from_numpy(v.view(np.float64).astype(rd)).cuda() for v in vals]
    size = n * n * 32 * (1 if r2c else 2)
    spaces = [torch.empty(size, dtype=dt, device="cuda") for _ in trs]
    outs = [torch.empty(2 * len(t), dtype=dt, device="cuda") for t in sets]
    sp.multi_transform_backward(trs, ins, spaces)
    sp.multi_transform_forward(trs, spaces, outs, [sp.SPFFT_FULL_SCALING] * len(trs))
    torch.cuda.synchronize()
    for k, t in enumerate(sets):
        o = orc.OraclePlan(r2c, dims, t)
        want = o.backward(vals[k])
        got = spaces[k].cpu().numpy().astype(np.float64)
        got = got.reshape(32, n, n) if r2c else got.view(np.complex128).reshape(32, n, n)
        assert orc.rel_l2(got, want) <= TOL[precision], k
        f = outs[k].cpu().numpy().astype(np.float64).view(np.complex128)
        assert orc.rel_l2(f, o.forward(got, True)) <= TOL[precision], k


def test_split_stage_host_pointers():
    """Pageable host input and output: the plane stage runs in chunks of
    planes (each chunk's y and x kernels) while the copies proceed."""
    n = 128
    dims = (n, n, n)
    t = sphere_triplets(40)
    o = orc.OraclePlan(False, dims, t)
    tr = build(dims, t, False)
    v = complex_values(0x5EED0B60, len(t))
    space = np.empty(2 * n ** 3)
    tr.backward(v.view(np.float64).ctypes.data, space.ctypes.data)
    want = o.backward(v)
    assert orc.rel_l2(space.view(np.complex128).reshape(n, n, n), want) <= 1e-12
    out = np.empty(2 * len(t))
    tr.forward(space.ctypes.data, out.ctypes.data, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.view(np.complex128), o.forward(want, True)) <= 1e-12
