"""Grids with Nx != Ny (each in 32, 64, 96, 128, 192) on the fixed-length
kernels: z kernels and the split plane stage (y kernels of length Ny, x
kernels of length Nx), chosen without any forcing. Against the FP64 oracle:
FP64 <= 1e-12, FP32 <= 1e-5 relative L2; single transforms (backward, forward
scaled and unscaled), uneven batches, host pointers."""
import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.test_gpu_parity import build, gpu_backward, gpu_forward, rdt
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL = {"double": 1e-12, "float": 1e-5}
SHAPES = [(64, 128, 96), (128, 64, 32), (192, 96, 64), (32, 192, 128), (96, 32, 192)]


def ellipsoid(dims, r2c):
    """All (x, y, z) with (x/ax)^2 + (y/ay)^2 + (z/az)^2 <= 1, axes 0.3 of each
    length (centred indices; R2C: the hermitian half)."""
    nx, ny, nz = dims
    ax, ay, az = 0.3 * nx, 0.3 * ny, 0.3 * nz
    t = sphere_triplets(1.0 * max(ax, ay, az), half=r2c)
    keep = ((t[:, 0] / ax) ** 2 + (t[:, 1] / ay) ** 2 + (t[:, 2] / az) ** 2) <= 1.0
    return np.ascontiguousarray(t[keep])


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("dims", SHAPES, ids=lambda d: "x".join(map(str, d)))
@pytest.mark.parametrize("kind", ["random", "ellipsoid"])
def test_nonsquare_single(kind, dims, r2c, precision):
    nx, ny, nz = dims
    if kind == "random":
        t = random_index_set(0x5EED0C00 + nx + ny, dims, r2c, stick_fraction=0.5,
                             fill_fraction=0.3, centered=True)
    else:
        t = ellipsoid(dims, r2c)
    print(f"non-square {dims} {kind} {'R2C' if r2c else 'C2C'} {precision}: {len(t)} values",
          flush=True)
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c, precision)
    v = complex_values(0x5EED0C10 + nz, len(t))
    want = o.backward(v)
    assert orc.rel_l2(gpu_backward(tr, v, dims, r2c, precision), want) <= TOL[precision]
    s = uniform(0x5EED0C20, nx * ny * nz).reshape(nz, ny, nx) if r2c else \
        complex_values(0x5EED0C20, nx * ny * nz).reshape(nz, ny, nx)
    for scaling in (sp.SPFFT_NO_SCALING, sp.SPFFT_FULL_SCALING):
        f = gpu_forward(tr, s, r2c, scaling, precision)
        assert orc.rel_l2(f, o.forward(s, scaling == sp.SPFFT_FULL_SCALING)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
def test_nonsquare_uneven_batch(r2c, precision):
    dims = (96, 64, 128)
    nx, ny, nz = dims
    sets = [random_index_set(0x5EED0C30 + k, dims, r2c, stick_fraction=f, fill_fraction=0.2)
            for k, f in enumerate((0.2, 0.5, 0.9))]
    trs = [build(dims, t, r2c, precision) for t in sets]
    vals = [complex_values(0x5EED0C40 + k, len(t)) for k, t in enumerate(sets)]
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


def test_nonsquare_host_pointers():
    dims = (128, 96, 64)
    nx, ny, nz = dims
    t = ellipsoid(dims, False)
    o = orc.OraclePlan(False, dims, t)
    tr = build(dims, t, False)
    v = complex_values(0x5EED0C50, len(t))
    space = np.empty(2 * nx * ny * nz)
    tr.backward(v.view(np.float64).ctypes.data, space.ctypes.data)
    want = o.backward(v)
    assert orc.rel_l2(space.view(np.complex128).reshape(nz, ny, nx), want) <= 1e-12
    out = np.empty(2 * len(t))
    tr.forward(space.ctypes.data, out.ctypes.data, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.view(np.complex128), o.forward(want, True)) <= 1e-12
