"""The second and third groups of fixed lengths: the remaining 2^a 3^b 5^c 7^d
lengths up to 512 whose lines fit a wavefront (36, 42, 70, 140, 210 and 270 to
512 even; 45, 63, 75, 81, 105, 125, 135, 147, 175, 189, 225 and 243 odd), and
the even lengths with one factor 11 or 13 (44 to 462, radix-11 and radix-13
passes), and odd multiples of 11 and 13, 11^2, 13^2 and their doubles (99 to
351), on the z kernels and the split plane stage, against the FP64 oracle:
FP64 <= 1e-12, FP32 <= 1e-5 relative L2. Lengths up to 243 run on cubic grids (spheres and
random column patterns); the longer ones in each of the three dimensions of
grids whose other two lengths are small, and once cubic. R2C grids with an
odd row count leave the last packed row pair of the x kernels with one row.
Also grids mixing the groups, uneven batches, host pointers and a check that
the fixed-length kernels ran."""
import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.test_gpu_parity import build, gpu_backward, rdt
from tests.test_gpu_split_lengths import TOL, check_single, kernels_run
from tests.workloads import complex_values, random_index_set, sphere_triplets

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SHORT = [36, 42, 45, 63, 70, 75, 81, 105, 125, 135, 140, 147, 175, 189, 210, 225, 243]
LONG = [270, 280, 288, 300, 320, 336, 360, 384, 400, 420, 432, 448, 450, 480, 500, 512]
# the two short lengths beside each long one (odd and even, all groups)
PARTNERS = [(45, 36), (42, 63), (75, 40), (81, 70), (36, 105), (125, 48), (63, 42), (45, 56),
            (135, 36), (40, 75), (42, 81), (147, 50), (36, 45), (63, 60), (70, 45), (81, 36),
            (44, 45), (52, 36), (66, 75), (78, 40), (88, 63), (104, 45), (110, 36), (132, 81),
            (99, 36), (45, 117), (121, 40), (63, 165), (169, 42), (195, 36), (45, 231), (99, 50)]
ELEVEN = [44, 52, 66, 78, 88, 104, 110, 130, 132, 154, 156, 176, 182, 198, 208, 220, 234, 260]
ELEVEN_LONG = [264, 312, 330, 352, 390, 416, 440, 462]
ODD11 = [99, 117, 121, 165, 169, 195, 231]
ODD11_LONG = [242, 273, 275, 286, 297, 325, 338, 351]
MIXED = [(45, 36, 512), (512, 63, 42), (75, 480, 81), (125, 135, 140), (243, 128, 45),
         (96, 175, 189), (105, 210, 225), (384, 45, 64), (63, 320, 147), (36, 42, 500),
         (44, 52, 66), (132, 128, 78), (462, 45, 104), (52, 416, 75), (88, 110, 440),
         (99, 121, 169), (351, 45, 99), (117, 286, 64)]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", SHORT)
def test_more_lengths_sphere(n, r2c, precision):
    t = sphere_triplets(0.3 * n, half=r2c)
    check_single((n, n, n), t, r2c, precision, 0x5EED0E00 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", SHORT)
def test_more_lengths_random_columns(n, r2c, precision):
    """Half the columns, 30% of each filled: empty rows and columns, partial
    sticks and the zero column all occur."""
    dims = (n, n, n)
    t = random_index_set(0x5EED0E20 + n, dims, r2c, stick_fraction=0.5, fill_fraction=0.3,
                         centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0E30 + n)


def long_dims(n, axis):
    a, b = PARTNERS[(LONG + ELEVEN_LONG + ODD11_LONG).index(n)]
    d = [a, b]
    d.insert(axis, n)
    return tuple(d)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["x", "y", "z"])
@pytest.mark.parametrize("n", LONG)
def test_more_lengths_long(n, axis, r2c, precision):
    dims = long_dims(n, axis)
    t = random_index_set(0x5EED0E40 + n + axis, dims, r2c, stick_fraction=0.5,
                         fill_fraction=0.4, centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0E50 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", ELEVEN)
def test_eleven_thirteen_sphere(n, r2c, precision):
    t = sphere_triplets(0.3 * n, half=r2c)
    check_single((n, n, n), t, r2c, precision, 0x5EED0F00 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", ELEVEN)
def test_eleven_thirteen_random_columns(n, r2c, precision):
    dims = (n, n, n)
    t = random_index_set(0x5EED0F20 + n, dims, r2c, stick_fraction=0.5, fill_fraction=0.3,
                         centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0F30 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["x", "y", "z"])
@pytest.mark.parametrize("n", ELEVEN_LONG)
def test_eleven_thirteen_long(n, axis, r2c, precision):
    dims = long_dims(n, axis)
    t = random_index_set(0x5EED0F40 + n + axis, dims, r2c, stick_fraction=0.5,
                         fill_fraction=0.4, centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0F50 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", ODD11)
def test_odd_eleven_thirteen_sphere(n, r2c, precision):
    t = sphere_triplets(0.3 * n, half=r2c)
    check_single((n, n, n), t, r2c, precision, 0x5EED1000 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", ODD11)
def test_odd_eleven_thirteen_random_columns(n, r2c, precision):
    dims = (n, n, n)
    t = random_index_set(0x5EED1020 + n, dims, r2c, stick_fraction=0.5, fill_fraction=0.3,
                         centered=True)
    check_single(dims, t, r2c, precision, 0x5EED1030 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("axis", [0, 1, 2], ids=["x", "y", "z"])
@pytest.mark.parametrize("n", ODD11_LONG)
def test_odd_eleven_thirteen_long(n, axis, r2c, precision):
    dims = long_dims(n, axis)
    t = random_index_set(0x5EED1040 + n + axis, dims, r2c, stick_fraction=0.5,
                         fill_fraction=0.4, centered=True)
    check_single(dims, t, r2c, precision, 0x5EED1050 + n)


@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("n", [270, 320, 384, 264, 351])
def test_more_lengths_long_cubic(n, r2c):
    t = sphere_triplets(0.25 * n, half=r2c)
    check_single((n, n, n), t, r2c, "double", 0x5EED0E60 + n)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("dims", MIXED, ids=lambda d: "x".join(map(str, d)))
def test_more_lengths_mixed_grids(dims, r2c, precision):
    t = random_index_set(0x5EED0E70 + sum(dims), dims, r2c, stick_fraction=0.4,
                         fill_fraction=0.5, centered=True)
    check_single(dims, t, r2c, precision, 0x5EED0E80 + dims[0])


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("r2c", [False, True], ids=["c2c", "r2c"])
@pytest.mark.parametrize("dims", [(45, 63, 75), (36, 512, 42), (225, 105, 81), (66, 78, 88),
                                  (99, 117, 121)],
                         ids=lambda d: "x".join(map(str, d)))
def test_more_lengths_uneven_batch(dims, r2c, precision):
    """Three transforms of different column counts as one fused chain."""
    nx, ny, nz = dims
    sets = [random_index_set(0x5EED0E90 + k, dims, r2c, stick_fraction=f, fill_fraction=0.25)
            for k, f in enumerate((0.1, 0.6, 0.95))]
    trs = [build(dims, t, r2c, precision) for t in sets]
    vals = [complex_values(0x5EED0EA0 + k, len(t)) for k, t in enumerate(sets)]
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


@pytest.mark.parametrize("dims,r2c", [((105, 105, 105), True), ((320, 36, 45), False)],
                         ids=["105-r2c", "320x36x45-c2c"])
def test_more_lengths_host_pointers(dims, r2c):
    """Host buffers: values staged in the work buffer, planes in chunks."""
    nx, ny, nz = dims
    t = random_index_set(0x5EED0EB0 + nx, dims, r2c, stick_fraction=0.5, fill_fraction=0.4,
                         centered=True)
    o = orc.OraclePlan(r2c, dims, t)
    tr = build(dims, t, r2c)
    v = complex_values(0x5EED0EC0 + nx, len(t))
    space = np.empty(nx * ny * nz * (1 if r2c else 2))
    tr.backward(v.view(np.float64).ctypes.data, space.ctypes.data)
    want = o.backward(v)
    got = space.reshape(nz, ny, nx) if r2c else space.view(np.complex128).reshape(nz, ny, nx)
    assert orc.rel_l2(got, want) <= 1e-12
    out = np.empty(2 * len(t))
    tr.forward(space.ctypes.data, out.ctypes.data, sp.SPFFT_FULL_SCALING)
    assert orc.rel_l2(out.view(np.complex128), o.forward(got, True)) <= 1e-12


@pytest.mark.parametrize("n", SHORT + LONG + ELEVEN + ELEVEN_LONG + ODD11 + ODD11_LONG)
def test_more_lengths_use_fixed_length_kernels(n):
    dims = (n, n, n) if n in SHORT + ELEVEN + ODD11 else long_dims(n, 2)
    t = random_index_set(0x5EED0ED0 + n, dims, False, stick_fraction=0.3, fill_fraction=0.3)
    tr = build(dims, t, False)
    v = complex_values(0x5EED0EE0, len(t))
    ran = kernels_run(lambda: gpu_backward(tr, v, dims, False))
    assert {"k_z_backward_fast", "k_ysplit_backward", "k_xsplit_backward"} <= ran, ran
