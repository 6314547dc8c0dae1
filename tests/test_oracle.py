"""Pins the CPU oracle (oracle/spfft_oracle.c) before anything is compared
with it. Each check is independent of the oracle's own formulas:

* 1D FFT          -> numpy pocketfft, O(N^2) long-double DFT, closed forms
* index maps      -> a dict/sorted restatement of the reference's std::map
                     walk (src/compression/indices.hpp:120-186)
* 3D C2C          -> dense numpy ifftn/fftn of the scattered grid, brute-force
                     plane-wave sums on tiny grids, delta -> plane wave
* 3D R2C          -> hermitian completion + dense inverse DFT; "lower half
                     wins" on non-hermitian input (symmetry_host.hpp:47-90)
* scaling         -> one IEEE multiply per component, bit-exact
The reference ships no golden vectors (tests generate data at run time and
compare with FFTW 3D at 1e-6 absolute, tests/test_util/test_check_values.hpp);
the bounds here are tighter.
"""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.workloads import complex_values, random_index_set, sphere_triplets, uniform


def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    # exact integer phase reduction, long double angle
    e = (np.outer(k, k) % n).astype(np.longdouble)
    ang = sign * 2 * np.pi * e / n
    w = np.cos(ang) + 1j * np.sin(ang)
    return (w.astype(np.complex128) @ x.astype(np.complex128))


SIZES = list(range(1, 34)) + [48, 60, 64, 96, 97, 100, 101, 125, 128, 143, 169, 192, 210, 256, 343]


@pytest.mark.parametrize("n", SIZES)
def test_fft1d_matches_numpy(n):
    x = complex_values(1000 + n, n)
    for sign in (-1, +1):
        got = orc.fft1d(x, sign)
        want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
        assert orc.rel_l2(got, want) < 2e-15, (n, sign)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 11, 12, 13, 16, 17, 30, 31])
def test_fft1d_matches_brute_force(n):
    x = complex_values(2000 + n, n)
    for sign in (-1, +1):
        assert orc.rel_l2(orc.fft1d(x, sign), brute_dft(x, sign)) < 2e-15


@pytest.mark.parametrize("n", [4, 12, 13, 100, 128])
def test_fft1d_closed_forms(n):
    # delta at m -> exp(sign 2 pi i k m / n); constant -> n * delta_0
    m = 3 % n
    d = np.zeros(n, complex)
    d[m] = 1.0
    k = np.arange(n)
    got = orc.fft1d(d, -1)
    want = np.exp(-2j * np.pi * ((k * m) % n) / n)
    assert np.max(np.abs(got - want)) < 1e-15
    c = orc.fft1d(np.ones(n, complex), +1)
    want = np.zeros(n, complex)
    want[0] = n
    assert np.max(np.abs(c - want)) < 1e-12


def ref_index_map(r2c, dims, t):
    """Reference semantics with Python containers (indices.hpp:120-186)."""
    nx, ny, nz = dims
    t = np.asarray(t).reshape(-1, 3)
    if len(t) > nx * ny * nz:
        return 5
    centered = bool((t < 0).any())
    maxX = (nx // 2 + 1 if (r2c or centered) else nx) - 1
    maxY = (ny // 2 + 1 if centered else ny) - 1
    maxZ = (nz // 2 + 1 if centered else nz) - 1
    minX = 0 if r2c else maxX - nx + 1
    minY, minZ = maxY - ny + 1, maxZ - nz + 1
    for x, y, z in t:
        if not (minX <= x <= maxX and minY <= y <= maxY and minZ <= z <= maxZ):
            return 7
    st = lambda d, i: d + i if i < 0 else i  # noqa: E731
    table = {}
    for x, y, z in t:
        table[st(nx, x) * ny + st(ny, y)] = 0
    keys = sorted(table)
    for i, k in enumerate(keys):
        table[k] = i
    slots = [table[st(nx, x) * ny + st(ny, y)] * nz + st(nz, z) for x, y, z in t]
    return np.array(slots, np.int32), np.array(keys, np.int32)


INDEX_CASES = [
    (False, (11, 12, 13), False), (False, (11, 12, 13), True), (True, (11, 12, 13), False),
    (True, (12, 12, 12), True), (False, (1, 1, 1), False), (False, (2, 2, 2), True),
    (True, (2, 1, 5), False), (False, (16, 3, 7), True),
]


@pytest.mark.parametrize("r2c,dims,centered", INDEX_CASES)
def test_index_map_matches_reference_semantics(r2c, dims, centered):
    t = random_index_set(7, dims, r2c, centered=centered)
    # shuffle the value order: the map must follow creation order
    perm = np.argsort(uniform(99, len(t)))
    t = t[perm]
    want = ref_index_map(r2c, dims, t)
    got = orc.index_map(r2c, dims, t)
    assert not isinstance(want, int)
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])


def test_index_map_errors():
    dims = (4, 4, 4)
    assert orc.index_map(False, dims, [[4, 0, 0]]) == 7       # x beyond N-1
    assert orc.index_map(False, dims, [[-1, 0, 0], [3, 0, 0]]) == 7  # centred max is N/2
    assert orc.index_map(True, dims, [[3, 0, 0]]) == 7        # R2C x <= N/2
    assert orc.index_map(True, dims, [[-1, 0, 0]]) == 7       # R2C x >= 0
    assert orc.index_map(False, (1, 1, 1), [[0, 0, 0], [0, 0, 0]]) == 5  # more values than grid
    assert not isinstance(orc.index_map(False, dims, [[-1, 2, 2]]), int)  # centred range [-1, 2]
    assert orc.index_map(False, dims, [[-1, -2, 2]]) == 7
    assert ref_index_map(False, dims, [[-1, 0, 0], [3, 0, 0]]) == 7


C2C_CASES = [((11, 12, 13), False), ((11, 12, 13), True), ((12, 12, 12), True), ((1, 1, 1), False),
             ((2, 2, 2), True), ((13, 1, 100), False), ((1, 2, 11), True), ((32, 16, 24), False)]


@pytest.mark.parametrize("dims,centered", C2C_CASES)
def test_c2c_backward_matches_dense_definition(dims, centered):
    t = random_index_set(11, dims, False, centered=centered)
    v = complex_values(12, len(t))
    p = orc.OraclePlan(False, dims, t)
    got = p.backward(v)
    want = orc.dense_backward_c2c(dims, t, v)
    assert orc.rel_l2(got, want) < 1e-14


@pytest.mark.parametrize("dims,centered", C2C_CASES)
def test_c2c_forward_matches_dense_definition(dims, centered):
    t = random_index_set(13, dims, False, centered=centered)
    nx, ny, nz = dims
    s = complex_values(14, nx * ny * nz).reshape(nz, ny, nx)
    p = orc.OraclePlan(False, dims, t)
    for scaling in (False, True):
        got = p.forward(s, scaling)
        want = orc.dense_forward(dims, t, s, scaling)
        assert orc.rel_l2(got, want) < 1e-14


def test_c2c_matches_brute_force_plane_waves():
    dims = (5, 4, 3)
    t = random_index_set(21, dims, False, centered=True)
    v = complex_values(22, len(t))
    got = orc.OraclePlan(False, dims, t).backward(v)
    want = orc.brute_dft3(dims, t, v, +1)
    assert orc.rel_l2(got, want) < 1e-14


def test_delta_gives_plane_wave():
    dims = (12, 10, 8)
    t = np.array([[-3, 2, 1]], np.int32)
    got = orc.OraclePlan(False, dims, t).backward(np.array([1.0 + 0j]))
    z, y, x = np.meshgrid(np.arange(8), np.arange(10), np.arange(12), indexing="ij")
    want = np.exp(2j * np.pi * (-3 * x / 12 + 2 * y / 10 + 1 * z / 8))
    assert np.max(np.abs(got - want)) < 1e-14


def test_sphere_32_r12_round_trip_and_definition():
    dims = (32, 32, 32)
    t = sphere_triplets(12)
    v = complex_values(0x5EED0001, len(t))
    p = orc.OraclePlan(False, dims, t)
    space = p.backward(v)
    assert orc.rel_l2(space, orc.dense_backward_c2c(dims, t, v)) < 1e-14
    back = p.forward(space, True)
    assert orc.rel_l2(back, v) < 1e-14


def test_scaling_is_one_multiply():
    dims = (12, 10, 8)
    t = random_index_set(31, dims, False)
    s = complex_values(32, 12 * 10 * 8)
    p = orc.OraclePlan(False, dims, t)
    a = p.forward(s, False)
    b = p.forward(s, True)
    scale = np.float64(1.0) / np.float64(12 * 10 * 8)
    np.testing.assert_array_equal(b.real, a.real * scale)
    np.testing.assert_array_equal(b.imag, a.imag * scale)


def hermitian_pairs(dims, seed):
    """R2C half-space set with hermitian-consistent values on the x = 0 plane."""
    nx, ny, nz = dims
    t = random_index_set(seed, dims, True, stick_fraction=1.0, fill_fraction=1.0)
    v = complex_values(seed + 1, len(t))
    return t, v


@pytest.mark.parametrize("dims", [(8, 8, 8), (11, 12, 13), (12, 11, 10), (1, 6, 5), (6, 1, 1),
                                  (2, 2, 2), (1, 1, 4)])
def test_r2c_backward_matches_hermitian_completion(dims):
    # full half space: the symmetry fill completes x = 0 from y in [1, ny/2]
    # and the (0,0) stick from z in [1, nz/2]; compare with the completion of
    # exactly those values.
    t, v = hermitian_pairs(dims, 41)
    nx, ny, nz = dims
    keep = np.ones(len(t), bool)
    for i, (x, y, z) in enumerate(t):
        if x == 0 and y == 0 and z > nz // 2:
            keep[i] = False
    p = orc.OraclePlan(True, dims, t[keep])
    got = p.backward(v[keep])
    want = orc.dense_backward_r2c_hermitian(dims, t[keep], v[keep])
    assert orc.rel_l2(got, want) < 1e-13


@pytest.mark.parametrize("dims", [(8, 8, 8), (11, 12, 13), (6, 1, 1), (1, 1, 4)])
def test_r2c_forward_matches_dense_rfft(dims):
    nx, ny, nz = dims
    t = random_index_set(51, dims, True)
    s = uniform(52, nx * ny * nz).reshape(nz, ny, nx)
    p = orc.OraclePlan(True, dims, t)
    got = p.forward(s, True)
    want = orc.dense_forward(dims, t, s, True)
    assert orc.rel_l2(got, want) < 1e-14


def test_r2c_lower_half_wins_on_non_hermitian_input():
    dims = (8, 8, 8)
    # (0,0) stick: z=+2 and z=-2 both given, inconsistent. The reference keeps
    # the lower storage index (z=2) and overwrites z=-2 with its conjugate.
    t = np.array([[0, 0, 2], [0, 0, -2], [0, 3, 1], [0, -3, 1], [1, 1, 1]], np.int32)
    v = np.array([1 + 2j, 5 - 7j, 0.5 + 0.25j, -3 + 1j, 0.3 - 0.1j])
    got = orc.OraclePlan(True, dims, t).backward(v)
    # x=0 plane: the y=+3 stick wins over y=-3 plane by plane (after the z
    # transform), i.e. the y=-3 entry is replaced by the mirror of y=+3.
    tk = np.array([[0, 0, 2], [0, 3, 1], [1, 1, 1]], np.int32)
    vk = np.array([1 + 2j, 0.5 + 0.25j, 0.3 - 0.1j])
    want = orc.dense_backward_r2c_hermitian(dims, tk, vk)
    assert orc.rel_l2(got, want) < 1e-14


def test_r2c_upper_half_fills_missing_lower():
    dims = (8, 8, 8)
    # only the upper (negative) partners given: they fill the lower ones
    t = np.array([[0, 0, -2], [0, -3, 1], [2, 1, 0]], np.int32)
    v = np.array([1 + 2j, 0.5 + 0.25j, 0.3 - 0.1j])
    got = orc.OraclePlan(True, dims, t).backward(v)
    want = orc.dense_backward_r2c_hermitian(dims, t, v)
    assert orc.rel_l2(got, want) < 1e-14


GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_golden_fixtures(path):
    """The C oracle against the pocketfft fixtures (tests/golden/make_fixtures.py)."""
    g = np.load(path)
    dims = tuple(int(x) for x in g["dims"])
    r2c = bool(g["r2c"])
    o = orc.OraclePlan(r2c, dims, g["triplets"])
    b = o.backward(g["values"]).ravel()
    if "backward" in g:
        assert orc.rel_l2(b, g["backward"]) <= 1e-13
    else:
        assert orc.rel_l2(b[g["backward_index"]], g["backward_sample"]) <= 1e-13
        assert abs(np.linalg.norm(b) - float(g["backward_norm"])) <= 1e-12 * float(g["backward_norm"])
    space = b.reshape(dims[2], dims[1], dims[0])
    assert orc.rel_l2(o.forward(space, True), g["forward_scaled"]) <= 1e-13


@pytest.mark.parametrize("dims,r2c", [((11, 12, 13), False), ((12, 11, 13), True),
                                      ((32, 32, 32), False), ((16, 1, 9), True)])
def test_fftw_api_provider_matches_builtin(dims, r2c):
    """The CPU baseline may route the oracle's 1D lines through a library
    exporting the FFTW3 C API (FFTW or MKL): same results as the built-in FFT
    to rounding, backward and forward."""
    t = random_index_set(11, dims, r2c, centered=True)
    v = complex_values(12, len(t))
    orc.use_builtin_fft()
    p0 = orc.OraclePlan(r2c, dims, t)
    if orc.use_fft_library() is None:
        pytest.skip("no library with the FFTW3 C API on this host")
    try:
        p1 = orc.OraclePlan(r2c, dims, t)
        w0, w1 = p0.backward(v), p1.backward(v)
        assert orc.rel_l2(w1, w0) <= 1e-14
        assert orc.rel_l2(p1.forward(w0, True), p0.forward(w0, True)) <= 1e-14
    finally:
        orc.use_builtin_fft()
