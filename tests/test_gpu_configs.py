"""GPU parity at the batched workload shapes, on index sets that take every z
kernel path, and for the orchestration contracts (host pointers, scaling,
preceding errors).

Bars: FP64 relative L2 <= 1e-12 and FP32 <= 1e-5 against the FP64 oracle;
a band of a batch against the same transform called alone <= 1e-14 (FP64) /
1e-6 (FP32), since batches may use other plane kernels; data movement and
scaling bit for bit.
"""
import ctypes
import os

import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import (complex_values, normal, random_index_set, sphere_triplets, uniform,
                             uniform01)

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5
SELF_TOL = {"double": 1e-14, "float": 1e-6}
TOL = {"double": FP64_TOL, "float": FP32_TOL}
CDT = {"double": torch.complex128, "float": torch.complex64}


def make(dims, t, r2c=False, precision="double"):
    cls = sp.Transform if precision == "double" else sp.TransformFloat
    return cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_R2C if r2c else sp.SPFFT_TRANS_C2C, *dims,
               len(t), sp.SPFFT_INDEX_TRIPLETS, t)


def to_dev(v, precision):
    return torch.from_numpy(np.ascontiguousarray(v.astype(np.complex128))).cuda().to(CDT[precision])


def rel_l2_dev(a, b):
    a = a.to(torch.complex128)
    b = b.to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def ptr(x):
    return torch.view_as_real(x).data_ptr()


def band_values(t, seed):
    """N(0,1) complex values damped by exp(-|G|^2 / 900), as plane-wave codes hold them."""
    w = np.exp(-np.sum(t.astype(np.float64) ** 2, axis=1) / 900.0)
    return (normal(seed, len(t)) + 1j * normal(seed + 7919, len(t))) * w


def run_shared_set_batch(n, radius, bands, precision, oracle_bands):
    """`bands` transforms sharing one sphere, one multi_transform call per
    direction. Every band is compared with the transform called alone; the
    listed bands with the oracle (backward and scaled forward)."""
    dims = (n, n, n)
    t = sphere_triplets(radius)
    trs = [make(dims, t, False, precision) for _ in range(bands)]
    vals = [band_values(t, 0x5EED0300 + 2 * b) for b in range(bands)]
    ins = [to_dev(v, precision) for v in vals]
    spaces = [torch.empty((n, n, n), dtype=CDT[precision], device="cuda") for _ in range(bands)]
    outs = [torch.empty(len(t), dtype=CDT[precision], device="cuda") for _ in range(bands)]
    sp.multi_transform_backward(trs, [ptr(x) for x in ins], [ptr(s) for s in spaces])
    sp.multi_transform_forward(trs, [ptr(s) for s in spaces], [ptr(o) for o in outs],
                               [sp.SPFFT_FULL_SCALING] * bands)
    torch.cuda.synchronize()
    single_s = torch.empty((n, n, n), dtype=CDT[precision], device="cuda")
    single_o = torch.empty(len(t), dtype=CDT[precision], device="cuda")
    worst_self = 0.0
    for b in range(bands):
        trs[b].backward(ptr(ins[b]), ptr(single_s))
        trs[b].forward(ptr(spaces[b]), ptr(single_o), sp.SPFFT_FULL_SCALING)
        torch.cuda.synchronize()
        worst_self = max(worst_self, rel_l2_dev(spaces[b], single_s), rel_l2_dev(outs[b], single_o))
    o = orc.OraclePlan(False, dims, t)
    errs = []
    for b in oracle_bands:
        want = o.backward(vals[b])
        got = spaces[b].cpu().numpy().astype(np.complex128)
        eb = orc.rel_l2(got, want)
        ef = orc.rel_l2(outs[b].cpu().numpy().astype(np.complex128), o.forward(want, True))
        errs.append((b, eb, ef))
    print(f"{n}^3 r={radius} x{bands} {precision}: band vs single {worst_self:.2e}; oracle "
          + ", ".join(f"band {b}: bwd {eb:.2e} fwd {ef:.2e}" for b, eb, ef in errs))
    assert worst_self <= SELF_TOL[precision]
    for b, eb, ef in errs:
        assert eb <= TOL[precision] and ef <= TOL[precision], b


@pytest.mark.parametrize("precision", ["double", "float"])
def test_cfg3_64_bands_shared_sphere_96(precision):
    """64 bands on 96^3 sharing the r=30 sphere (113,081 values each), values
    N(0,1) exp(-|G|^2/900), one multi_transform call per direction."""
    run_shared_set_batch(96, 30, 64, precision, oracle_bands=[0, 21, 42, 63])


@pytest.mark.parametrize("precision", ["double", "float"])
def test_64_bands_shared_sphere_128(precision):
    """The batch behind the batched-FP32 versus plain-FP64 throughput ratio:
    64 bands on 128^3 sharing the r=40 sphere."""
    run_shared_set_batch(128, 40, 64, precision, oracle_bands=[0, 31, 63, 17])


def gsorted_sphere(radius):
    """Sphere triplets in plane-wave order (sorted by |G|^2, ties by index):
    no stick's values form one run in user order."""
    t = sphere_triplets(radius)
    g2 = np.sum(t.astype(np.int64) ** 2, axis=1)
    return np.ascontiguousarray(t[np.lexsort((np.arange(len(t)), g2))])


def repeated_triplets(radius, seed):
    """A sphere where every 7th triplet is repeated later in the list: the last
    copy's value wins on backward, every copy receives the value on forward."""
    t = sphere_triplets(radius)
    rep = t[::7]
    t = np.concatenate([t, rep])
    return np.ascontiguousarray(t[np.argsort(uniform(seed, len(t)), kind="stable")])


@pytest.mark.parametrize("plane", ["1", "0"], ids=["whole_plane", "class_split"])
@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", [64, 96, 128])
def test_batched_entry_table_sets(n, precision, plane, monkeypatch):
    """Batches whose sticks are not runs in user order, so the batched z
    kernels take the entry-table path: a shuffled random set, a sphere sorted
    by |G|, a sphere with repeated triplets (last wins) and a centred sphere
    (runs) in one call."""
    monkeypatch.setenv("SPFFT_AMD_PLANE", plane)
    dims = (n, n, n)
    r = 0.3 * n
    sets = [
        random_index_set(900 + n, dims, False, stick_fraction=0.3, fill_fraction=0.6,
                         centered=True),
        gsorted_sphere(r),
        repeated_triplets(r * 0.9, 901 + n),
        sphere_triplets(r * 0.8),
    ]
    sets[0] = np.ascontiguousarray(sets[0][np.argsort(uniform(902 + n, len(sets[0])))])
    trs = [make(dims, t, False, precision) for t in sets]
    vals = [complex_values(910 + k, len(t)) for k, t in enumerate(sets)]
    ins = [to_dev(v, precision) for v in vals]
    spaces = [torch.empty((n, n, n), dtype=CDT[precision], device="cuda") for _ in sets]
    outs = [torch.empty(len(t), dtype=CDT[precision], device="cuda") for t in sets]
    sp.multi_transform_backward(trs, [ptr(x) for x in ins], [ptr(s) for s in spaces])
    sp.multi_transform_forward(trs, [ptr(s) for s in spaces], [ptr(o) for o in outs],
                               [sp.SPFFT_FULL_SCALING] * len(sets))
    torch.cuda.synchronize()
    for k, t in enumerate(sets):
        o = orc.OraclePlan(False, dims, t)
        want = o.backward(vals[k])
        got = spaces[k].cpu().numpy().astype(np.complex128)
        assert orc.rel_l2(got, want) <= TOL[precision], k
        f = outs[k].cpu().numpy().astype(np.complex128)
        assert orc.rel_l2(f, o.forward(want, True)) <= TOL[precision], k


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n", [32, 128])
def test_host_pointers_against_oracle(n, precision):
    """Pageable host input and output buffers (staged by the library) against
    the oracle, backward and forward."""
    dims = (n, n, n)
    t = sphere_triplets(0.3 * n)
    v = complex_values(950 + n, len(t))
    tr = make(dims, t, False, precision)
    rd = np.float64 if precision == "double" else np.float32
    host_in = np.ascontiguousarray(v.view(np.float64).astype(rd))
    host_space = np.zeros(2 * n ** 3, rd)
    tr.backward(host_in, host_space)
    o = orc.OraclePlan(False, dims, t)
    want = o.backward(v)
    got = host_space.astype(np.float64).view(np.complex128).reshape(n, n, n)
    assert orc.rel_l2(got, want) <= TOL[precision]
    space_in = np.ascontiguousarray(want.view(np.float64).ravel().astype(rd))
    host_out = np.zeros(2 * len(t), rd)
    tr.forward(space_in, host_out, sp.SPFFT_FULL_SCALING)
    f = host_out.astype(np.float64).view(np.complex128)
    assert orc.rel_l2(f, o.forward(want, True)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("n,r2c", [(12, False), (32, False), (128, False), (128, True)])
def test_full_scaling_is_one_multiply(n, r2c, precision):
    """Forward with SPFFT_FULL_SCALING equals the unscaled forward times
    (T)(1.0 / (double)(Nx Ny Nz)), bit for bit (one rounding per component)."""
    dims = (n, n, n)
    t = sphere_triplets(0.3 * n, half=r2c)
    tr = make(dims, t, r2c, precision)
    rd = np.float64 if precision == "double" else np.float32
    count = n ** 3 * (1 if r2c else 2)
    space = torch.from_numpy(uniform(960 + n, count).astype(rd)).cuda()
    a = torch.empty(2 * len(t), dtype=space.dtype, device="cuda")
    b = torch.empty_like(a)
    tr.forward(space.data_ptr(), a.data_ptr(), sp.SPFFT_NO_SCALING)
    tr.forward(space.data_ptr(), b.data_ptr(), sp.SPFFT_FULL_SCALING)
    s = rd(1.0 / float(n ** 3))
    np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy() * s)


def _hip_runtime():
    """The HIP runtime already mapped into this process (the one torch and the
    library share)."""
    for line in open("/proc/self/maps"):
        path = line.split()[-1] if len(line.split()) >= 6 else ""
        if os.path.basename(path).startswith("libamdhip64.so"):
            return ctypes.CDLL(path, mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
    raise RuntimeError("no HIP runtime mapped")


def test_gpu_preceding_error():
    """An error left by an earlier host-side HIP call (an invalid device
    ordinal, no kernel involved) is reported as SPFFT_GPU_PRECEDING_ERROR by
    the next call and then cleared (reference execution_gpu.cpp:256-258)."""
    dims = (32, 32, 32)
    t = sphere_triplets(10)
    tr = make(dims, t)
    x = to_dev(complex_values(970, len(t)), "double")
    space = torch.empty((32, 32, 32), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    hip = _hip_runtime()
    hip.hipSetDevice.argtypes = [ctypes.c_int]
    assert hip.hipSetDevice(100000) != 0
    with pytest.raises(sp.SpfftError) as e:
        tr.backward(ptr(x), ptr(space))
    assert e.value.code == sp.SPFFT_GPU_PRECEDING_ERROR
    tr.backward(ptr(x), ptr(space))  # the error was consumed
    torch.cuda.synchronize()
    want = orc.OraclePlan(False, dims, t).backward(x.cpu().numpy())
    assert orc.rel_l2(space.cpu().numpy(), want) <= FP64_TOL


@pytest.mark.parametrize("precision", ["double", "float"])
def test_host_and_gpu_transforms_on_one_grid(precision):
    """A grid with SPFFT_PU_HOST | SPFFT_PU_GPU (reference
    grid_internal.cpp:64-67) serves a host and a GPU transform; both agree
    with the oracle, and the GPU transform's host mirror is the grid's host
    array."""
    dims = (32, 32, 32)
    t = sphere_triplets(11)
    cls = sp.Grid if precision == "double" else sp.GridFloat
    g = cls(32, 32, 32, 32 * 32, sp.SPFFT_PU_HOST | sp.SPFFT_PU_GPU, 4)
    assert g.processing_unit() == sp.SPFFT_PU_HOST | sp.SPFFT_PU_GPU
    th = g.create_transform(sp.SPFFT_PU_HOST, sp.SPFFT_TRANS_C2C, *dims, 32, len(t),
                            sp.SPFFT_INDEX_TRIPLETS, t)
    tg = g.create_transform(sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, 32, len(t),
                            sp.SPFFT_INDEX_TRIPLETS, t)
    v = complex_values(980, len(t))
    rd = np.float64 if precision == "double" else np.float32
    host_in = np.ascontiguousarray(v.view(np.float64).astype(rd))
    want = orc.OraclePlan(False, dims, t).backward(v)
    out_h = np.zeros(2 * 32 ** 3, rd)
    th.backward(host_in, out_h)
    assert orc.rel_l2(out_h.astype(np.float64).view(np.complex128).reshape(32, 32, 32),
                      want) <= TOL[precision]
    x = to_dev(v, precision)
    space = torch.empty((32, 32, 32), dtype=CDT[precision], device="cuda")
    tg.backward(ptr(x), ptr(space))
    torch.cuda.synchronize()
    assert orc.rel_l2(space.cpu().numpy().astype(np.complex128), want) <= TOL[precision]
    assert tg.space_domain_data(sp.SPFFT_PU_HOST) == th.space_domain_data(sp.SPFFT_PU_HOST)
