"""The allocation-free batched entry point (include/spfft_amd/no_alloc.h)
against the multi_transform entry points it varies: same transforms, same
inputs, outputs bit for bit (both take the plane kernels tuned by the first
batched call), on two calls in a row, the second with a smaller batch on a
smaller grid. Also the size limit of its fixed device scratch."""
import numpy as np
import pytest

import spfft_amd as sp
from tests.workloads import normal, sphere_triplets

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def make_batch(n, radius, bands, seed):
    t = sphere_triplets(radius)
    trs = [sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, n, n, n, len(t),
                        sp.SPFFT_INDEX_TRIPLETS, t) for _ in range(bands)]
    vals = [torch.from_numpy(normal(seed + 2 * b, len(t)) + 1j * normal(seed + 2 * b + 1, len(t)))
            .to("cuda") for b in range(bands)]
    return t, trs, vals


def ptr(x):
    return torch.view_as_real(x).data_ptr()


def rel(a, b):
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def run_pair(n, radius, bands, seed):
    t, trs, vals = make_batch(n, radius, bands, seed)
    print(f"no-alloc: {n}^3 r={radius} x{bands} ({len(t)} values per band)", flush=True)
    space = lambda: [torch.empty((n, n, n), dtype=torch.complex128, device="cuda")  # noqa: E731
                     for _ in range(bands)]
    outs = lambda: [torch.empty(len(t), dtype=torch.complex128, device="cuda")  # noqa: E731
                    for _ in range(bands)]
    ref_s, got_s = space(), space()
    ref_o, got_o = outs(), outs()
    scal = [sp.SPFFT_FULL_SCALING if b % 2 == 0 else sp.SPFFT_NO_SCALING for b in range(bands)]
    sp.multi_transform_backward(trs, [ptr(v) for v in vals], [ptr(s) for s in ref_s])
    sp.multi_transform_forward(trs, [ptr(s) for s in ref_s], [ptr(o) for o in ref_o], scal)
    torch.cuda.synchronize()
    sp.multi_transform_noalloc(trs, [ptr(v) for v in vals], [ptr(s) for s in got_s], False)
    sp.multi_transform_noalloc(trs, [ptr(s) for s in got_s], [ptr(o) for o in got_o], True, scal)
    torch.cuda.synchronize()
    for b in range(bands):
        if bands > 1:
            # both calls run the same fused launch chain
            assert torch.equal(got_s[b], ref_s[b]), b
            assert torch.equal(got_o[b], ref_o[b]), b
        else:
            # a single transform through multi_transform takes the plain path
            assert rel(got_s[b], ref_s[b]) <= 1e-14, b
            assert rel(got_o[b], ref_o[b]) <= 1e-14, b
    # the round trip returns the input (scaled bands) or N^3 times it
    for b in range(bands):
        want = vals[b] if b % 2 == 0 else vals[b] * n ** 3
        assert rel(got_o[b], want) <= 1e-12, b


def test_noalloc_matches_multi_transform_two_calls():
    run_pair(128, 40, 6, 0x5EED0700)
    run_pair(64, 20, 3, 0x5EED0710)


def test_noalloc_single_transform():
    run_pair(96, 30, 1, 0x5EED0720)


def test_noalloc_split_stage_lengths():
    """Lengths served by the split plane stage only: a mixed radix schedule
    (120), then lines whose waves have spare lanes (90), smaller."""
    run_pair(120, 37, 4, 0x5EED0740)
    run_pair(90, 28, 3, 0x5EED0750)


def test_noalloc_scratch_limit():
    """64 x 64 x 128 transforms with all 4,096 sticks (one value each): 512
    blocks of 8 sticks, so 224 + 4,096 bytes of tables per transform. 240 fit
    the 1 MiB scratch (1,036,800 bytes) and match multi_transform; 250
    (1,080,000 bytes) are rejected before any launch."""
    x, y = np.meshgrid(np.arange(64), np.arange(64), indexing="ij")
    t = np.ascontiguousarray(np.stack([x.ravel(), y.ravel(), np.zeros(64 * 64, dtype=np.int64)],
                                      axis=1).astype(np.int32))
    trs = [sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, 64, 64, 128, len(t),
                        sp.SPFFT_INDEX_TRIPLETS, t) for _ in range(250)]
    vals = [torch.from_numpy(normal(0x5EED0730 + b, len(t)) + 0j).to("cuda") for b in range(250)]
    spaces = [torch.empty((128, 64, 64), dtype=torch.complex128, device="cuda") for _ in range(250)]
    with pytest.raises(sp.SpfftError) as e:
        sp.multi_transform_noalloc(trs, [ptr(v) for v in vals], [ptr(s) for s in spaces], False)
    assert e.value.code == sp.SPFFT_INVALID_PARAMETER_ERROR
    ref = [torch.empty_like(s) for s in spaces[:240]]
    sp.multi_transform_backward(trs[:240], [ptr(v) for v in vals[:240]], [ptr(s) for s in ref])
    sp.multi_transform_noalloc(trs[:240], [ptr(v) for v in vals[:240]],
                               [ptr(s) for s in spaces[:240]], False)
    torch.cuda.synchronize()
    for b in range(240):
        assert torch.equal(spaces[b], ref[b]), b
