"""Gradient of L = sum V |psi|^2, psi = backward(c), through the adjoint kernel
chain (the unnormalised forward on the sparse set) behind
torch.autograd.Function.

Checks (FP64): central finite differences on 64 random coefficients
(real and imaginary parts separately, step 1e-2: the loss is quadratic, so
the central difference is exact up to rounding of order eps*L/h; agreement
1e-6 relative to the gradient's RMS), and the full gradient against the FP64 oracle gradient
2 F_unscaled(V psi) computed on the CPU (relative L2 <= 1e-12). FP32: the same
oracle gradient at relative L2 <= 1e-5. The recompute and store variants agree
bit for bit.
"""
import numpy as np
import pytest

import spfft_amd as sp
from oracle import oracle as orc
from tests.workloads import normal, sphere_triplets, uniform01

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def setup(n, radius, precision, seed=5):
    from spfft_amd.autograd import sparse_energy  # noqa: F401
    dims = (n, n, n)
    t = sphere_triplets(radius)
    cls = sp.Transform if precision == "double" else sp.TransformFloat
    tr = cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t), sp.SPFFT_INDEX_TRIPLETS, t)
    c = normal(seed, len(t)) + 1j * normal(seed + 1, len(t))
    v = uniform01(seed + 2, n ** 3).reshape(n, n, n)
    return dims, t, tr, c, v


def oracle_grad(dims, t, c, v):
    o = orc.OraclePlan(False, dims, t)
    psi = o.backward(c)
    return 2.0 * o.forward(v * psi, False), float(np.sum(v * np.abs(psi) ** 2))


@pytest.mark.parametrize("n,radius", [(32, 10), (64, 20)])
def test_energy_gradient_fd_and_oracle_fp64(n, radius):
    from spfft_amd.autograd import sparse_energy
    dims, t, tr, c, v = setup(n, radius, "double")
    ct = torch.tensor(c, dtype=torch.complex128, device="cuda", requires_grad=True)
    vt = torch.tensor(v, dtype=torch.float64, device="cuda")
    e = sparse_energy(ct, vt, tr, dims)
    e.backward()
    g = ct.grad.detach().cpu().numpy()
    want, e_want = oracle_grad(dims, t, c, v)
    assert abs(e.item() - e_want) <= 1e-12 * abs(e_want)
    assert orc.rel_l2(g, want) <= 1e-12

    # central finite differences on 64 random coefficients, re and im parts
    pick = (uniform01(77, 64) * len(t)).astype(int)
    h = 1e-2
    gnorm = np.linalg.norm(g) / np.sqrt(len(g))
    with torch.no_grad():
        for k in pick:
            for part in (1.0, 1j):
                cp = ct.detach().clone()
                cm = ct.detach().clone()
                cp[k] += h * part
                cm[k] -= h * part
                fd = (sparse_energy(cp, vt, tr, dims).item() -
                      sparse_energy(cm, vt, tr, dims).item()) / (2 * h)
                an = g[k].real if part == 1.0 else g[k].imag
                assert abs(fd - an) <= 1e-6 * max(gnorm, abs(an)), (k, part, fd, an)


def test_energy_gradient_fp32_and_store_vs_recompute():
    from spfft_amd.autograd import sparse_energy
    dims, t, tr, c, v = setup(64, 20, "float", seed=9)
    want, _ = oracle_grad(dims, t, c, v)
    grads = []
    for recompute in (True, False):
        ct = torch.tensor(c, dtype=torch.complex64, device="cuda", requires_grad=True)
        vt = torch.tensor(v, dtype=torch.float32, device="cuda")
        sparse_energy(ct, vt, tr, dims, recompute).backward()
        grads.append(ct.grad.detach().cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])
    assert orc.rel_l2(grads[0].astype(np.complex128), want) <= 1e-5


def test_sparse_backward_adjoint_identity():
    """<B c, y> == <c, B^H y> for random c, y (the adjoint is exact)."""
    from spfft_amd.autograd import sparse_backward
    dims, t, tr, c, _ = setup(32, 10, "double", seed=13)
    y = normal(21, 32 ** 3).reshape(32, 32, 32) + 1j * normal(22, 32 ** 3).reshape(32, 32, 32)
    ct = torch.tensor(c, dtype=torch.complex128, device="cuda", requires_grad=True)
    psi = sparse_backward(ct, tr, dims)
    yt = torch.tensor(y, dtype=torch.complex128, device="cuda")
    lhs = torch.sum(torch.conj(yt) * psi)
    (g,) = torch.autograd.grad(torch.real(lhs), ct)
    # d Re<y, Bc> / dc* gives B^H y in the conjugate-Wirtinger convention
    o = orc.OraclePlan(False, dims, t)
    assert orc.rel_l2(g.detach().cpu().numpy(), o.forward(y, False)) <= 1e-12


@pytest.mark.parametrize("precision", ["double", "float"])
def test_sphere_192_r60_gradient_vs_oracle(precision):
    """The 192^3, r=60 configuration (904,089 coefficients), one band."""
    from spfft_amd.autograd import sparse_energy
    dims, t, tr, c, v = setup(192, 60, precision, seed=31)
    cdt = torch.complex128 if precision == "double" else torch.complex64
    rdt = torch.float64 if precision == "double" else torch.float32
    ct = torch.tensor(c, dtype=cdt, device="cuda", requires_grad=True)
    vt = torch.tensor(v, dtype=rdt, device="cuda")
    sparse_energy(ct, vt, tr, dims).backward()
    want, _ = oracle_grad(dims, t, c, v)
    err = orc.rel_l2(ct.grad.detach().cpu().numpy().astype(np.complex128), want)
    print(f"192^3 r=60 {precision}: gradient rel L2 vs FP64 oracle {err:.2e}")
    assert err <= (1e-12 if precision == "double" else 1e-5)


@pytest.fixture(scope="module")
def cfg5_bands():
    """8 bands on 192^3 with the r=60 sphere (904,089 coefficients each),
    c iid complex N(0,1), V iid U[0,1); the FP64 oracle energy and gradient of
    every band."""
    n, bands = 192, 8
    dims = (n, n, n)
    t = sphere_triplets(60)
    C = np.stack([normal(0x5EED0500 + 2 * b, len(t)) + 1j * normal(0x5EED0501 + 2 * b, len(t))
                  for b in range(bands)])
    v = uniform01(0x5EED05FF, n ** 3).reshape(n, n, n)
    o = orc.OraclePlan(False, dims, t)
    grads, energy = [], 0.0
    for b in range(bands):
        psi = o.backward(C[b])
        grads.append(2.0 * o.forward(v * psi, False))
        energy += float(np.sum(v * np.abs(psi) ** 2))
    return dims, t, C, v, np.stack(grads), energy


@pytest.mark.parametrize("precision", ["double", "float"])
def test_cfg5_eight_band_gradient_fd_and_oracle(cfg5_bands, precision):
    """The differentiated workload at its stated shape: 8 bands of 192^3 r=60
    through the batched chain (one multi_transform call per direction),
    gradient against the FP64 oracle for every band, and central finite
    differences on 64 random (band, coefficient) pairs, real and imaginary
    parts. The loss is quadratic, so the central difference is exact for any
    step: FP64 uses h = 1 (agreement 1e-6 of the gradient's RMS), FP32 h = 16
    with the energy accumulated in FP64 (agreement 1e-4: the FP32 fields'
    rounding, about 2e-7 relative, changes with every evaluation)."""
    from spfft_amd.autograd import sparse_energy_batch
    dims, t, C, v, want, e_want = cfg5_bands
    bands = C.shape[0]
    cls = sp.Transform if precision == "double" else sp.TransformFloat
    trs = [cls(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t), sp.SPFFT_INDEX_TRIPLETS, t)
           for _ in range(bands)]
    cdt = torch.complex128 if precision == "double" else torch.complex64
    rdt = torch.float64 if precision == "double" else torch.float32
    tol = 1e-12 if precision == "double" else 1e-5
    ct = torch.tensor(C, dtype=cdt, device="cuda", requires_grad=True)
    vt = torch.tensor(v, dtype=rdt, device="cuda")
    e = sparse_energy_batch(ct, vt, trs, dims)
    e.backward()
    g = ct.grad.detach().cpu().numpy().astype(np.complex128)
    e_err = abs(e.item() - e_want) / e_want
    errs = [orc.rel_l2(g[b], want[b]) for b in range(bands)]
    print(f"cfg5 {precision}: energy rel err {e_err:.2e}; gradient rel L2 vs FP64 oracle per band "
          + " ".join(f"{x:.1e}" for x in errs))
    assert e_err <= tol
    assert max(errs) <= tol

    pick_b = (uniform01(0x5EED0510, 64) * bands).astype(int)
    pick_k = (uniform01(0x5EED0511, 64) * len(t)).astype(int)
    h = 1.0 if precision == "double" else 16.0
    rel = 1e-6 if precision == "double" else 1e-4
    gnorm = np.linalg.norm(g) / np.sqrt(g.size)
    worst = 0.0
    with torch.no_grad():
        for b, k in zip(pick_b, pick_k):
            for part in (1.0, 1j):
                cp = ct.detach().clone()
                cm = ct.detach().clone()
                cp[b, k] += h * part
                cm[b, k] -= h * part
                fd = (sparse_energy_batch(cp, vt, trs, dims).item() -
                      sparse_energy_batch(cm, vt, trs, dims).item()) / (2 * h)
                an = g[b, k].real if part == 1.0 else g[b, k].imag
                worst = max(worst, abs(fd - an) / max(gnorm, abs(an)))
                assert abs(fd - an) <= rel * max(gnorm, abs(an)), (b, k, part, fd, an)
    print(f"cfg5 {precision}: worst finite-difference mismatch {worst:.1e} of the gradient RMS")


def test_cfg5_recompute_and_store_agree(cfg5_bands):
    """recompute="auto" recomputes the fields when they exceed the memory
    budget and keeps them otherwise; both give the same gradient bit for bit."""
    from spfft_amd.autograd import recompute_fields, sparse_energy_batch
    dims, t, C, v, _, _ = cfg5_bands
    field_bytes = C.shape[0] * 192 ** 3 * 16
    assert recompute_fields(field_bytes, "auto", memory_budget=field_bytes - 1)
    assert not recompute_fields(field_bytes, "auto", memory_budget=field_bytes)
    assert recompute_fields(0, True) and not recompute_fields(1 << 40, False)
    trs = [sp.Transform(-1, sp.SPFFT_PU_GPU, sp.SPFFT_TRANS_C2C, *dims, len(t),
                        sp.SPFFT_INDEX_TRIPLETS, t) for _ in range(C.shape[0])]
    vt = torch.tensor(v, dtype=torch.float64, device="cuda")
    grads = []
    for budget in (0, 1 << 40):
        ct = torch.tensor(C, dtype=torch.complex128, device="cuda", requires_grad=True)
        sparse_energy_batch(ct, vt, trs, dims, memory_budget=budget).backward()
        grads.append(ct.grad.detach().cpu().numpy())
    np.testing.assert_array_equal(grads[0], grads[1])
