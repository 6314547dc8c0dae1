"""PyTorch autograd shim over the HIP transforms (test/integration layer).

SparseBackward: psi = B c, where B is the backward sparse 3D DFT
(exp(+2 pi i k.r / N), unnormalised) from the sparse coefficients c to the
dense grid. Its adjoint is B^H = the unnormalised forward transform restricted
to the sparse set, i.e. the library's forward kernel chain with gather and
scatter swapped; with PyTorch's conjugate-Wirtinger convention the gradient
of a real loss is grad_c = B^H grad_psi.

SparseEnergy: L = sum_r V(r) |psi(r)|^2 with psi = B c, so
grad_c = 2 B^H (V psi). With recompute=True (the default) psi is not kept
between forward and backward: the backward recomputes it, so memory per band
stays at one dense field regardless of the batch size.

Tensors are complex128 (double transforms) or complex64 (float transforms) on
the GPU; the dense field is shaped (Nz, Ny, Nx).
"""
from __future__ import annotations

import torch

import spfft_amd as sp


def _dtype(transform):
    return torch.complex128 if transform._precision == "double" else torch.complex64


def _backward(transform, dims, c):
    nx, ny, nz = dims
    c = c.contiguous()
    psi = torch.empty((nz, ny, nx), dtype=c.dtype, device=c.device)
    transform.backward(torch.view_as_real(c).data_ptr(), torch.view_as_real(psi).data_ptr())
    return psi


def _adjoint(transform, n, field):
    """B^H field: unnormalised forward restricted to the sparse set."""
    field = field.contiguous()
    out = torch.empty(n, dtype=field.dtype, device=field.device)
    transform.forward(torch.view_as_real(field).data_ptr(), torch.view_as_real(out).data_ptr(),
                      sp.SPFFT_NO_SCALING)
    return out


class SparseBackward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, transform, dims):
        ctx.transform = transform
        ctx.n = c.numel()
        return _backward(transform, dims, c)

    @staticmethod
    def backward(ctx, grad_psi):
        return _adjoint(ctx.transform, ctx.n, grad_psi), None, None


class SparseEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, c, potential, transform, dims, recompute=True):
        psi = _backward(transform, dims, c)
        energy = torch.sum(potential * (psi.real ** 2 + psi.imag ** 2))
        ctx.transform, ctx.dims, ctx.recompute = transform, dims, recompute
        if recompute:
            ctx.save_for_backward(c, potential)
        else:
            ctx.save_for_backward(c, potential, psi)
        return energy

    @staticmethod
    def backward(ctx, grad_energy):
        if ctx.recompute:
            c, potential = ctx.saved_tensors
            psi = _backward(ctx.transform, ctx.dims, c)
        else:
            c, potential, psi = ctx.saved_tensors
        g = _adjoint(ctx.transform, c.numel(), (2.0 * potential) * psi)
        return grad_energy * g, None, None, None, None


def sparse_backward(c, transform, dims):
    return SparseBackward.apply(c, transform, dims)


def sparse_energy(c, potential, transform, dims, recompute=True):
    return SparseEnergy.apply(c, potential, transform, dims, recompute)
