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
the GPU; the dense field is shaped (Nz, Ny, Nx). Only C2C transforms qualify;
shapes, dtypes, devices and the transform type are checked before any pointer
reaches the library (ValueError otherwise). The library enqueues on the HIP
default stream; when torch's current stream is another one, the default stream
first waits for it and the current stream then waits for the transform.
"""
from __future__ import annotations

import torch

import spfft_amd as sp


def _dtype(transform):
    return torch.complex128 if transform._precision == "double" else torch.complex64


def _check_tensor(name, x, transform, numel):
    want = _dtype(transform)
    if x.dtype != want:
        raise ValueError(f"{name}: dtype {x.dtype}, the transform needs {want}")
    if x.device.type != "cuda":
        raise ValueError(f"{name}: tensor on {x.device}, the transform runs on the GPU")
    if transform.device_id() != (x.device.index if x.device.index is not None
                                 else torch.cuda.current_device()):
        raise ValueError(f"{name}: tensor on {x.device}, the transform on device "
                         f"{transform.device_id()}")
    if x.numel() != numel:
        raise ValueError(f"{name}: {x.numel()} elements, the transform needs {numel}")


def _check_transform(transform, dims):
    if transform.transform_type != sp.SPFFT_TRANS_C2C:
        raise ValueError("only C2C transforms have a complex field to differentiate")
    if tuple(dims) != (transform.dim_x(), transform.dim_y(), transform.dim_z()):
        raise ValueError(f"dims {tuple(dims)} differ from the transform's "
                         f"{(transform.dim_x(), transform.dim_y(), transform.dim_z())}")


class _OnDefaultStream:
    """Orders a library call (HIP default stream) against torch's current stream."""

    def __enter__(self):
        self.cur = torch.cuda.current_stream()
        self.dflt = torch.cuda.default_stream()
        self.other = self.cur.cuda_stream != self.dflt.cuda_stream
        if self.other:
            self.dflt.wait_stream(self.cur)

    def __exit__(self, *exc):
        if self.other:
            self.cur.wait_stream(self.dflt)
        return False


def _backward(transform, dims, c):
    nx, ny, nz = dims
    _check_transform(transform, dims)
    _check_tensor("c", c, transform, transform.num_values)
    c = c.contiguous()
    psi = torch.empty((nz, ny, nx), dtype=c.dtype, device=c.device)
    with _OnDefaultStream():
        transform.backward(torch.view_as_real(c).data_ptr(), torch.view_as_real(psi).data_ptr())
    return psi


def _adjoint(transform, n, field):
    """B^H field: unnormalised forward restricted to the sparse set."""
    _check_tensor("field", field, transform,
                  transform.dim_x() * transform.dim_y() * transform.dim_z())
    field = field.contiguous()
    out = torch.empty(n, dtype=field.dtype, device=field.device)
    with _OnDefaultStream():
        transform.forward(torch.view_as_real(field).data_ptr(),
                          torch.view_as_real(out).data_ptr(), sp.SPFFT_NO_SCALING)
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
