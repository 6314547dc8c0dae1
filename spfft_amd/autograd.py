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

SparseEnergyBatch: the same loss summed over bands, L = sum_b sum_r V |B_b c_b|^2,
for bands that each have their own transform (own grid), evaluated with the
fused batched chain (multi_transform_backward for psi, multi_transform_forward
without scaling for the adjoint). recompute="auto" keeps the bands' fields for
the backward only when they fit a memory budget (by default a quarter of the
device memory free at the call), else recomputes them.

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


def _field_budget(memory_budget):
    if memory_budget is not None:
        return int(memory_budget)
    free, _ = torch.cuda.mem_get_info()
    return free // 4


def _batch_backward(transforms, dims, C):
    nx, ny, nz = dims
    n = transforms[0].num_values
    for t in transforms:
        _check_transform(t, dims)
        if t.num_values != n or _dtype(t) != _dtype(transforms[0]):
            raise ValueError("all bands need transforms of one precision and value count")
    if C.dim() != 2 or C.shape[0] != len(transforms):
        raise ValueError(f"C: shape {tuple(C.shape)}, need ({len(transforms)}, {n})")
    _check_tensor("C", C, transforms[0], len(transforms) * n)
    C = C.contiguous()
    psi = torch.empty((len(transforms), nz, ny, nx), dtype=C.dtype, device=C.device)
    with _OnDefaultStream():
        sp.multi_transform_backward(transforms, [torch.view_as_real(C[b]).data_ptr()
                                                 for b in range(len(transforms))],
                                    [torch.view_as_real(psi[b]).data_ptr()
                                     for b in range(len(transforms))])
    return psi


def _batch_adjoint(transforms, n, fields):
    out = torch.empty((len(transforms), n), dtype=fields.dtype, device=fields.device)
    fields = fields.contiguous()
    with _OnDefaultStream():
        sp.multi_transform_forward(transforms, [torch.view_as_real(fields[b]).data_ptr()
                                                for b in range(len(transforms))],
                                   [torch.view_as_real(out[b]).data_ptr()
                                    for b in range(len(transforms))],
                                   [sp.SPFFT_NO_SCALING] * len(transforms))
    return out


def recompute_fields(field_bytes, recompute="auto", memory_budget=None):
    """True when the batched energy's backward recomputes the fields: always
    for recompute=True, never for False, and for "auto" when the fields
    (field_bytes) exceed the budget (memory_budget bytes, or a quarter of the
    device memory free now)."""
    if recompute == "auto":
        return field_bytes > _field_budget(memory_budget)
    return bool(recompute)


class SparseEnergyBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, C, potential, transforms, dims, recompute="auto", memory_budget=None):
        psi = _batch_backward(transforms, dims, C)
        energy = torch.sum(potential.to(torch.float64) *
                           torch.sum((psi.real ** 2 + psi.imag ** 2).to(torch.float64), dim=0))
        ctx.transforms, ctx.dims = transforms, dims
        ctx.recompute = recompute_fields(psi.numel() * psi.element_size(), recompute,
                                         memory_budget)
        if ctx.recompute:
            ctx.save_for_backward(C, potential)
        else:
            ctx.save_for_backward(C, potential, psi)
        return energy

    @staticmethod
    def backward(ctx, grad_energy):
        if ctx.recompute:
            C, potential = ctx.saved_tensors
            psi = _batch_backward(ctx.transforms, ctx.dims, C)
        else:
            C, potential, psi = ctx.saved_tensors
        g = _batch_adjoint(ctx.transforms, C.shape[1], (2.0 * potential) * psi)
        return grad_energy.to(potential.dtype) * g, None, None, None, None, None


def sparse_energy_batch(C, potential, transforms, dims, recompute="auto", memory_budget=None):
    """sum_b sum_r V |psi_b|^2 over the bands' transforms (C: bands x values),
    accumulated and returned in float64 for either precision; the gradient has
    C's dtype."""
    return SparseEnergyBatch.apply(C, potential, transforms, dims, recompute, memory_budget)
