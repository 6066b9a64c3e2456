"""Differentiable projectors (autograd.py:17-119 of the reference).

forward_project's gradient is the hand-written back projector and vice
versa; both run through libmbir_hip.so.  The pair is an exact adjoint pair by
construction (checked by the adjoint and finite-difference tests).
"""

import torch


class _ForwardProjectFunction(torch.autograd.Function):
    """A: voxel cylinders at fixed indices -> sinogram; backward is A^T."""

    @staticmethod
    def forward(ctx, voxel_values, model, pixel_indices):
        ctx.model = model
        ctx.pixel_indices = pixel_indices
        return model.sparse_forward_project(voxel_values, pixel_indices)

    @staticmethod
    def backward(ctx, grad_sinogram):
        grad = ctx.model.sparse_back_project(grad_sinogram.contiguous(), ctx.pixel_indices)
        return grad, None, None


class _BackProjectFunction(torch.autograd.Function):
    """A^T: sinogram -> voxel cylinders at fixed indices; backward is A."""

    @staticmethod
    def forward(ctx, sinogram, model, pixel_indices):
        ctx.model = model
        ctx.pixel_indices = pixel_indices
        return model.sparse_back_project(sinogram, pixel_indices)

    @staticmethod
    def backward(ctx, grad_values):
        grad = ctx.model.sparse_forward_project(grad_values.contiguous(), ctx.pixel_indices)
        return grad, None, None


def forward_project_differentiable(model, volume):
    """Differentiable full-volume forward projection; the volume may live on
    any device/dtype (moved differentiably to the model's device as float32)."""
    volume = volume.to(device=model.torch_device, dtype=torch.float32)
    indices = model.full_indices_device()
    voxel_values = volume.reshape(-1, volume.shape[-1])[indices]
    return _ForwardProjectFunction.apply(voxel_values, model, indices)


def back_project_differentiable(model, sinogram):
    """Differentiable full back projection (adjoint of the above)."""
    recon_shape = model.get_params('recon_shape')
    sinogram = sinogram.to(device=model.torch_device, dtype=torch.float32)
    indices = model.full_indices_device()
    cyl = _BackProjectFunction.apply(sinogram, model, indices)
    volume = torch.zeros(recon_shape[0] * recon_shape[1], cyl.shape[-1], dtype=cyl.dtype,
                         device=cyl.device)
    volume = volume.index_put((indices,), cyl)
    return volume.reshape(tuple(recon_shape[:2]) + (cyl.shape[-1],))


class TorchProjector(torch.nn.Module):
    """The forward operator as an nn.Module; ``adjoint`` is the back projector."""

    def __init__(self, model):
        super().__init__()
        self.model = model

    def forward(self, volume):
        return forward_project_differentiable(self.model, volume)

    def adjoint(self, sinogram):
        return back_project_differentiable(self.model, sinogram)
