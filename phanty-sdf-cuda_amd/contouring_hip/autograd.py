"""Differentiable dual contouring on PyTorch-ROCm: vertices as a function of the SDF samples.

DualContouringFunction.forward runs the HIP pipeline on a float32 CUDA tensor phi of shape
(nz, ny, nx); backward turns dL/dV into dL/dphi with the adjoint kernels (implicit
differentiation of the truncated QEF solve through the hermite points and the averaged trilinear
normals, at fixed topology), recomputing per-cell QEF data instead of storing it. No counterpart
in the reference.
"""
import torch

from .backend import HipDualContouringBackend

_default = {}


def default_backend(precision="fp32-fallback", device=0):
    key = (precision, device)
    if key not in _default:
        _default[key] = HipDualContouringBackend(precision, device=device)
    return _default[key]


class DualContouringFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, backend, spacing, origin):
        # the adjoint reads the forward's samples again: keep the contiguous copy the context
        # points at alive until backward (a temporary would be freed, and torch's allocator could
        # hand its memory to the gradient)
        phic = phi.detach().contiguous()
        (V, F), _ = backend.extract_tensor(phic, spacing, origin)
        ctx.phic = phic
        ctx.backend = backend
        ctx.generation = backend.generation
        ctx.spacing = spacing
        ctx.origin = origin
        ctx.save_for_backward(phi)
        ctx.mark_non_differentiable(F)
        return V, F

    @staticmethod
    def backward(ctx, gV, gF):
        (phi,) = ctx.saved_tensors
        be = ctx.backend
        if be.generation != ctx.generation:
            # another extraction used the context since: rebuild this forward's topology
            be.extract_tensor(ctx.phic, ctx.spacing, ctx.origin)
        # The forward extracted phi.contiguous() and the adjoint writes dL/dphi in that layout
        # (x fastest): allocate it contiguous whatever phi's strides are (autograd accepts any).
        gphi = torch.empty(phi.shape, dtype=torch.float32, device=phi.device,
                           memory_format=torch.contiguous_format)
        be.backward(gV.contiguous(), gphi)
        ctx.phic = None
        return gphi, None, None, None


def dual_contour(phi, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), backend=None,
                 precision="fp32-fallback"):
    """(V [A,3] float32, F [F,4] int32) with V differentiable with respect to phi."""
    if backend is None:
        backend = default_backend(precision, phi.device.index or 0)
    return DualContouringFunction.apply(phi, backend, tuple(spacing), tuple(origin))
