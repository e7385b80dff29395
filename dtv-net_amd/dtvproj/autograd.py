"""Differentiable matched projector pair.

The backward of A is the matched A^T kernel and the backward of A^T is the A
kernel, so gradients are exact for the discrete operator (the reference
instead wires its FDK-weighted backprojector as A's gradient,
StandardGeometry.py:18-21, 31-34; BeijingGeometry.py:20-25, 35-39).
"""
from __future__ import annotations

import torch

from . import ops
from .geometry import ConeGeometry


class _Project(torch.autograd.Function):
    @staticmethod
    def forward(ctx, volume, geom):
        ctx.geom = geom
        return ops.forward(volume.contiguous(), geom)

    @staticmethod
    def backward(ctx, grad):
        return ops.adjoint(grad.contiguous(), ctx.geom), None


class _Backproject(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, geom):
        ctx.geom = geom
        return ops.adjoint(sino.contiguous(), geom)

    @staticmethod
    def backward(ctx, grad):
        return ops.forward(grad.contiguous(), ctx.geom), None


def project(volume: torch.Tensor, geom: ConeGeometry) -> torch.Tensor:
    """A(volume), differentiable (gradient = A^T)."""
    return _Project.apply(volume, geom)


def backproject(sino: torch.Tensor, geom: ConeGeometry) -> torch.Tensor:
    """A^T(sino), differentiable (gradient = A)."""
    return _Backproject.apply(sino, geom)


class ConeProjector(torch.nn.Module):
    """nn.Module wrapper: ``forward(x)`` = A x, ``T(y)`` = A^T y."""

    def __init__(self, geom: ConeGeometry):
        super().__init__()
        self.geom = geom

    def forward(self, volume):
        return project(volume, self.geom)

    def T(self, sino):
        return backproject(sino, self.geom)
