"""Drop-in for ``model/ConeBeamLayers/Standard/StandardGeometry.py``.

Same module globals (``angles``, ``volumeSize``, ``detectorSize``), autograd
Functions (``ForwardProjection``, ``BackProjection``) and modules
(``StandardGeometry``, ``StandardGeometryWithFBP``) as the reference
(StandardGeometry.py:6-93), running on the gfx950 kernels.  Globals are read
at call time, so callers may reassign them as the reference's drivers do.

``adjoint_mode`` selects the operator used as A^T:

* ``"reference"`` (default) -- the reference's FDK-weighted voxel-driven
  backprojector (backwardKernel.cu:22-53), as the reference wires it
  (StandardGeometry.py:17-21, 30-34), so callers written for the reference
  keep their iterate trajectories and hand-tuned step sizes (``lamb = 1.0``
  here; ASD-POCS ``beta``, DTV-CP ``tau``/``sigma``);
* ``"matched"`` -- the exact transpose of A, so the backward of
  ``ForwardProjection`` is the true gradient and vice versa.  Its scale is
  that of A^T A (about ``unroll.power_norm`` of the geometry), not of the
  FDK backprojector, so a reference step size must be rescaled, e.g.
  ``lamb = 1 / power_norm``.
"""
import numpy as np
import torch

import JITStandardGeometry as projector
from dtvproj import fbp as _fbp
from options import standardAngleNum, standardDetectorSize, standardVolumeSize, standardSID, \
    standardSDD

angles = torch.from_numpy(np.linspace(0, 360, standardAngleNum, False))
volumeSize = torch.tensor(standardVolumeSize)
detectorSize = torch.tensor(standardDetectorSize)
adjoint_mode = "reference"


def _args(device):
    return (angles.to(device), volumeSize.to(device), detectorSize.to(device), standardSID,
            standardSDD, device.index)


def _transpose(sino):
    device = sino.device
    if adjoint_mode == "reference":
        return projector.backward(sino.contiguous(), *_args(device))
    if adjoint_mode != "matched":
        raise RuntimeError(f"unknown adjoint_mode {adjoint_mode!r}")
    return projector.adjoint(sino.contiguous(), *_args(device))


class ForwardProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input):
        return projector.forward(input.contiguous(), *_args(input.device))

    @staticmethod
    def backward(ctx, grad):
        return _transpose(grad)


class BackProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input):
        return _transpose(input)

    @staticmethod
    def backward(ctx, grad):
        return projector.forward(grad.contiguous(), *_args(grad.device))


class StandardGeometry(torch.nn.Module):
    """One gradient step x - lamb * A^T (A x - p) (StandardGeometry.py:36-43)."""

    def __init__(self):
        super(StandardGeometry, self).__init__()
        self.lamb = 1.0

    def forward(self, x, p):
        residual = ForwardProjection.apply(x) - p
        return x - self.lamb * BackProjection.apply(residual)


class StandardGeometryWithFBP(torch.nn.Module):
    """Cosine-weighted, Ram-Lak-filtered residual backprojection
    (StandardGeometry.py:45-66)."""

    def __init__(self):
        super(StandardGeometryWithFBP, self).__init__()
        self.lamb = torch.nn.Parameter(torch.tensor(1.0), requires_grad=False)
        self.cosWeight = torch.nn.Parameter(self.__cosWeight__(), requires_grad=False)
        self.ramp = torch.nn.Parameter(self.__conv__(int(detectorSize[0])), requires_grad=False)

    def forward(self, x, p):
        residual = ForwardProjection.apply(x) - p
        # cosine weight and the Ram-Lak row convolution in one gfx950 pass
        # (the reference's residual * cosWeight, then conv2d with padding U/2)
        residual = _fbp.row_filter(residual, self.ramp, self.cosWeight[0, 0, 0])
        return self.lamb * BackProjection.apply(residual)

    def __cosWeight__(self):
        """SDD / sqrt(SDD^2 + (iv - V/2)^2 + (iu - U/2)^2) per pixel."""
        U, V = int(detectorSize[0]), int(detectorSize[1])
        iv = np.arange(V, dtype=np.float64)[:, None] - V / 2
        iu = np.arange(U, dtype=np.float64)[None, :] - U / 2
        w = standardSDD / np.sqrt(standardSDD ** 2 + iv ** 2 + iu ** 2)
        out = np.broadcast_to(w.astype(np.float32), (1, 1, standardAngleNum, V, U))
        return torch.from_numpy(np.ascontiguousarray(out))

    def __filter__(self):
        """Triangular frequency ramp (kept for interface parity; unused)."""
        U, V = int(detectorSize[0]), int(detectorSize[1])
        mid = U / 2
        ramp = (mid - np.abs(np.arange(U) - mid)).astype(np.float32)
        out = np.broadcast_to(ramp, (1, 1, standardAngleNum, V, U))
        return torch.from_numpy(np.ascontiguousarray(out))

    def __conv__(self, projWidth):
        """Spatial Ram-Lak kernel of width projWidth+1: 1/8 at the centre,
        -0.5/(pi^2 n^2) at odd offsets n, 0 at even offsets."""
        n = np.arange(projWidth + 1, dtype=np.float64) - np.floor(projWidth / 2)
        taps = np.where(n % 2 == 0, 0.0, -0.5 / (np.pi * np.pi * np.where(n == 0, 1.0, n) ** 2))
        taps[n == 0] = 1.0 / 8
        return torch.from_numpy(taps.astype(np.float32).reshape(1, 1, 1, projWidth + 1))
