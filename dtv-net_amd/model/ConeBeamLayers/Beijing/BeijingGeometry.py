"""Drop-in for ``model/ConeBeamLayers/Beijing/BeijingGeometry.py``.

Same module globals (``parameters``, ``volumeSize``, ``detectorSize``),
autograd Functions and modules as the reference (BeijingGeometry.py:7-84),
running on the gfx950 kernels.  ``parameters`` is loaded from the .mat file
named by ``beijingParameterRoot`` (key ``projection_matrix``) when it exists;
otherwise a synthetic circular stand-in with the same [angles*planes, 12]
layout is generated (the reference's file is not distributed with it).

Layout: the projector's sinogram is [angles*planes, V, U] with the planes of
one angle consecutive; the reference's public layout interleaves them as
[angles, V, planes, U] (the permutes at BeijingGeometry.py:18, 23, 31, 39),
which is kept here for any batch size (the reference fixes batch 1).

``adjoint_mode``: ``"reference"`` (default) uses the reference's
backprojector and its factor placement (``sampleInterval`` in
ForwardProjection.forward and BackProjection.backward only,
BeijingGeometry.py:17, 22-25, 38), so reference callers keep their tuned
step sizes (``lamb = 10e-3``); ``"matched"`` uses the exact transpose of
``sampleInterval * A`` as A^T, so autograd gradients are exact (a reference
step size then needs rescaling to the scale of A^T A).
"""
import os

import numpy as np
import torch

import JITBeijingGeometry as projector
from dtvproj import fbp as _fbp
from dtvproj import geometry as _G
from options import beijingAngleNum, beijingPlanes, beijingSubDetectorSize, beijingVolumeSize, \
    beijingParameterRoot, beijingSID, beijingSDD, sampleInterval


def _load_parameters():
    if os.path.exists(beijingParameterRoot):
        import scipy.io as sco
        mat = sco.loadmat(beijingParameterRoot)
        p = np.array(mat["projection_matrix"]).astype(np.float32)
    else:
        # source and detector distances taken in voxels; pixel pitches chosen
        # so the tiled detector covers the volume's shadow with 20% margin
        U, V = beijingSubDetectorSize
        X, Y, Z = beijingVolumeSize
        mag = beijingSDD / beijingSID
        pu = 1.2 * np.hypot(X, Y) * mag / (U * beijingPlanes)
        pv = 1.2 * Z * mag / V
        p = _G.beijing_default_vectors(beijingAngleNum, beijingPlanes, (U, V), beijingSID,
                                       beijingSDD, (pu, pv))
    return torch.from_numpy(p).contiguous()


parameters = _load_parameters()
volumeSize = torch.IntTensor(beijingVolumeSize)
detectorSize = torch.IntTensor(beijingSubDetectorSize)
adjoint_mode = "reference"


def _args(device):
    return volumeSize.to(device), detectorSize.to(device), parameters.to(device), device.index


def _to_public(sino):
    """[B,1,A*P,V,U] (planes consecutive) -> [B,1,A*P,V,U] in the
    reference's interleaved order (angles, V, planes, U)."""
    B = sino.shape[0]
    U, V = beijingSubDetectorSize
    return sino.reshape(B, beijingAngleNum, beijingPlanes, V, U).permute(0, 1, 3, 2, 4) \
        .reshape(B, 1, beijingAngleNum * beijingPlanes, V, U)


def _to_projector(sino):
    B = sino.shape[0]
    U, V = beijingSubDetectorSize
    return sino.reshape(B, beijingAngleNum, V, beijingPlanes, U).permute(0, 1, 3, 2, 4) \
        .reshape(B, 1, beijingAngleNum * beijingPlanes, V, U).contiguous()


def _A(x):
    return _to_public(projector.forward(x.contiguous(), *_args(x.device)) * sampleInterval)


def _AT(sino):
    s = _to_projector(sino)
    if adjoint_mode == "reference":
        return projector.backward(s, *_args(s.device))
    if adjoint_mode != "matched":
        raise RuntimeError(f"unknown adjoint_mode {adjoint_mode!r}")
    return projector.adjoint(s, *_args(s.device)) * sampleInterval


class ForwardProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input):
        return _A(input)

    @staticmethod
    def backward(ctx, grad):
        if adjoint_mode == "reference":
            # the reference omits sampleInterval here (BeijingGeometry.py:22-25)
            s = _to_projector(grad)
            return projector.backward(s, *_args(s.device))
        return _AT(grad)


class BackProjection(torch.autograd.Function):
    @staticmethod
    def forward(ctx, input):
        return _AT(input)

    @staticmethod
    def backward(ctx, grad):
        return _A(grad)


class BeijingGeometry(torch.nn.Module):
    """The step only: -lamb * A^T (A x - p); callers add x
    (BeijingGeometry.py:41-48)."""

    def __init__(self):
        super(BeijingGeometry, self).__init__()
        self.lamb = 10e-3

    def forward(self, x, p):
        residual = ForwardProjection.apply(x) - p
        return -self.lamb * BackProjection.apply(residual)


class BeijingGeometryWithFBP(torch.nn.Module):
    """Cosine weight and Ram-Lak filter over the tiled detector row of all
    planes, then x - lamb * A^T (BeijingGeometry.py:50-64)."""

    def __init__(self):
        super(BeijingGeometryWithFBP, self).__init__()
        self.lamb = torch.nn.Parameter(torch.tensor(1.0), requires_grad=False)
        self.cosWeight = torch.nn.Parameter(self.__cosWeight__(), requires_grad=False)
        self.ramp = torch.nn.Parameter(self.__conv__(beijingSubDetectorSize[0] * beijingPlanes),
                                       requires_grad=False)

    def forward(self, x, p):
        U, V = beijingSubDetectorSize
        B = x.shape[0]
        residual = ForwardProjection.apply(x) - p
        # rows of the tiled detector (all planes side by side, the public
        # layout): cosine weight and Ram-Lak row convolution in one pass
        residual = residual.reshape(B, 1, beijingAngleNum, V, U * beijingPlanes)
        residual = _fbp.row_filter(residual, self.ramp, self.cosWeight[0, 0, 0])
        residual = residual.view(B, 1, beijingAngleNum * beijingPlanes, V, U)
        return x - self.lamb * BackProjection.apply(residual)

    def __cosWeight__(self):
        """SDD / sqrt(SDD^2 + (i - W/2)^2 + (j - V/2)^2) over rows i < V and
        tiled columns j < W = U*planes -- the reference pairs the row index
        with the row-length centre and vice versa (BeijingGeometry.py:66-72)."""
        U, V = beijingSubDetectorSize
        W = U * beijingPlanes
        i = np.arange(V, dtype=np.float64)[:, None]
        j = np.arange(W, dtype=np.float64)[None, :]
        w = beijingSDD / np.sqrt(beijingSDD ** 2 + (i - W / 2) ** 2 + (j - V / 2) ** 2)
        out = np.broadcast_to(w.astype(np.float32), (1, 1, beijingAngleNum, V, W))
        return torch.from_numpy(np.ascontiguousarray(out))

    def __conv__(self, projWidth):
        n = np.arange(projWidth + 1, dtype=np.float64) - np.floor(projWidth / 2)
        taps = np.where(n % 2 == 0, 0.0, -0.5 / (np.pi * np.pi * np.where(n == 0, 1.0, n) ** 2))
        taps[n == 0] = 1.0 / 8
        return torch.from_numpy(taps.astype(np.float32).reshape(1, 1, 1, projWidth + 1))
