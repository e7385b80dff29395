"""Drop-in for the reference's pybind module ``JITBeijingGeometry``.

Same functions and argument meaning as
``model/ConeBeamLayers/Beijing/plug/jit.cpp:5-7`` (``plug/include/
forward.h:3``, ``backward.h:3``):

    forward(volume, volumeSize, detectorSize, projectVector, device) -> sino
    backward(sino, volumeSize, detectorSize, projectVector, device) -> volume

``forward`` omits the ``sampleInterval`` factor exactly as the reference
kernel does (the Python layer applies it, BeijingGeometry.py:17).
``backward`` is the reference's triple-product backprojector; ``adjoint``
is added for the exact transpose of ``forward``.
"""
from __future__ import annotations

import torch

from dtvproj import _capi
from dtvproj.ops import _stream, check_input
from JITStandardGeometry import _check_volume, _sizes


def _vectors(projectVector):
    check_input(projectVector, "projectVector")
    if projectVector.dim() != 2 or projectVector.size(1) != 12:
        raise RuntimeError("project vector's shape must be [angle's number, 12]")
    if projectVector.dtype != torch.float32:
        projectVector = projectVector.to(torch.float32).contiguous()
    return projectVector


def forward(volume, volumeSize, detectorSize, projectVector, device):
    """Ray-driven forward projection (Beijing/plug/kernel/forwardKernel.cu:78-135)."""
    check_input(volume, "volume")   # argument order of the reference's CHECK_INPUTs
    vs, ds = _sizes(volumeSize, detectorSize)
    pv = _vectors(projectVector)
    B = _check_volume(volume, "volume", vs[0] * vs[1] * vs[2])
    n = int(pv.size(0))
    out = torch.empty((B, 1, n, ds[1], ds[0]), dtype=torch.float32, device=volume.device)
    with torch.cuda.device(volume.device):
        rc = _capi.lib().dtv_beijing_forward(
            volume.data_ptr(), out.data_ptr(), B, _capi.int3(vs), _capi.int2(ds), pv.data_ptr(),
            n, int(device) if device is not None else -1, _stream(volume.device))
    _capi.check(rc, "JITBeijingGeometry.forward")
    return out


def backward(sino, volumeSize, detectorSize, projectVector, device):
    """Reference backprojector (Beijing/plug/kernel/backwardKernel.cu:47-104)."""
    check_input(sino, "sino")   # argument order of the reference's CHECK_INPUTs
    vs, ds = _sizes(volumeSize, detectorSize)
    pv = _vectors(projectVector)
    n = int(pv.size(0))
    B = _check_volume(sino, "sino", n * ds[0] * ds[1])
    out = torch.empty((B, 1, vs[2], vs[1], vs[0]), dtype=torch.float32, device=sino.device)
    with torch.cuda.device(sino.device):
        rc = _capi.lib().dtv_beijing_backward(
            sino.data_ptr(), out.data_ptr(), B, _capi.int3(vs), _capi.int2(ds), pv.data_ptr(),
            n, int(device) if device is not None else -1, _stream(sino.device))
    _capi.check(rc, "JITBeijingGeometry.backward")
    return out


def adjoint(sino, volumeSize, detectorSize, projectVector, device):
    """Exact transpose of :func:`forward` (no reference counterpart)."""
    check_input(sino, "sino")   # argument order of the reference's CHECK_INPUTs
    vs, ds = _sizes(volumeSize, detectorSize)
    pv = _vectors(projectVector)
    n = int(pv.size(0))
    B = _check_volume(sino, "sino", n * ds[0] * ds[1])
    out = torch.empty((B, 1, vs[2], vs[1], vs[0]), dtype=torch.float32, device=sino.device)
    with torch.cuda.device(sino.device):
        rc = _capi.lib().dtv_beijing_adjoint(
            sino.data_ptr(), out.data_ptr(), B, _capi.int3(vs), _capi.int2(ds), pv.data_ptr(),
            n, int(device) if device is not None else -1, _stream(sino.device))
    _capi.check(rc, "JITBeijingGeometry.adjoint")
    return out
