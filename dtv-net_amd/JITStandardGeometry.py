"""Drop-in for the reference's pybind module ``JITStandardGeometry``.

Same functions, argument order and meaning as
``model/ConeBeamLayers/Standard/plug/jit.cpp:5-7`` (declared in
``plug/include/forward.h:3`` and ``backward.h:3``), implemented by the gfx950
library through its C ABI:

    forward(volume, angles, volumeSize, detectorSize, sid, sdd, device) -> sino
    backward(sino, angles, volumeSize, detectorSize, sid, sdd, device) -> volume

``backward`` is the reference's voxel-driven FDK-weighted backprojector (not
the adjoint of ``forward``); ``adjoint`` is added for the exact transpose.
Tensors must be on the GPU and contiguous (the reference's CHECK_INPUT);
the same assertion messages are raised as RuntimeError.
"""
from __future__ import annotations

import torch

from dtvproj import _capi
from dtvproj.ops import _stream, check_input


def _sizes(volumeSize, detectorSize):
    check_input(volumeSize, "_volumeSize")
    if volumeSize.dim() != 1 or volumeSize.size(0) != 3:
        raise RuntimeError("volume size's length must be 3")
    check_input(detectorSize, "_detectorSize")
    if detectorSize.dim() != 1 or detectorSize.size(0) != 2:
        raise RuntimeError("detector size's length must be 2")
    vs = [int(v) for v in volumeSize.cpu().tolist()]
    ds = [int(v) for v in detectorSize.cpu().tolist()]
    return vs, ds


def _angles(angles):
    check_input(angles, "angles")
    if angles.dtype != torch.float64:
        angles = angles.to(torch.float64).contiguous()
    return angles


def _check_volume(t, name, per_item):
    check_input(t, name)
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float32 tensor, got {t.dtype}")
    if t.dim() < 1 or t.numel() != t.size(0) * per_item:
        raise RuntimeError(f"{name} has {t.numel()} elements, expected batch x {per_item}")
    return int(t.size(0))


def forward(volume, angles, volumeSize, detectorSize, sid, sdd, device):
    """Ray-driven forward projection (Standard/plug/kernel/forwardKernel.cu:88-143)."""
    check_input(volume, "volume")   # argument order of the reference's CHECK_INPUTs
    angles = _angles(angles)
    vs, ds = _sizes(volumeSize, detectorSize)
    B = _check_volume(volume, "volume", vs[0] * vs[1] * vs[2])
    A = int(angles.numel())
    out = torch.empty((B, 1, A, ds[1], ds[0]), dtype=torch.float32, device=volume.device)
    with torch.cuda.device(volume.device):
        rc = _capi.lib().dtv_standard_forward(
            volume.data_ptr(), out.data_ptr(), B, angles.data_ptr(), A, _capi.int3(vs),
            _capi.int2(ds), float(sid), float(sdd), int(device) if device is not None else -1,
            _stream(volume.device))
    _capi.check(rc, "JITStandardGeometry.forward")
    return out


def backward(sino, angles, volumeSize, detectorSize, sid, sdd, device):
    """Reference backprojector (Standard/plug/kernel/backwardKernel.cu:55-110)."""
    check_input(sino, "sino")   # argument order of the reference's CHECK_INPUTs
    angles = _angles(angles)
    vs, ds = _sizes(volumeSize, detectorSize)
    A = int(angles.numel())
    B = _check_volume(sino, "sino", A * ds[0] * ds[1])
    out = torch.empty((B, 1, vs[2], vs[1], vs[0]), dtype=torch.float32, device=sino.device)
    with torch.cuda.device(sino.device):
        rc = _capi.lib().dtv_standard_backward(
            sino.data_ptr(), out.data_ptr(), B, angles.data_ptr(), A, _capi.int3(vs),
            _capi.int2(ds), float(sid), float(sdd), int(device) if device is not None else -1,
            _stream(sino.device))
    _capi.check(rc, "JITStandardGeometry.backward")
    return out


def adjoint(sino, angles, volumeSize, detectorSize, sid, sdd, device):
    """Exact transpose of :func:`forward` (no reference counterpart)."""
    check_input(sino, "sino")   # argument order of the reference's CHECK_INPUTs
    angles = _angles(angles)
    vs, ds = _sizes(volumeSize, detectorSize)
    A = int(angles.numel())
    B = _check_volume(sino, "sino", A * ds[0] * ds[1])
    out = torch.empty((B, 1, vs[2], vs[1], vs[0]), dtype=torch.float32, device=sino.device)
    with torch.cuda.device(sino.device):
        rc = _capi.lib().dtv_standard_adjoint(
            sino.data_ptr(), out.data_ptr(), B, angles.data_ptr(), A, _capi.int3(vs),
            _capi.int2(ds), float(sid), float(sdd), int(device) if device is not None else -1,
            _stream(sino.device))
    _capi.check(rc, "JITStandardGeometry.adjoint")
    return out
