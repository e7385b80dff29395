"""Tensor-level projector calls through the C ABI.

Inputs are device tensors (ROCm ``torch.device('cuda', i)``), FP32 and
C-contiguous; outputs are allocated here with ``torch.empty`` on the input's
device and every element is written by the kernels.  Work is enqueued on the
current torch stream of that device.  There is no CPU path: a CPU tensor or
a missing library raises RuntimeError.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi
from .geometry import ConeGeometry, KIND_BEIJING, KIND_STANDARD


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def check_input(t: torch.Tensor, name: str) -> None:
    """CHECK_CUDA / CHECK_CONTIGUOUS of the reference (forwardKernel.cu:15-17)."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _check_f32(t: torch.Tensor, name: str) -> None:
    check_input(t, name)
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be a float32 tensor, got {t.dtype}")


def _batch_of(t: torch.Tensor, per_item: int, name: str) -> int:
    if t.dim() < 1 or t.shape[0] < 1:
        raise RuntimeError(f"{name} must have a leading batch dimension")
    b = int(t.shape[0])
    if t.numel() != b * per_item:
        raise RuntimeError(
            f"{name} has {t.numel()} elements, expected batch {b} x {per_item}")
    return b


def _vol_numel(geom: ConeGeometry) -> int:
    X, Y, Z = geom.volume_size
    return X * Y * Z


def _sino_numel(geom: ConeGeometry) -> int:
    U, V = geom.detector_size
    return geom.n_views * U * V


def _launch_arg(launch):
    """None -> NULL (library defaults); dict/tuple -> dtv_launch struct."""
    if launch is None:
        return None
    if isinstance(launch, _capi.DtvLaunch):
        return ctypes.byref(launch)
    if isinstance(launch, dict):
        return ctypes.byref(_capi.launch_struct(**launch))
    raise TypeError("launch must be None, a dict or a DtvLaunch")


def forward_plan(geom: ConeGeometry, batch: int = 1, launch=None):
    """The launch the deterministic :func:`forward` runs for ``geom`` and the
    workspace bytes it needs: ``(DtvLaunch, bytes)``.  No GPU work."""
    plan = _capi.DtvLaunch()
    nbytes = ctypes.c_size_t(0)
    rc = _capi.lib().dtv_cone_forward_plan(
        int(batch), _capi.int3(geom.volume_size), _capi.int2(geom.detector_size), geom.n_views,
        _launch_arg(launch), ctypes.byref(plan), ctypes.byref(nbytes))
    _capi.check(rc, "dtv_cone_forward_plan")
    return plan, int(nbytes.value)


def forward(volume: torch.Tensor, geom: ConeGeometry, launch=None,
            deterministic=None) -> torch.Tensor:
    """A: ray sums of ``volume`` [B,1,Z,Y,X] -> sinogram [B,1,A,V,U].

    ``launch`` optionally fixes the kernel (``dict(algorithm=, brick=,
    block=, view_parts=)``); every choice gives the same result to FP32
    rounding.  The brick kernels add per-brick partial sums with FP32
    atomics, so the last bits vary from run to run; ``deterministic=True``
    (the default when ``torch.use_deterministic_algorithms(True)`` is in
    force) adds them in int64 fixed point through a workspace from torch's
    allocator instead (``dtv_cone_forward_ws``): bitwise reproducible, 2-8%
    slower on large volumes."""
    _check_f32(volume, "volume")
    B = _batch_of(volume, _vol_numel(geom), "volume")
    out = torch.empty(geom.sino_shape(B), dtype=torch.float32, device=volume.device)
    views = geom.device_views(volume.device)
    sc = _capi.scheme_struct(geom.scheme)
    if deterministic is None:
        deterministic = torch.are_deterministic_algorithms_enabled()
    lib = _capi.lib()
    with torch.cuda.device(volume.device):
        if deterministic:
            plan, nbytes = forward_plan(geom, B, launch)
            ws = torch.empty(max(nbytes // 4, 1), dtype=torch.float32, device=volume.device)
            rc = lib.dtv_cone_forward_ws(
                volume.data_ptr(), out.data_ptr(), B, _capi.int3(geom.volume_size),
                _capi.int2(geom.detector_size), views.data_ptr(), geom.n_views,
                ctypes.byref(sc), _launch_arg(plan), ws.data_ptr(), nbytes,
                _stream(volume.device))
        else:
            rc = lib.dtv_cone_forward_ex(
                volume.data_ptr(), out.data_ptr(), B, _capi.int3(geom.volume_size),
                _capi.int2(geom.detector_size), views.data_ptr(), geom.n_views,
                ctypes.byref(sc), _launch_arg(launch), _stream(volume.device))
    _capi.check(rc, "dtv_cone_forward")
    return out


def adjoint_plan(geom: ConeGeometry, batch: int = 1, launch=None):
    """The launch :func:`adjoint` runs for ``geom`` (and an optional fixed
    ``launch``) with the view split it picks, and the workspace bytes it
    needs: ``(DtvLaunch, bytes)``.  No GPU work."""
    plan = _capi.DtvLaunch()
    nbytes = ctypes.c_size_t(0)
    rc = _capi.lib().dtv_cone_adjoint_plan(
        int(batch), _capi.int3(geom.volume_size), _capi.int2(geom.detector_size), geom.n_views,
        _launch_arg(launch), ctypes.byref(plan), ctypes.byref(nbytes))
    _capi.check(rc, "dtv_cone_adjoint_plan")
    return plan, int(nbytes.value)


def adjoint(sino: torch.Tensor, geom: ConeGeometry, launch=None,
            partition: bool = True) -> torch.Tensor:
    """Matched A^T: exact transpose of :func:`forward`.

    ``partition`` (default): ``dtv_cone_adjoint_ws`` with a workspace from
    torch's allocator.  Each sample is marched by the one brick holding its
    floor cell, and the bricks' upper-halo sums (and the partial volumes of a
    split beyond two parts) are added in a fixed order, so the result is
    bitwise reproducible for any split.  ``partition=False``:
    ``dtv_cone_adjoint_ex``, the unpartitioned bricks (each also marches the
    samples that only touch its halo) that the reference-shaped entry points
    use; their view split is at most two parts, which add into the output
    with atomics."""
    _check_f32(sino, "sino")
    B = _batch_of(sino, _sino_numel(geom), "sino")
    out = torch.empty(geom.volume_shape(B), dtype=torch.float32, device=sino.device)
    views = geom.device_views(sino.device)
    sc = _capi.scheme_struct(geom.scheme)
    lib = _capi.lib()
    L = launch
    ws_bytes = 0
    if partition:
        plan, ws_bytes = adjoint_plan(geom, B, L)
        if ws_bytes:
            L = plan
    with torch.cuda.device(sino.device):
        if ws_bytes:
            ws = torch.empty(ws_bytes // 4, dtype=torch.float32, device=sino.device)
            rc = lib.dtv_cone_adjoint_ws(
                sino.data_ptr(), out.data_ptr(), B, _capi.int3(geom.volume_size),
                _capi.int2(geom.detector_size), views.data_ptr(), geom.n_views,
                ctypes.byref(sc), _launch_arg(L), ws.data_ptr(), ws_bytes, _stream(sino.device))
        else:
            rc = lib.dtv_cone_adjoint_ex(
                sino.data_ptr(), out.data_ptr(), B, _capi.int3(geom.volume_size),
                _capi.int2(geom.detector_size), views.data_ptr(), geom.n_views,
                ctypes.byref(sc), _launch_arg(L), _stream(sino.device))
    _capi.check(rc, "dtv_cone_adjoint")
    return out


def _ref_angles(geom: ConeGeometry, device):
    key = ("angles", str(device))
    t = geom._dev_cache.get(key)
    if t is None:
        t = torch.from_numpy(np.ascontiguousarray(geom.ref["angles"], dtype=np.float64)).to(device)
        geom._dev_cache[key] = t
    return t


def _ref_vectors(geom: ConeGeometry, device):
    key = ("vectors", str(device))
    t = geom._dev_cache.get(key)
    if t is None:
        t = torch.from_numpy(np.ascontiguousarray(geom.ref["vectors"], dtype=np.float32)).to(device)
        geom._dev_cache[key] = t
    return t


def bp_ref(sino: torch.Tensor, geom: ConeGeometry) -> torch.Tensor:
    """Reference-mode backprojector (FDK-style weights, not the adjoint):
    Standard backwardKernel.cu:22-53 or Beijing backwardKernel.cu:18-45."""
    _check_f32(sino, "sino")
    B = _batch_of(sino, _sino_numel(geom), "sino")
    out = torch.empty(geom.volume_shape(B), dtype=torch.float32, device=sino.device)
    dev = sino.device
    with torch.cuda.device(dev):
        if geom.kind == KIND_STANDARD:
            rc = _capi.lib().dtv_standard_backward(
                sino.data_ptr(), out.data_ptr(), B, _ref_angles(geom, dev).data_ptr(),
                geom.n_views, _capi.int3(geom.volume_size), _capi.int2(geom.detector_size),
                geom.ref["sid"], geom.ref["sdd"], -1, _stream(dev))
            _capi.check(rc, "dtv_standard_backward")
        elif geom.kind == KIND_BEIJING:
            rc = _capi.lib().dtv_beijing_backward(
                sino.data_ptr(), out.data_ptr(), B, _capi.int3(geom.volume_size),
                _capi.int2(geom.detector_size), _ref_vectors(geom, dev).data_ptr(),
                geom.n_views, -1, _stream(dev))
            _capi.check(rc, "dtv_beijing_backward")
        else:
            raise RuntimeError("the reference backprojector exists only for the Standard "
                               "and Beijing geometries; use adjoint() for table geometries")
    return out


def host_views_standard(angles, volume_size, detector_size, sid, sdd):
    """Canonical view table computed by the library's host code (no GPU)."""
    a = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros((a.shape[0], 12), dtype=np.float64)
    sc = _capi.DtvScheme()
    rc = _capi.lib().dtv_views_standard(
        a.ctypes.data_as(_capi.c_double_p), a.shape[0], _capi.int3(volume_size),
        _capi.int2(detector_size), float(sid), float(sdd),
        out.ctypes.data_as(_capi.c_double_p), ctypes.byref(sc))
    _capi.check(rc, "dtv_views_standard")
    return out, sc


def host_views_beijing(project_vector, volume_size, detector_size):
    P = np.ascontiguousarray(project_vector, dtype=np.float32)
    out = np.zeros((P.shape[0], 12), dtype=np.float64)
    sc = _capi.DtvScheme()
    rc = _capi.lib().dtv_views_beijing(
        P.ctypes.data_as(_capi.c_float_p), P.shape[0], _capi.int3(volume_size),
        _capi.int2(detector_size), out.ctypes.data_as(_capi.c_double_p), ctypes.byref(sc))
    _capi.check(rc, "dtv_views_beijing")
    return out, sc
