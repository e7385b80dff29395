"""Directional finite differences, total variation and the fused update step.

Torch operators (differentiable, any device) restate the reference's sparse
difference matrices (``model/PDHG/Utils/Gradients.py:8-47``) without building
them: for C-order flattened volumes (x fastest) D_a f = f[i + s_a] - f[i]
with

* ``mode="reference"``: s = (1, Y, X*Y) and the reference's rows -- D_x has
  no +1 where (i+1) % X == 0, D_y strides by Y (= volumeSize[1]) over the
  whole flat array (crossing slice ends), D_z stops at the last slice; rows
  without a +1 give -f[i];
* ``mode="neumann"``: conventional forward differences, zero at the far
  boundary of each axis.

``fista_tv_step`` runs the fused gfx950 kernel that, per voxel, evaluates the
gradient of the smoothed total variation TV_eps(y) = sum_c sqrt(|D y|_c^2 +
eps^2), takes the step x = y - L (g + lam grad TV_eps(y)), clamps at zero if
asked and applies the momentum y' = x + beta (x - x_prev) -- one HBM pass.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi
from .ops import _stream, check_input

MODES = {"neumann": 0, "reference": 1}


def _flat(f):
    shp = f.shape
    return f.reshape(*shp[:-3], shp[-3] * shp[-2] * shp[-1]), shp


def D(f: torch.Tensor, axis: str, mode: str = "neumann") -> torch.Tensor:
    """Forward difference along ``axis`` of f [..., Z, Y, X]."""
    Z, Y, X = f.shape[-3:]
    if mode == "neumann":
        dim = {"x": -1, "y": -2, "z": -3}[axis]
        d = torch.diff(f, dim=dim)
        pad = [0, 0] * 3
        pad[2 * (-dim - 1) + 1] = 1
        return torch.nn.functional.pad(d, pad)
    if mode != "reference":
        raise ValueError(mode)
    ff, shp = _flat(f)
    N = ff.shape[-1]
    out = -ff
    if axis == "x":
        nxt = torch.zeros_like(ff)
        nxt[..., :-1] = ff[..., 1:]
        mask = (torch.arange(N, device=f.device) + 1) % X != 0
        out = out + nxt * mask
    elif axis == "y":
        out = torch.cat([ff[..., Y:] - ff[..., :N - Y], -ff[..., N - Y:]], -1)
    elif axis == "z":
        s = X * Y
        out = torch.cat([ff[..., s:] - ff[..., :N - s], -ff[..., N - s:]], -1)
    else:
        raise ValueError(axis)
    return out.reshape(shp)


def DT(p: torch.Tensor, axis: str, mode: str = "neumann") -> torch.Tensor:
    """Transpose of :func:`D` applied to p [..., Z, Y, X]."""
    Z, Y, X = p.shape[-3:]
    if mode == "neumann":
        dim = {"x": -1, "y": -2, "z": -3}[axis]
        n = p.shape[dim]
        head = p.narrow(dim, 0, n - 1)
        out = torch.nn.functional.pad(head, _pad_for(dim, front=1)) - \
            torch.nn.functional.pad(head, _pad_for(dim, back=1))
        return out
    pp, shp = _flat(p)
    N = pp.shape[-1]
    out = -pp
    if axis == "x":
        mask = (torch.arange(N - 1, device=p.device) + 1) % X != 0
        prev = torch.zeros_like(pp)
        prev[..., 1:] = pp[..., :-1] * mask
        out = out + prev
    elif axis == "y":
        out = torch.cat([-pp[..., :Y], pp[..., :N - Y] - pp[..., Y:]], -1)
    elif axis == "z":
        s = X * Y
        out = torch.cat([-pp[..., :s], pp[..., :N - s] - pp[..., s:]], -1)
    else:
        raise ValueError(axis)
    return out.reshape(shp)


def _pad_for(dim, front=0, back=0):
    pad = [0, 0] * 3
    k = -dim - 1
    pad[2 * k] = front
    pad[2 * k + 1] = back
    return pad


def grad3(f, mode="neumann"):
    return D(f, "x", mode), D(f, "y", mode), D(f, "z", mode)


def tv_eps(f, eps, mode="neumann"):
    dx, dy, dz = grad3(f, mode)
    return torch.sqrt(dx * dx + dy * dy + dz * dz + eps * eps).sum()


def tv_eps_grad(f, eps, mode="neumann"):
    """Gradient of :func:`tv_eps` (explicit, no autograd)."""
    dx, dy, dz = grad3(f, mode)
    n = torch.sqrt(dx * dx + dy * dy + dz * dz + eps * eps)
    return DT(dx / n, "x", mode) + DT(dy / n, "y", mode) + DT(dz / n, "z", mode)


def asd_potential(f, eps=1.0, mode="reference"):
    """sum_c [r_c - eps log(1 + r_c / eps)] with r = |D f| per voxel: the
    functional whose gradient is the TV direction of ASD-POCS."""
    dx, dy, dz = grad3(f, mode)
    r = torch.sqrt(dx * dx + dy * dy + dz * dz)
    return (r - eps * torch.log1p(r / eps)).sum()


def asd_tv_grad(f, eps=1.0, mode="reference"):
    """sum_a D_a^T (D_a f / (|D f| + eps)): gradTV of model/PDHG/ASDPOCS.py:38-42
    (there eps = 1 and the matrices are Gradients.py:8-47, ``mode="reference"``)."""
    dx, dy, dz = grad3(f, mode)
    n = torch.sqrt(dx * dx + dy * dy + dz * dz) + eps
    return DT(dx / n, "x", mode) + DT(dy / n, "y", mode) + DT(dz / n, "z", mode)


def tv_descent_torch(f, step, eps=1.0, mode="reference"):
    """Unfused torch formulation of :func:`tv_descent`."""
    return f - step * asd_tv_grad(f, eps, mode)


def tv_descent(f, step, eps=1.0, mode="reference", out=None):
    """One ASD-POCS TV step f - step * asd_tv_grad(f) on the fused gfx950
    kernel (``dtv_tv_descent``).  f [B, 1, Z, Y, X] FP32 on the GPU; ``out``
    (optional) must not share storage with f."""
    check_input(f, "f")
    if f.dtype != torch.float32:
        raise RuntimeError("f must be float32")
    out = torch.empty_like(f) if out is None else out
    check_input(out, "out")
    if out.shape != f.shape or out.dtype != torch.float32:
        raise RuntimeError("out must be float32 with the shape of f")
    if out.data_ptr() == f.data_ptr():
        raise RuntimeError("out must not share storage with f (stencil reads)")
    _same_device(f, out)
    Z, Y, X = f.shape[-3:]
    B = f.numel() // (X * Y * Z)
    with torch.cuda.device(f.device):
        rc = _capi.lib().dtv_tv_descent(f.data_ptr(), out.data_ptr(), B, _capi.int3((X, Y, Z)),
                                        ctypes.c_float(step), ctypes.c_float(eps), MODES[mode],
                                        _stream(f.device))
    _capi.check(rc, "dtv_tv_descent")
    return out


def pdhg_dtv_step_torch(f, fbar, atw, p, sigma, tau, lam, mode="neumann"):
    """One primal-dual step of the directional-TV unroll after A and A^T, as
    torch ops (differentiable): p_a = clamp(p_a + sigma D_a fbar, -lam_a,
    lam_a); f' = relu(f - tau (atw + sum_a D_a^T p_a)); returns
    (f', 2 f' - f, [p_x, p_y, p_z])."""
    axes = ("x", "y", "z")
    p = [torch.clamp(p[i] + sigma * D(fbar, a, mode), -lam[i], lam[i])
         for i, a in enumerate(axes)]
    grad = atw + DT(p[0], "x", mode) + DT(p[1], "y", mode) + DT(p[2], "z", mode)
    fn = torch.nn.functional.relu(f - tau * grad)
    return fn, 2 * fn - f, p


def pdhg_dtv_step(f, fbar, atw, p, sigma, tau, lam, mode="neumann"):
    """:func:`pdhg_dtv_step_torch` in place on the fused gfx950 kernels
    (``dtv_pdhg_dtv_step``): f, fbar [B, 1, Z, Y, X] and p [3, B, 1, Z, Y, X]
    FP32 on the GPU are overwritten with f', 2 f' - f and the new p; atw is
    read.  The four tensors must be distinct and contiguous."""
    ts = {"f": f, "fbar": fbar, "atw": atw, "p": p}
    for name, t in ts.items():
        check_input(t, name)
        if t.dtype != torch.float32:
            raise RuntimeError(f"{name} must be float32")
    if fbar.shape != f.shape or atw.shape != f.shape or p.shape != (3,) + tuple(f.shape):
        raise RuntimeError("fbar and atw must have the shape of f, p the shape [3, *f.shape]")
    if len({t.untyped_storage().data_ptr() for t in ts.values()}) != 4:
        raise RuntimeError("f, fbar, atw and p must not share storage")
    _same_device(f, fbar, atw, p)
    Z, Y, X = f.shape[-3:]
    B = f.numel() // (X * Y * Z)
    lam3 = (ctypes.c_float * 3)(*[float(v) for v in lam])
    with torch.cuda.device(f.device):
        rc = _capi.lib().dtv_pdhg_dtv_step(f.data_ptr(), fbar.data_ptr(), atw.data_ptr(),
                                           p.data_ptr(), B, _capi.int3((X, Y, Z)),
                                           ctypes.c_float(sigma), ctypes.c_float(tau), lam3,
                                           MODES[mode], _stream(f.device))
    _capi.check(rc, "dtv_pdhg_dtv_step")
    return f, fbar, p


class L1Item(ctypes.Structure):
    """One vector of a batched l1-ball projection (mirrors dtv_l1_item)."""
    _fields_ = [("v", ctypes.c_void_p), ("out", ctypes.c_void_p), ("n", ctypes.c_int64),
                ("radius", ctypes.c_double), ("sigma", ctypes.c_float),
                ("reserved", ctypes.c_int32)]


L1_MAX_ITEMS = 16


def _l1ball_device(vs, radii, outs, mode, sigma=1.0, tol=1e-4, max_iter=100):
    """Run ``dtv_l1ball`` over the (contiguous FP32 CUDA) tensors ``vs``."""
    n = len(vs)
    if not 1 <= n <= L1_MAX_ITEMS:
        raise RuntimeError(f"l1ball takes 1..{L1_MAX_ITEMS} vectors per call")
    dev = vs[0].device
    items = (L1Item * n)()
    for k, (v, o, b) in enumerate(zip(vs, outs, radii)):
        check_input(v, "v")
        check_input(o, "out")
        if v.dtype != torch.float32 or o.dtype != torch.float32 or o.numel() != v.numel():
            raise RuntimeError("l1ball takes float32 tensors, out of the size of v")
        if v.device != dev or o.device != dev:
            raise RuntimeError("l1ball vectors must be on one device")
        items[k].v = v.data_ptr()
        items[k].out = o.data_ptr()
        items[k].n = v.numel()
        items[k].radius = float(b)
        items[k].sigma = float(sigma)
    L = _capi.lib()
    nbytes = L.dtv_l1ball_workspace_bytes(n)
    ws = torch.empty(nbytes // 8 + 1, dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        rc = L.dtv_l1ball(ctypes.addressof(items), n, mode, float(tol), int(max_iter),
                          ws.data_ptr(), ws.numel() * 8, _stream(dev))
    _capi.check(rc, "dtv_l1ball")
    return outs


def l1ball(v, b, tol=1e-4, max_iter=100, out=None):
    """Shrink v onto the l1 ball of radius b as model/PDHG/Utils/L1ball.py:4-17
    does: if sum|v| < b, v itself; else the threshold lam solving
    sum relu(|v| - lam) = b by the reference's damped Newton iteration from
    lam = 0 (slope -count(|v| > lam) + 0.001, stop once |objective| <= tol or
    after max_iter steps), then sign(v) relu(|v| - lam).

    Runs on the fused gfx950 kernels (``dtv_l1ball``: the whole iteration on
    the GPU, no host synchronisation); ``v`` is a contiguous FP32 CUDA
    tensor.  :func:`l1ball_torch` is the same iteration as torch ops."""
    out = torch.empty_like(v) if out is None else out
    return _l1ball_device([v], [b], [out], 0, tol=tol, max_iter=max_iter)[0]


def l1ball_dual(ps, sigma, radii, tol=1e-4, max_iter=100):
    """The dual update of model/PDHG/DTVCP.py:44-46 for every tensor in ``ps``
    (up to 16, updated in place): p <- p - sigma sign(p) l1ball(|p| / sigma,
    b), one ``dtv_l1ball`` call (mode 1) for all of them."""
    _l1ball_device(list(ps), list(radii), list(ps), 1, sigma=sigma, tol=tol,
                   max_iter=max_iter)
    return ps


def l1ball_torch(v, b, tol=1e-4, max_iter=100):
    """:func:`l1ball` as torch ops on v's device, lam a 0-d FP64 tensor (the
    reference keeps it as a Python float), with a convergence check on the
    host every 8 steps (the reference synchronises on every step)."""
    a = v.abs()
    total = a.sum()
    if float(total) < b:
        return v
    lam = torch.zeros((), dtype=torch.float64, device=v.device)
    obj = total.double() - b
    active = obj.abs() > tol
    for it in range(max_iter):
        if it % 8 == 0 and not bool(active):
            break
        r = torch.clamp(a - lam.to(a.dtype), min=0)
        obj = r.sum().double() - b
        slope = 0.001 - (r > 0).sum().double()
        lam = torch.where(active, lam - obj / slope, lam)
        active = active & (obj.abs() > tol)
    return torch.sign(v) * torch.clamp(a - lam.to(a.dtype), min=0)


def _check_outputs(like, *outs):
    """Each (tensor, name) in ``outs``: a contiguous FP32 CUDA tensor of the
    shape of ``like`` (the kernels write ``like.numel()`` elements)."""
    for o, n in outs:
        check_input(o, n)
        if o.dtype != torch.float32 or o.shape != like.shape:
            raise RuntimeError(f"{n} must be float32 with the shape {tuple(like.shape)}")


def _same_device(*ts):
    dev = ts[0].device
    if any(t.device != dev for t in ts):
        raise RuntimeError("all tensors must be on one device")


def shrink(v, thr):
    """Soft threshold sign(v) max(|v| - thr, 0)."""
    return torch.sign(v) * torch.clamp(v.abs() - thr, min=0)


def shrink_step(y, g, xprev, step, thr, beta, xnew=None, ynew=None):
    """Fused soft-threshold FISTA step (``dtv_shrink_step``, model/FISTA/FISTA.py:21-25):
    d = y - step g; x = sign(d) relu(|d| - thr); y' = x + beta (x - xprev).
    Returns (x, y'); xnew may be xprev (updated in place)."""
    for t, n in ((y, "y"), (g, "g"), (xprev, "xprev")):
        check_input(t, n)
        if t.dtype != torch.float32 or t.shape != y.shape:
            raise RuntimeError(f"{n} must be float32 with the shape of y")
    xnew = torch.empty_like(y) if xnew is None else xnew
    ynew = torch.empty_like(y) if ynew is None else ynew
    _check_outputs(y, (xnew, "xnew"), (ynew, "ynew"))
    _same_device(y, g, xprev, xnew, ynew)
    if ynew.data_ptr() == xprev.data_ptr():
        raise RuntimeError("ynew must not share storage with xprev")
    if ynew.data_ptr() == xnew.data_ptr():
        raise RuntimeError("xnew and ynew must be distinct tensors")
    with torch.cuda.device(y.device):
        rc = _capi.lib().dtv_shrink_step(y.data_ptr(), g.data_ptr(), xprev.data_ptr(),
                                         xnew.data_ptr(), ynew.data_ptr(), y.numel(),
                                         float(step), float(thr), float(beta), _stream(y.device))
    _capi.check(rc, "dtv_shrink_step")
    return xnew, ynew


def shrink_step_torch(y, g, xprev, step, thr, beta):
    """Unfused torch formulation of :func:`shrink_step`."""
    x = shrink(y - step * g, thr)
    return x, x + beta * (x - xprev)


def fdpg_step(f, g, sigma, v, t, beta, mode="reference", out=None):
    """One fused step of the directional-TV dual-gradient unroll
    (``dtv_fdpg_step``, model/FISTA/DTVFISTA.py:29-41) after g = A^T(A f - b):
    z = relu(f - sigma g); m_a = max(|D_a z| - v_a) over the whole batch;
    f_new = z + sum_a t_a D_a^T (sign(D_a z) m_a); returns
    f_new + beta (f_new - f) (a new tensor, or ``out``)."""
    for x, n in ((f, "f"), (g, "g")):
        check_input(x, n)
        if x.dtype != torch.float32 or x.shape != f.shape:
            raise RuntimeError(f"{n} must be float32 with the shape of f")
    out = torch.empty_like(f) if out is None else out
    _check_outputs(f, (out, "out"))
    _same_device(f, g, out)
    if out.data_ptr() in (f.data_ptr(), g.data_ptr()):
        raise RuntimeError("out must not share storage with f or g")
    Z, Y, X = f.shape[-3:]
    B = f.numel() // (X * Y * Z)
    L = _capi.lib()
    ws = torch.empty(L.dtv_fdpg_workspace_bytes() // 4 + 1, dtype=torch.int32, device=f.device)
    v3 = (ctypes.c_float * 3)(*[float(x) for x in v])
    t3 = (ctypes.c_float * 3)(*[float(x) for x in t])
    with torch.cuda.device(f.device):
        rc = L.dtv_fdpg_step(f.data_ptr(), g.data_ptr(), out.data_ptr(), B, _capi.int3((X, Y, Z)),
                             float(sigma), v3, t3, float(beta), MODES[mode], ws.data_ptr(),
                             ws.numel() * 4, _stream(f.device))
    _capi.check(rc, "dtv_fdpg_step")
    return out


def fdpg_step_torch(f, g, sigma, v, t, beta, mode="reference"):
    """Unfused torch formulation of :func:`fdpg_step` (the reference's
    statements with D_a / D_a^T in place of the sparse matrices)."""
    z = torch.relu(f - sigma * g)
    axes = ("x", "y", "z")
    fnew = z
    for i, a in enumerate(axes):
        d = D(z, a, mode)
        p = torch.sign(d) * torch.max(d.abs() - v[i])
        fnew = fnew + t[i] * DT(p, a, mode)
    return fnew + beta * (fnew - f)


# --------------------------------------------------------------------------
# Fused gfx950 step
# --------------------------------------------------------------------------

class TvItem(ctypes.Structure):
    """One instance of a batched fused step (mirrors dtv_tv_item)."""
    _fields_ = [
        ("y", ctypes.c_void_p), ("g", ctypes.c_void_p), ("xprev", ctypes.c_void_p),
        ("xnew", ctypes.c_void_p), ("ynew", ctypes.c_void_p),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("nonneg", ctypes.c_int32),
        ("step", ctypes.c_float), ("lam", ctypes.c_float), ("eps", ctypes.c_float),
        ("beta", ctypes.c_float),
    ]


def fista_tv_step(y, g, xprev, step, lam, eps, beta, nonneg=True, mode="neumann",
                  xnew=None, ynew=None):
    """x = [y - step (g + lam grad TV_eps(y))]_+ ; y' = x + beta (x - xprev).

    All volumes [B, 1, Z, Y, X] FP32 on the GPU; returns (x, y')."""
    for t, n in ((y, "y"), (g, "g"), (xprev, "xprev")):
        check_input(t, n)
        if t.dtype != torch.float32 or t.shape != y.shape:
            raise RuntimeError(f"{n} must be float32 with the shape of y")
    xnew = torch.empty_like(y) if xnew is None else xnew
    ynew = torch.empty_like(y) if ynew is None else ynew
    for o, n in ((xnew, "xnew"), (ynew, "ynew")):
        check_input(o, n)
        if o.shape != y.shape or o.dtype != torch.float32:
            raise RuntimeError(f"{n} must be float32 with the shape of y")
        if o.data_ptr() == y.data_ptr() or o.data_ptr() == g.data_ptr():
            raise RuntimeError(f"{n} must not share storage with y or g (stencil reads)")
    if xnew.data_ptr() == ynew.data_ptr():
        raise RuntimeError("xnew and ynew must be distinct tensors")
    _same_device(y, g, xprev, xnew, ynew)
    Z, Y, X = y.shape[-3:]
    B = y.numel() // (X * Y * Z)
    with torch.cuda.device(y.device):
        rc = _capi.lib().dtv_tv_fista_step(
            y.data_ptr(), g.data_ptr(), xprev.data_ptr(), xnew.data_ptr(), ynew.data_ptr(), B,
            _capi.int3((X, Y, Z)), ctypes.c_float(step), ctypes.c_float(lam),
            ctypes.c_float(eps), ctypes.c_float(beta), int(bool(nonneg)), MODES[mode],
            _stream(y.device))
    _capi.check(rc, "dtv_tv_fista_step")
    return xnew, ynew


def fista_tv_step_torch(y, g, xprev, step, lam, eps, beta, nonneg=True, mode="neumann"):
    """Unfused torch formulation of :func:`fista_tv_step` (the plain path)."""
    x = y - step * (g + lam * tv_eps_grad(y, eps, mode))
    if nonneg:
        x = torch.clamp(x, min=0)
    return x, x + beta * (x - xprev)
