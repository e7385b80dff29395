"""FP64 PyTorch-CPU restatement of the reference cone-beam projectors.

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker for the
HIP kernels and the CPU baseline, never imported by the product.

Each front-end restates one reference kernel *in that kernel's own units*:

* ``std_*``  -- Standard circular geometry, world unit = one detector pixel
  (``model/ConeBeamLayers/Standard/plug/kernel/forwardKernel.cu:22-86`` for A,
  ``backwardKernel.cu:22-53`` + ``include/helper_geometry.h:26-42`` for the
  reference backprojector).
* ``bj_*``   -- Beijing projection vectors (``Beijing/plug/kernel/
  forwardKernel.cu:17-76``, ``backwardKernel.cu:18-45``).
* ``circ_*`` -- the physical-unit circular geometry (mm) of the benchmark
  configurations; it follows the Beijing sampling rule in world mm.
* ``table_*`` -- an arbitrary canonical view table (voxel frame).

Sampling (all front-ends): samples at alpha_k = alpha_start + k*step (exact
products, where the reference accumulates ``alpha += step`` in FP32), value =
trilinear interpolation with texel centres at integer voxel coordinates and a
zero border (``tex3D(p + 0.5)`` with border addressing and linear filtering).
The *matched adjoint* is the exact transpose of the same enumeration.

``dtype`` may be torch.float32 to obtain the FP32 restatement used as the
CPU baseline.
"""
from __future__ import annotations

import math

import numpy as np
import torch

PI_REF = 3.14159265359   # forwardKernel.cu:14 ``#define PI 3.14159265359``
F64 = torch.float64


class Rays:
    """Rays in a native unit system.

    Sample positions in voxel coordinates are
    p(alpha) = (src + alpha*d) / div_s + off, alpha_k = a_start + k*step,
    k = 0, 1, ... while alpha_k < a_end.  The ray sum is multiplied by
    ``scale``.  Arrays are flattened over rays in sinogram order
    [view, v, u].
    """

    def __init__(self, src, d, a_start, a_end, step, div_s, off, scale):
        self.src = src.reshape(-1, 3)
        self.d = d.reshape(-1, 3)
        self.a_start = a_start.reshape(-1)
        self.a_end = a_end.reshape(-1)
        self.step = step
        self.div_s = div_s
        self.off = off
        self.scale = scale

    @property
    def n(self):
        return self.a_start.shape[0]


def _norm3(x, y, z):
    n = torch.sqrt(x * x + y * y + z * z)
    return x / n, y / n, z / n


# --------------------------------------------------------------------------
# Ray set-up, one per reference kernel
# --------------------------------------------------------------------------

def std_rays(angles, volume_size, detector_size, sid, sdd, dtype=F64):
    """Standard forwardKernel.cu:22-86 in world units (1 unit = 1 pixel)."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    a = torch.as_tensor(np.asarray(angles, dtype=np.float64), dtype=dtype)
    theta = a / 180 * PI_REF - PI_REF / 2                     # :30
    exx, exy = torch.cos(theta), torch.sin(theta)             # :31
    eyx, eyy = -exy, exx                                      # :32
    detX = torch.arange(U, dtype=dtype) + (-U / 2.0)          # :33
    detY = torch.arange(V, dtype=dtype) + (-V / 2.0)          # :34
    A = a.shape[0]
    wx = exx.view(A, 1, 1) * sdd + eyx.view(A, 1, 1) * detX.view(1, 1, U)   # :37
    wy = exy.view(A, 1, 1) * sdd + eyy.view(A, 1, 1) * detX.view(1, 1, U)
    rz = detY.view(1, V, 1)
    wx, wy, rz = torch.broadcast_tensors(wx, wy, rz)
    dx, dy, dz = _norm3(wx, wy, rz)                           # :37-39
    sx = (-exx * sid).view(A, 1, 1).expand(A, V, U)           # :40
    sy = (-exy * sid).view(A, 1, 1).expand(A, V, U)
    sz = torch.zeros_like(sx)
    use_x = dx.abs() >= dy.abs()                              # :45
    a0 = torch.where(use_x, (-X / 2.0 - sx) / dx, (-Y / 2.0 - sy) / dy)   # :46-56
    a1 = torch.where(use_x, (X / 2.0 - sx) / dx, (Y / 2.0 - sy) / dy)
    a_start = torch.minimum(a0, a1) - 3                       # :57
    a_end = torch.maximum(a0, a1) + 3                         # :58
    s = sid / sdd                                             # :63
    src = torch.stack([sx, sy, sz], -1)
    d = torch.stack([dx, dy, dz], -1)
    off = torch.tensor([X / 2.0, Y / 2.0, Z / 2.0], dtype=dtype)   # p -= volumeCenter
    return Rays(src, d, a_start, a_end, s, s, off, s)


def bj_rays(project_vector, volume_size, detector_size, dtype=F64):
    """Beijing forwardKernel.cu:17-76 in kernel voxel units."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    P = torch.as_tensor(np.asarray(project_vector, dtype=np.float32)).to(dtype)
    n = P.shape[0]
    S = torch.stack([P[:, 0], P[:, 1], -P[:, 2]], -1)        # :27
    D = torch.stack([P[:, 3], P[:, 4], -P[:, 5]], -1)        # :28
    v = P[:, 6:9]                                             # :29
    u = P[:, 9:12]                                            # :30
    detX = torch.arange(U, dtype=dtype) + (-U / 2.0)          # :24
    detY = torch.arange(V, dtype=dtype) + (-V / 2.0)          # :25
    pix = (D.view(n, 1, 1, 3) + (0.5 + detX).view(1, 1, U, 1) * u.view(n, 1, 1, 3)
           + (0.5 + detY).view(1, V, 1, 1) * v.view(n, 1, 1, 3))   # :33
    r = pix - S.view(n, 1, 1, 3)
    dx, dy, dz = _norm3(r[..., 0], r[..., 1], r[..., 2])      # :35
    ax_, ay_, az_ = dx.abs(), dy.abs(), dz.abs()
    m = torch.maximum(ax_, torch.maximum(az_, ay_))           # :40
    use_x = ax_ == m                                          # :41
    use_y = (~use_x) & (ay_ == m)                             # :47
    Sx = S[:, 0].view(n, 1, 1).expand(n, V, U)
    Sy = S[:, 1].view(n, 1, 1).expand(n, V, U)
    Sz = S[:, 2].view(n, 1, 1).expand(n, V, U)
    a0 = torch.where(use_x, (-X / 2.0 - Sx) / dx,
                     torch.where(use_y, (-Y / 2.0 - Sy) / dy, (-Z / 2.0 - Sz) / dz))
    a1 = torch.where(use_x, (X / 2.0 - Sx) / dx,
                     torch.where(use_y, (Y / 2.0 - Sy) / dy, (Z / 2.0 - Sz) / dz))
    a_start = torch.minimum(a0, a1) - 3                       # :59
    a_end = torch.maximum(a0, a1) + 3                         # :60
    src = torch.stack([Sx, Sy, Sz], -1)
    d = torch.stack([dx, dy, dz], -1)
    off = torch.tensor([X / 2.0, Y / 2.0, Z / 2.0], dtype=dtype)
    return Rays(src, d, a_start, a_end, 1.0, 1.0, off, 1.0)


def circ_rays(angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd,
              dtype=F64):
    """Physical-unit circular geometry (mm), Beijing sampling rule."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    vx = float(voxel_mm)
    if np.ndim(pitch_mm) == 0:
        du = dv = float(pitch_mm)
    else:
        du, dv = (float(p) for p in pitch_mm)
    a = torch.as_tensor(np.asarray(angles, dtype=np.float64), dtype=dtype)
    theta = a / 180 * PI_REF - PI_REF / 2
    c, sn = torch.cos(theta), torch.sin(theta)
    A = a.shape[0]
    ex = torch.stack([c, sn, torch.zeros_like(c)], -1).view(A, 1, 1, 3)
    ey = torch.stack([-sn, c, torch.zeros_like(c)], -1).view(A, 1, 1, 3)
    ez = torch.tensor([0.0, 0.0, 1.0], dtype=dtype).view(1, 1, 1, 3)
    src = -ex * sod
    centre = ex * (sdd - sod)
    tu = ((torch.arange(U, dtype=dtype) + 0.5 - U / 2.0) * du).view(1, 1, U, 1)
    tv = ((torch.arange(V, dtype=dtype) + 0.5 - V / 2.0) * dv).view(1, V, 1, 1)
    pos = centre + ey * tu + ez * tv
    r = pos - src
    dx, dy, dz = _norm3(r[..., 0], r[..., 1], r[..., 2])
    ax_, ay_, az_ = dx.abs(), dy.abs(), dz.abs()
    m = torch.maximum(ax_, torch.maximum(az_, ay_))
    use_x = ax_ == m
    use_y = (~use_x) & (ay_ == m)
    srcb = src.expand(A, V, U, 3)
    Sx, Sy, Sz = srcb[..., 0], srcb[..., 1], srcb[..., 2]
    hx, hy, hz = X * vx / 2.0, Y * vx / 2.0, Z * vx / 2.0
    a0 = torch.where(use_x, (-hx - Sx) / dx, torch.where(use_y, (-hy - Sy) / dy, (-hz - Sz) / dz))
    a1 = torch.where(use_x, (hx - Sx) / dx, torch.where(use_y, (hy - Sy) / dy, (hz - Sz) / dz))
    a_start = torch.minimum(a0, a1) - 3 * vx
    a_end = torch.maximum(a0, a1) + 3 * vx
    d = torch.stack([dx, dy, dz], -1)
    off = torch.tensor([X / 2.0 - 0.5, Y / 2.0 - 0.5, Z / 2.0 - 0.5], dtype=dtype)
    return Rays(srcb.contiguous(), d, a_start, a_end, vx, vx, off, vx)


def table_rays(views, detector_size, axis_mode, box_lo, box_hi, margin, scale,
               dtype=F64):
    """Canonical voxel-frame view table [n, 12] = (S, D0, U, V)."""
    U, V = (int(v) for v in detector_size)
    T = torch.as_tensor(np.asarray(views, dtype=np.float64)).to(dtype)
    n = T.shape[0]
    S, D0, Uv, Vv = T[:, 0:3], T[:, 3:6], T[:, 6:9], T[:, 9:12]
    iu = torch.arange(U, dtype=dtype).view(1, 1, U, 1)
    iv = torch.arange(V, dtype=dtype).view(1, V, 1, 1)
    pix = D0.view(n, 1, 1, 3) + iu * Uv.view(n, 1, 1, 3) + iv * Vv.view(n, 1, 1, 3)
    r = pix - S.view(n, 1, 1, 3)
    dx, dy, dz = _norm3(r[..., 0], r[..., 1], r[..., 2])
    ax_, ay_, az_ = dx.abs(), dy.abs(), dz.abs()
    if axis_mode == 2:
        use_x = ax_ >= ay_
        use_y = ~use_x
    else:
        m = torch.maximum(ax_, torch.maximum(az_, ay_))
        use_x = ax_ == m
        use_y = (~use_x) & (ay_ == m)
    Sb = S.view(n, 1, 1, 3).expand(n, V, U, 3)
    Sx, Sy, Sz = Sb[..., 0], Sb[..., 1], Sb[..., 2]
    lo, hi = box_lo, box_hi
    a0 = torch.where(use_x, (lo[0] - Sx) / dx, torch.where(use_y, (lo[1] - Sy) / dy, (lo[2] - Sz) / dz))
    a1 = torch.where(use_x, (hi[0] - Sx) / dx, torch.where(use_y, (hi[1] - Sy) / dy, (hi[2] - Sz) / dz))
    a_start = torch.minimum(a0, a1) - margin
    a_end = torch.maximum(a0, a1) + margin
    d = torch.stack([dx, dy, dz], -1)
    off = torch.zeros(3, dtype=dtype)
    return Rays(Sb.contiguous(), d, a_start, a_end, 1.0, 1.0, off, scale)


# --------------------------------------------------------------------------
# Trilinear sample enumeration
# --------------------------------------------------------------------------

def _k_window(rays, sl, N):
    """Conservative [kmin, kmax] per ray covering every sample whose
    trilinear support can touch the volume (p in [-1, N] on every axis)."""
    src = rays.src[sl]
    d = rays.d[sl]
    a_start = rays.a_start[sl]
    dtype = src.dtype
    lo_k = torch.full_like(a_start, -math.inf)
    hi_k = torch.full_like(a_start, math.inf)
    for ax in range(3):
        # alpha where p_ax = -1 and p_ax = N_ax
        e0 = ((-1.0 - rays.off[ax]) * rays.div_s - src[:, ax])
        e1 = ((N[ax] - rays.off[ax]) * rays.div_s - src[:, ax])
        dd = d[:, ax]
        nz = dd != 0
        t0 = torch.where(nz, e0 / torch.where(nz, dd, torch.ones_like(dd)), torch.zeros_like(dd))
        t1 = torch.where(nz, e1 / torch.where(nz, dd, torch.ones_like(dd)), torch.zeros_like(dd))
        k0 = (torch.minimum(t0, t1) - a_start) / rays.step
        k1 = (torch.maximum(t0, t1) - a_start) / rays.step
        # a ray parallel to this axis' planes is inside the slab or nowhere
        inside = (e0 <= 0) & (e1 >= 0)
        k0 = torch.where(nz, k0, torch.where(inside, torch.full_like(k0, -math.inf),
                                             torch.full_like(k0, math.inf)))
        k1 = torch.where(nz, k1, torch.where(inside, torch.full_like(k1, math.inf),
                                             torch.full_like(k1, -math.inf)))
        lo_k = torch.maximum(lo_k, k0)
        hi_k = torch.minimum(hi_k, k1)
    kmax_count = torch.ceil((rays.a_end[sl] - a_start) / rays.step) + 1
    kmin = torch.clamp(torch.floor(lo_k) - 1, min=0)
    kmax = torch.minimum(torch.ceil(hi_k) + 1, kmax_count)
    empty = ~(kmax >= kmin)
    kmin = torch.where(empty, torch.zeros_like(kmin), kmin)
    kmax = torch.where(empty, torch.full_like(kmax, -1), kmax)
    return kmin.to(torch.int64), kmax.to(torch.int64)


def _taps(p, N):
    """Trilinear taps of points p [M, 3] (texel centres at integers, zero
    border): flat indices [M, 8] (clamped) and weights [M, 8] (0 if out)."""
    X, Y, Z = N
    i0 = torch.floor(p)
    f = p - i0
    i0 = i0.to(torch.int64)
    flats = []
    ws = []
    for dz in (0, 1):
        iz = i0[:, 2] + dz
        wz = f[:, 2] if dz else 1 - f[:, 2]
        vz = (iz >= 0) & (iz < Z)
        for dy in (0, 1):
            iy = i0[:, 1] + dy
            wy = f[:, 1] if dy else 1 - f[:, 1]
            vy = (iy >= 0) & (iy < Y)
            for dx in (0, 1):
                ix = i0[:, 0] + dx
                wx = f[:, 0] if dx else 1 - f[:, 0]
                vx = (ix >= 0) & (ix < X)
                ok = vx & vy & vz
                flat = (iz * Y + iy) * X + ix
                flats.append(torch.where(ok, flat, torch.zeros_like(flat)))
                ws.append(torch.where(ok, wx * wy * wz, torch.zeros_like(wx)))
    return torch.stack(flats, 1), torch.stack(ws, 1)


def _march(rays, N, visit, chunk=1 << 17):
    """Call visit(ray_index[M], flat[M, 8], w[M, 8]) for every sample."""
    R = rays.n
    for c0 in range(0, R, chunk):
        sl = slice(c0, min(R, c0 + chunk))
        kmin, kmax = _k_window(rays, sl, N)
        span = int((kmax - kmin + 1).clamp(min=0).max().item()) if kmin.numel() else 0
        src = rays.src[sl]
        d = rays.d[sl]
        a_start = rays.a_start[sl]
        a_end = rays.a_end[sl]
        ridx = torch.arange(c0, sl.stop)
        for j in range(span):
            k = kmin + j
            live = (k <= kmax)
            alpha = a_start + k.to(a_start.dtype) * rays.step
            live &= alpha < a_end
            if not bool(live.any()):
                continue
            al = alpha[live]
            p = (src[live] + al[:, None] * d[live]) / rays.div_s + rays.off
            flat, w = _taps(p, N)
            visit(ridx[live], flat, w)


def forward(vol, rays, N):
    """Ray sums of ``vol`` [B, 1, Z, Y, X] -> [B, R]."""
    B = vol.shape[0]
    out = torch.zeros(B, rays.n, dtype=vol.dtype)
    flat_vol = vol.reshape(B, -1)

    def visit(ri, flat, w):
        for b in range(B):
            out[b].index_add_(0, ri, (w * flat_vol[b][flat]).sum(1))

    _march(rays, N, visit)
    return out * rays.scale


def adjoint(sino_flat, rays, N):
    """Exact transpose of ``forward``: sino [B, R] -> volume [B, Z*Y*X]."""
    B = sino_flat.shape[0]
    X, Y, Z = N
    out = torch.zeros(B, X * Y * Z, dtype=sino_flat.dtype)

    def visit(ri, flat, w):
        for b in range(B):
            out[b].index_add_(0, flat.reshape(-1), (w * sino_flat[b][ri][:, None]).reshape(-1))

    _march(rays, N, visit)
    return out * rays.scale


def sample_count(rays, N):
    """Number of samples whose trilinear support touches the volume."""
    cnt = [0]

    def visit(ri, flat, w):
        cnt[0] += int(((w != 0).any(1)).sum().item())

    _march(rays, N, visit)
    return cnt[0]


def _shape_out(flat, B, A, V, U):
    return flat.reshape(B, 1, A, V, U)


# --------------------------------------------------------------------------
# Public front-ends: forward A and matched adjoint A^T
# --------------------------------------------------------------------------

def std_forward(vol, angles, volume_size, detector_size, sid, sdd, dtype=F64):
    X, Y, Z = volume_size
    U, V = detector_size
    rays = std_rays(angles, volume_size, detector_size, sid, sdd, dtype)
    out = forward(vol.to(dtype), rays, (X, Y, Z))
    return _shape_out(out, vol.shape[0], len(angles), V, U)


def std_adjoint(sino, angles, volume_size, detector_size, sid, sdd, dtype=F64):
    X, Y, Z = volume_size
    rays = std_rays(angles, volume_size, detector_size, sid, sdd, dtype)
    out = adjoint(sino.to(dtype).reshape(sino.shape[0], -1), rays, (X, Y, Z))
    return out.reshape(sino.shape[0], 1, Z, Y, X)


def bj_forward(vol, project_vector, volume_size, detector_size, dtype=F64):
    X, Y, Z = volume_size
    U, V = detector_size
    rays = bj_rays(project_vector, volume_size, detector_size, dtype)
    out = forward(vol.to(dtype), rays, (X, Y, Z))
    return _shape_out(out, vol.shape[0], len(project_vector), V, U)


def bj_adjoint(sino, project_vector, volume_size, detector_size, dtype=F64):
    X, Y, Z = volume_size
    rays = bj_rays(project_vector, volume_size, detector_size, dtype)
    out = adjoint(sino.to(dtype).reshape(sino.shape[0], -1), rays, (X, Y, Z))
    return out.reshape(sino.shape[0], 1, Z, Y, X)


def circ_forward(vol, angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd,
                 dtype=F64):
    X, Y, Z = volume_size
    U, V = detector_size
    rays = circ_rays(angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd, dtype)
    out = forward(vol.to(dtype), rays, (X, Y, Z))
    return _shape_out(out, vol.shape[0], len(angles), V, U)


def circ_adjoint(sino, angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd,
                 dtype=F64):
    X, Y, Z = volume_size
    rays = circ_rays(angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd, dtype)
    out = adjoint(sino.to(dtype).reshape(sino.shape[0], -1), rays, (X, Y, Z))
    return out.reshape(sino.shape[0], 1, Z, Y, X)


def circ_forward_at(vol, angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd,
                    rays_idx, dtype=F64):
    """``circ_forward`` evaluated at the rays ``rays_idx`` (flat indices in
    sinogram order [view, v, u]) only; returns [B, K].  The same ray
    construction and sample enumeration, restricted to the chosen rays."""
    X, Y, Z = volume_size
    rays = circ_rays(angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd, dtype)
    idx = torch.as_tensor(np.asarray(rays_idx), dtype=torch.int64).view(-1)
    sub = Rays(rays.src[idx], rays.d[idx], rays.a_start[idx], rays.a_end[idx], rays.step,
               rays.div_s, rays.off, rays.scale)
    return forward(vol.to(dtype), sub, (X, Y, Z))


def circ_adjoint_at(sino, angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd,
                    voxels, dtype=F64):
    """``circ_adjoint`` evaluated at the voxels [K, 3] of (ix, iy, iz) only;
    returns [B, K].

    Same enumeration as ``_march``: samples k >= 0 with alpha_k < a_end,
    alpha_k = a_start + k*step, at p = (src + alpha_k d)/div_s + off, each
    adding w = prod_ax max(0, 1 - |p_ax - v_ax|) (the trilinear weight that
    ``_taps`` gives voxel v) times the ray's value.  The rays that can reach
    a voxel are found per view from the perspective image of its support
    cube [v - 1, v + 1]^3 (a convex box in front of the source maps inside
    the rectangle spanned by its corners' images), widened by one pixel;
    along a candidate ray only the k whose alpha falls in the intersection of
    the three slabs |p_ax - v_ax| < 1 are visited.  So full-size workloads
    (256^3, 60 views) cost seconds for a few thousand voxels."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    vx = float(voxel_mm)
    if np.ndim(pitch_mm) == 0:
        du = dv = float(pitch_mm)
    else:
        du, dv = (float(p) for p in pitch_mm)
    rays = circ_rays(angles, volume_size, voxel_mm, detector_size, pitch_mm, sod, sdd, dtype)
    vox = torch.as_tensor(np.asarray(voxels), dtype=torch.int64).view(-1, 3)
    K = vox.shape[0]
    B = sino.shape[0]
    sflat = sino.to(dtype).reshape(B, -1)
    out = torch.zeros(B, K, dtype=dtype)
    vf = vox.to(dtype)
    corners = torch.stack([((vf + torch.tensor(c, dtype=dtype)) - rays.off) * rays.div_s
                           for c in ((a, b, e) for a in (-1, 1) for b in (-1, 1)
                                     for e in (-1, 1))], 1)          # [K, 8, 3] world mm
    a = torch.as_tensor(np.asarray(angles, dtype=np.float64), dtype=dtype)
    theta = a / 180 * PI_REF - PI_REF / 2
    for j in range(a.shape[0]):
        c, sn = torch.cos(theta[j]), torch.sin(theta[j])
        rel_x = corners[..., 0] + c * sod                             # q - src, src = -ex*sod
        rel_y = corners[..., 1] + sn * sod
        depth = rel_x * c + rel_y * sn
        if not bool((depth > 0).all()):
            raise ValueError("a voxel's support reaches behind the source")
        iu = sdd * (-rel_x * sn + rel_y * c) / depth / du + U / 2.0 - 0.5
        iv = sdd * corners[..., 2] / depth / dv + V / 2.0 - 0.5
        u0 = (torch.floor(iu.min(1).values).to(torch.int64) - 1).clamp(min=0)
        u1 = (torch.ceil(iu.max(1).values).to(torch.int64) + 1).clamp(max=U - 1)
        v0 = (torch.floor(iv.min(1).values).to(torch.int64) - 1).clamp(min=0)
        v1 = (torch.ceil(iv.max(1).values).to(torch.int64) + 1).clamp(max=V - 1)
        _adjoint_at_candidates(sflat, rays, vf, j, U, V, u0, u1, v0, v1, out)
    return out * rays.scale


def _adjoint_at_candidates(sflat, rays, vf, j, U, V, u0, u1, v0, v1, out):
    """Adds into out [B, K] the adjoint contributions of view j's candidate
    rays, pixels [u0, u1] x [v0, v1] per voxel (of vf [K, 3]): along each ray
    only the k whose alpha lies in the three slabs |p_ax - v_ax| < 1 are
    visited, each adding the trilinear weight times the ray's value (the
    enumeration of ``_march``)."""
    K = vf.shape[0]
    B = sflat.shape[0]
    dtype = vf.dtype
    wu = int((u1 - u0 + 1).clamp(min=0).max())
    wv = int((v1 - v0 + 1).clamp(min=0).max())
    if wu <= 0 or wv <= 0:
        return
    ou = torch.arange(wu).view(1, 1, wu)
    ov = torch.arange(wv).view(1, wv, 1)
    uu = u0.view(K, 1, 1) + ou
    vv = v0.view(K, 1, 1) + ov
    ok = (uu <= u1.view(K, 1, 1)) & (vv <= v1.view(K, 1, 1))
    kk = torch.arange(K).view(K, 1, 1).expand(K, wv, wu)[ok]
    r = (j * V + vv.expand(K, wv, wu)[ok]) * U + uu.expand(K, wv, wu)[ok]
    src, d = rays.src[r], rays.d[r]
    a_start, a_end = rays.a_start[r], rays.a_end[r]
    v = vf[kk]
    lo = torch.full_like(a_start, -math.inf)
    hi = torch.full_like(a_start, math.inf)
    for ax in range(3):
        dd = d[:, ax]
        nz = dd != 0
        safe = torch.where(nz, dd, torch.ones_like(dd))
        e0 = ((v[:, ax] - 1 - rays.off[ax]) * rays.div_s - src[:, ax]) / safe
        e1 = ((v[:, ax] + 1 - rays.off[ax]) * rays.div_s - src[:, ax]) / safe
        lo = torch.where(nz, torch.maximum(lo, torch.minimum(e0, e1)), lo)
        hi = torch.where(nz, torch.minimum(hi, torch.maximum(e0, e1)), hi)
    live = hi > lo
    if not bool(live.any()):
        return
    kk, r, src, d, a_start, a_end, v = (t[live] for t in (kk, r, src, d, a_start, a_end, v))
    k_lo = torch.floor((lo[live] - a_start) / rays.step) - 1
    k_hi = torch.ceil((hi[live] - a_start) / rays.step) + 1
    span = int((k_hi - k_lo + 1).max())
    if span > 64:
        raise ValueError("unexpectedly long sample run at one voxel")
    acc = torch.zeros(kk.shape[0], dtype=dtype)
    for t in range(span):
        k = k_lo + t
        alpha = a_start + k * rays.step
        p = (src + alpha[:, None] * d) / rays.div_s + rays.off
        w = torch.clamp(1 - (p - v).abs(), min=0).prod(1)
        keep = (k >= 0) & (k <= k_hi) & (alpha < a_end)
        acc += torch.where(keep, w, torch.zeros_like(w))
    for b in range(B):
        out[b].index_add_(0, kk, acc * sflat[b][r])


def table_adjoint_at(sino, views, volume_size, detector_size, axis_mode, box_lo, box_hi,
                     margin, scale, voxels, dtype=F64):
    """``table_adjoint`` (canonical view table [n, 12] = S, D0, U, V in the
    voxel frame; the Standard, circular and Beijing schemes all reduce to
    it) evaluated at the voxels [K, 3] of (ix, iy, iz) only; returns [B, K].
    Per view the rays that can reach a voxel are the pixels whose centres
    fall in the bounding rectangle, widened by one pixel, of the perspective
    images of its support cube's 8 corners (a convex box in front of the
    source maps inside the hull of its corners' images); then the same
    per-candidate enumeration as ``circ_adjoint_at``."""
    U, V = (int(v) for v in detector_size)
    rays = table_rays(views, detector_size, axis_mode, box_lo, box_hi, margin, scale, dtype)
    T = torch.as_tensor(np.asarray(views, dtype=np.float64)).to(dtype)
    vox = torch.as_tensor(np.asarray(voxels), dtype=torch.int64).view(-1, 3)
    K = vox.shape[0]
    B = sino.shape[0]
    sflat = sino.to(dtype).reshape(B, -1)
    out = torch.zeros(B, K, dtype=dtype)
    vf = vox.to(dtype)
    corners = torch.stack([vf + torch.tensor(c, dtype=dtype)
                           for c in ((a, b, e) for a in (-1, 1) for b in (-1, 1)
                                     for e in (-1, 1))], 1)          # [K, 8, 3] voxel frame
    for j in range(T.shape[0]):
        S, D0, Uv, Vv = T[j, 0:3], T[j, 3:6], T[j, 6:9], T[j, 9:12]
        n = torch.linalg.cross(Uv, Vv)
        h = torch.dot(D0 - S, n)          # the detector plane is (x - S).n = h
        w = corners - S                   # [K, 8, 3]
        wn = (w * n).sum(-1)
        if not bool((wn * h > 0).all()):
            raise ValueError("a voxel's support reaches behind the source")
        P = S + w * (h / wn).unsqueeze(-1) - D0   # on the detector plane, from D0
        # P = iu U + iv V: solve the 2x2 normal equations
        uu, uv_, vv_ = torch.dot(Uv, Uv), torch.dot(Uv, Vv), torch.dot(Vv, Vv)
        det = uu * vv_ - uv_ * uv_
        pu, pv = (P * Uv).sum(-1), (P * Vv).sum(-1)
        iu = (vv_ * pu - uv_ * pv) / det
        iv = (uu * pv - uv_ * pu) / det
        u0 = (torch.floor(iu.min(1).values).to(torch.int64) - 1).clamp(min=0)
        u1 = (torch.ceil(iu.max(1).values).to(torch.int64) + 1).clamp(max=U - 1)
        v0 = (torch.floor(iv.min(1).values).to(torch.int64) - 1).clamp(min=0)
        v1 = (torch.ceil(iv.max(1).values).to(torch.int64) + 1).clamp(max=V - 1)
        _adjoint_at_candidates(sflat, rays, vf, j, U, V, u0, u1, v0, v1, out)
    return out * rays.scale


def table_forward(vol, views, volume_size, detector_size, axis_mode, box_lo, box_hi,
                  margin, scale, dtype=F64):
    X, Y, Z = volume_size
    U, V = detector_size
    rays = table_rays(views, detector_size, axis_mode, box_lo, box_hi, margin, scale, dtype)
    out = forward(vol.to(dtype), rays, (X, Y, Z))
    return _shape_out(out, vol.shape[0], len(views), V, U)


def table_forward_at(vol, views, volume_size, detector_size, axis_mode, box_lo, box_hi,
                     margin, scale, rays_idx, dtype=F64):
    """``table_forward`` evaluated at the rays ``rays_idx`` (flat indices in
    sinogram order [view, v, u]) only; returns [B, K].  Rays are built for the
    views that hold chosen rays only (the same construction per view), then
    the same sample enumeration."""
    X, Y, Z = volume_size
    U, V = (int(v) for v in detector_size)
    idx = np.asarray(rays_idx, dtype=np.int64).reshape(-1)
    per = U * V
    vsel, pos = np.unique(idx // per, return_inverse=True)
    T = np.asarray(views, dtype=np.float64)[vsel]
    rays = table_rays(T, detector_size, axis_mode, box_lo, box_hi, margin, scale, dtype)
    sub_idx = torch.as_tensor(pos.reshape(-1) * per + idx % per, dtype=torch.int64)
    sub = Rays(rays.src[sub_idx], rays.d[sub_idx], rays.a_start[sub_idx],
               rays.a_end[sub_idx], rays.step, rays.div_s, rays.off, rays.scale)
    return forward(vol.to(dtype), sub, (X, Y, Z))

def table_adjoint(sino, views, volume_size, detector_size, axis_mode, box_lo, box_hi,
                  margin, scale, dtype=F64):
    X, Y, Z = volume_size
    rays = table_rays(views, detector_size, axis_mode, box_lo, box_hi, margin, scale, dtype)
    out = adjoint(sino.to(dtype).reshape(sino.shape[0], -1), rays, (X, Y, Z))
    return out.reshape(sino.shape[0], 1, Z, Y, X)


# --------------------------------------------------------------------------
# Reference-mode backprojectors (FDK-style weights; not the adjoint of A)
# --------------------------------------------------------------------------

def _bilinear(img, x, y):
    """Bilinear sample of img [V, U] at texel-centred coords (x along U)."""
    V, U = img.shape
    x0 = torch.floor(x)
    y0 = torch.floor(y)
    fx = x - x0
    fy = y - y0
    x0 = x0.to(torch.int64)
    y0 = y0.to(torch.int64)
    flat = img.reshape(-1)
    acc = torch.zeros_like(x)
    for dy in (0, 1):
        yy = y0 + dy
        wy = fy if dy else 1 - fy
        for dx in (0, 1):
            xx = x0 + dx
            wx = fx if dx else 1 - fx
            ok = (xx >= 0) & (xx < U) & (yy >= 0) & (yy < V)
            idx = torch.where(ok, yy * U + xx, torch.zeros_like(xx))
            acc = acc + torch.where(ok, wx * wy * flat[idx], torch.zeros_like(wx))
    return acc


def intersect_lines_2d(p1, p2, p3, p4):
    """Line (p1,p2) with line (p3,p4); NaN when nearly parallel
    (Standard/plug/include/helper_geometry.h:26-42)."""
    dnom = (p1[0] - p2[0]) * (p3[1] - p4[1]) - (p1[1] - p2[1]) * (p3[0] - p4[0])
    a = p1[0] * p2[1] - p1[1] * p2[0]
    b = p3[0] * p4[1] - p3[1] * p4[0]
    x = a * (p3[0] - p4[0]) - (p1[0] - p2[0]) * b
    y = a * (p3[1] - p4[1]) - (p1[1] - p2[1]) * b
    par = (dnom < 1e-6) & (dnom > -1e-6)
    nan = torch.full_like(dnom, float("nan"))
    return torch.where(par, nan, x / dnom), torch.where(par, nan, y / dnom)


def std_bp_ref(sino, angles, volume_size, detector_size, sid, sdd, dtype=F64, voxels=None):
    """Standard backwardKernel.cu:22-53 (voxel-driven, bilinear, weight
    (SDD/L)^2 * pi / A), summed over views.

    ``voxels`` (optional int array [K, 3] of (ix, iy, iz)): evaluate only
    those voxels and return [B, K] (the same arithmetic per voxel)."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    sino = sino.to(dtype)
    B = sino.shape[0]
    A = len(angles)
    s = sid / sdd                                                       # :29
    ang = torch.as_tensor(np.asarray(angles, dtype=np.float64), dtype=dtype)
    if voxels is None:
        ix = torch.arange(X, dtype=dtype).view(1, X)
        iy = torch.arange(Y, dtype=dtype).view(Y, 1)
        kk = torch.arange(Z, dtype=dtype).view(Z, 1, 1)
        shape = (Z, Y, X)
    else:
        vox = torch.as_tensor(np.asarray(voxels), dtype=dtype).view(-1, 3)
        ix, iy, kk = vox[:, 0].view(1, -1), vox[:, 1].view(1, -1), vox[:, 2].view(1, 1, -1)
        shape = (1, vox.shape[0])
    cx = ((-X / 2.0) + ix) * s                                          # :30
    cy = ((-Y / 2.0) + iy) * s
    cx, cy = torch.broadcast_tensors(cx, cy)
    out = torch.zeros(B, 1, *shape, dtype=dtype)
    for a in range(A):
        theta = ang[a] / 180 * PI_REF - PI_REF / 2                       # :32
        exx, exy = torch.cos(theta), torch.sin(theta)
        eyx, eyy = -exy, exx
        srcx, srcy = exx * sid, exy * sid                               # :35
        detx, dety = exx * (sdd - sid), exy * (sdd - sid)               # :36
        px, py = intersect_lines_2d((cx, cy), (-srcx + 0 * cx, -srcy + 0 * cy),
                                    (detx + 0 * cx, dety + 0 * cy),
                                    (detx + eyx + 0 * cx, dety + eyy + 0 * cy))   # :37
        x = (px * eyx + py * eyy) - (-U / 2.0)                          # :38
        bx, by = srcx + cx, srcy + cy                                   # :40
        dz = sdd / (bx * exx + by * exy) * s                            # :41
        z0 = (-Z / 2.0) * dz - (-V / 2.0)                               # :42
        coff = dz / s                                                   # :44
        zk = z0.unsqueeze(0) + kk * dz.unsqueeze(0)                     # :51 (exact)
        if voxels is None:
            xk = x.unsqueeze(0).expand(Z, Y, X)
            wk = (coff * coff * PI_REF / A).unsqueeze(0)
        else:
            zk, xk, wk = zk[0], x, coff * coff * PI_REF / A
        for b in range(B):
            val = _bilinear(sino[b, 0, a], xk.reshape(-1), zk.reshape(-1)).reshape(shape)
            out[b, 0] += val * wk                                       # :49-50
    return out if voxels is None else out[:, 0, 0]


def _det3(a, b, c):
    """Scalar triple product a . (b x c) on [..., 3] tensors."""
    return (a[..., 0] * (b[..., 1] * c[..., 2] - b[..., 2] * c[..., 1])
            - a[..., 1] * (b[..., 0] * c[..., 2] - b[..., 2] * c[..., 0])
            + a[..., 2] * (b[..., 0] * c[..., 1] - b[..., 1] * c[..., 0]))


def bj_bp_ref(sino, project_vector, volume_size, detector_size, dtype=F64, voxels=None):
    """Beijing backwardKernel.cu:18-45 (voxel-driven, triple-product
    projection, skip rule on *volume* dimensions, weight fr * 2pi/(n/21)).

    ``voxels`` (optional int array [K, 3] of (ix, iy, iz)): evaluate only
    those voxels and return [B, K] (the same arithmetic per voxel)."""
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    P = torch.as_tensor(np.asarray(project_vector, dtype=np.float32)).to(dtype)
    n = P.shape[0]
    angles_num = n // 21                                                # host :96
    sino = sino.to(dtype)
    B = sino.shape[0]
    if voxels is None:
        iz = torch.arange(Z, dtype=dtype).view(Z, 1, 1)
        iy = torch.arange(Y, dtype=dtype).view(1, Y, 1)
        ix = torch.arange(X, dtype=dtype).view(1, 1, X)
        c = torch.stack(torch.broadcast_tensors(-X / 2.0 + ix, -Y / 2.0 + iy, -Z / 2.0 + iz), -1)
        shape = (Z, Y, X)
    else:
        vox = torch.as_tensor(np.asarray(voxels), dtype=dtype).view(-1, 3)
        c = torch.stack([-X / 2.0 + vox[:, 0], -Y / 2.0 + vox[:, 1], -Z / 2.0 + vox[:, 2]], -1)
        shape = (vox.shape[0],)
    out = torch.zeros(B, 1, *shape, dtype=dtype)
    for j in range(n):
        S = torch.stack([P[j, 0], P[j, 1], -P[j, 2]])                  # :29
        D = torch.stack([P[j, 3], P[j, 4], -P[j, 5]])                  # :30
        v = P[j, 6:9]
        u = P[j, 9:12]
        cs = c - S
        sc = S - c
        sd = (S - D).expand_as(c)
        ub = u.expand_as(c)
        vb = v.expand_as(c)
        fscale = 1.0 / _det3(ub, vb, sc)                               # :34
        sx = fscale * _det3(cs, vb, sd) - (-U / 2.0)                   # :35
        sy = fscale * _det3(ub, cs, sd) - (-V / 2.0)                   # :36
        fr = fscale * _det3(ub, vb, sd)                                # :37
        skip = (sx < -1) | (sx > X + 1) | (sy < -1) | (sy > Y + 1)     # :38
        for b in range(B):
            val = _bilinear(sino[b, 0, j], (sx - 0.5).reshape(-1), (sy - 0.5).reshape(-1))
            out[b, 0] += torch.where(skip, torch.zeros_like(fr), fr * val.reshape(shape))
    out = out * (2 * PI_REF / angles_num)                               # :43
    return out if voxels is None else out[:, 0]
