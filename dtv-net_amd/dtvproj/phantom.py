"""Seeded synthetic inputs: random-ellipsoid phantoms, Gaussian volumes and
sinograms, Poisson noise.

All generators run on the CPU with a ``torch.Generator`` so CPU and GPU
paths see bit-identical inputs; move the result with ``.to(device)``.
"""
from __future__ import annotations

import torch


def _rotation(gen: torch.Generator, n: int) -> torch.Tensor:
    """n uniformly random rotations (unit quaternions -> matrices)."""
    q = torch.randn(n, 4, generator=gen, dtype=torch.float64)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
    ], 1).view(n, 3, 3)
    return R


def ellipsoids(volume_size, voxel_mm: float, n: int, seed: int,
               semi_axes_mm=(5.0, 25.0), attenuation=(0.1, 1.0), inner=0.8):
    """Parameters of ``n`` random ellipsoids: centres uniform in the inner
    ``inner`` fraction of the field of view, semi-axes uniform in
    ``semi_axes_mm``, uniformly random orientation, additive attenuation
    uniform in ``attenuation`` (per cm; the phantom stores per-mm values
    ``attenuation / 10`` so line integrals in mm stay photon-count friendly).
    """
    gen = torch.Generator().manual_seed(int(seed))
    X, Y, Z = volume_size
    ext = torch.tensor([X, Y, Z], dtype=torch.float64) * voxel_mm
    centres = (torch.rand(n, 3, generator=gen, dtype=torch.float64) - 0.5) * ext * inner
    lo, hi = semi_axes_mm
    axes = lo + (hi - lo) * torch.rand(n, 3, generator=gen, dtype=torch.float64)
    R = _rotation(gen, n)
    alo, ahi = attenuation
    mu = (alo + (ahi - alo) * torch.rand(n, generator=gen, dtype=torch.float64)) / 10.0
    return dict(centres=centres, axes=axes, rot=R, mu=mu)


def render(volume_size, voxel_mm: float, params, dtype=torch.float32) -> torch.Tensor:
    """Voxelise ellipsoids at voxel centres ((i + 0.5 - N/2) * voxel_mm)
    into a [1, 1, Z, Y, X] volume."""
    X, Y, Z = volume_size
    out = torch.zeros(Z, Y, X, dtype=torch.float64)
    xs = (torch.arange(X, dtype=torch.float64) + 0.5 - X / 2.0) * voxel_mm
    ys = (torch.arange(Y, dtype=torch.float64) + 0.5 - Y / 2.0) * voxel_mm
    zs = (torch.arange(Z, dtype=torch.float64) + 0.5 - Z / 2.0) * voxel_mm
    n = params["mu"].shape[0]
    for e in range(n):
        c = params["centres"][e]
        a = params["axes"][e]
        R = params["rot"][e]
        # bounding box of the ellipsoid in voxel indices
        r = float(a.max())
        sl = []
        for coord, cc in ((zs, c[2]), (ys, c[1]), (xs, c[0])):
            idx = torch.nonzero((coord >= cc - r) & (coord <= cc + r)).flatten()
            if idx.numel() == 0:
                sl = None
                break
            sl.append(slice(int(idx[0]), int(idx[-1]) + 1))
        if sl is None:
            continue
        pz = zs[sl[0]].view(-1, 1, 1) - c[2]
        py = ys[sl[1]].view(1, -1, 1) - c[1]
        px = xs[sl[2]].view(1, 1, -1) - c[0]
        # body coordinates q = R^T (p - c)
        q0 = R[0, 0] * px + R[1, 0] * py + R[2, 0] * pz
        q1 = R[0, 1] * px + R[1, 1] * py + R[2, 1] * pz
        q2 = R[0, 2] * px + R[1, 2] * py + R[2, 2] * pz
        inside = (q0 / a[0]) ** 2 + (q1 / a[1]) ** 2 + (q2 / a[2]) ** 2 <= 1.0
        out[sl[0], sl[1], sl[2]] += inside.to(torch.float64) * params["mu"][e]
    return out.to(dtype).view(1, 1, Z, Y, X)


def phantom(volume_size, voxel_mm: float, n: int, seed: int, dtype=torch.float32):
    """Seeded ``n``-ellipsoid phantom [1, 1, Z, Y, X] (CPU)."""
    return render(volume_size, voxel_mm, ellipsoids(volume_size, voxel_mm, n, seed), dtype)


def gaussian(shape, seed: int, dtype=torch.float32) -> torch.Tensor:
    """i.i.d. N(0, 1) tensor from a seeded CPU generator."""
    gen = torch.Generator().manual_seed(int(seed))
    return torch.randn(*shape, generator=gen, dtype=torch.float64).to(dtype)


def poisson_noise(sino: torch.Tensor, i0: float, seed: int) -> torch.Tensor:
    """Noisy line integrals -log(N / I0), N ~ Poisson(I0 exp(-sino)),
    counts clamped at 1 (seeded, on the CPU)."""
    gen = torch.Generator().manual_seed(int(seed))
    lam = i0 * torch.exp(-sino.detach().double().cpu())
    counts = torch.poisson(lam, generator=gen).clamp(min=1.0)
    return (-torch.log(counts / i0)).to(sino.dtype).to(sino.device)
