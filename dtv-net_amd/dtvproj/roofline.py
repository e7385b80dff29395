"""Algorithmic byte accounting for the projector kernels.

A sample is *in support* when its 2x2x2 trilinear stencil touches the volume
(-1 < p < N on every axis).  Per in-support sample, A reads 8 FP32 taps
(32 B); the matched A^T scatters the same 8 weights (32 B).  A writes 4 B
per ray; A^T writes 4 B per voxel.  The reference backprojector reads 4 taps
(16 B) per voxel-view and writes 4 B per voxel.

The sample count is exact: for every ray, the number of integers k in
[0, kend) with -1 < P0 + k d < N on each axis is counted in closed form
(vectorised over rays, FP64), using the same per-ray set-up as the kernels.
"""
from __future__ import annotations

import numpy as np

from .geometry import AXIS_XY, ConeGeometry

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, GB/s
# MI355X vector issue: one wave64 VALU instruction per compute unit per cycle
# (four SIMD16 units, four cycles each), 256 CUs at 2.4 GHz
VALU_PEAK_WAVE_INSTS = 256 * 2.4e9


def _open_count(lo, hi):
    """Number of integers strictly inside (lo, hi)."""
    n = np.ceil(hi) - np.floor(lo) - 1.0
    return np.maximum(n, 0.0)


def sample_count(geom: ConeGeometry, chunk_views: int = 8) -> int:
    """Exact number of in-support samples of one forward projection."""
    X, Y, Z = geom.volume_size
    U, V = geom.detector_size
    N = np.array([X, Y, Z], dtype=np.float64)
    sc = geom.scheme
    lo_box = np.array(sc.box_lo, dtype=np.float64)
    hi_box = np.array(sc.box_hi, dtype=np.float64)
    iu = np.arange(U, dtype=np.float64).reshape(1, 1, U, 1)
    iv = np.arange(V, dtype=np.float64).reshape(1, V, 1, 1)
    total = 0
    for a0 in range(0, geom.n_views, chunk_views):
        T = geom.views[a0:a0 + chunk_views]
        n = T.shape[0]
        S = T[:, 0:3].reshape(n, 1, 1, 3)
        pix = T[:, 3:6].reshape(n, 1, 1, 3) + iu * T[:, 6:9].reshape(n, 1, 1, 3) \
            + iv * T[:, 9:12].reshape(n, 1, 1, 3)
        r = pix - S
        d = r / np.sqrt((r * r).sum(-1, keepdims=True))
        ad = np.abs(d)
        if sc.axis_mode == AXIS_XY:
            ax = np.where(ad[..., 0] >= ad[..., 1], 0, 1)
        else:
            m = np.maximum(ad[..., 0], np.maximum(ad[..., 2], ad[..., 1]))
            ax = np.where(ad[..., 0] == m, 0, np.where(ad[..., 1] == m, 1, 2))
        Sb = np.broadcast_to(S, d.shape)
        dax = np.take_along_axis(d, ax[..., None], -1)[..., 0]
        sax = np.take_along_axis(Sb, ax[..., None], -1)[..., 0]
        a0_ = (lo_box[ax] - sax) / dax
        a1_ = (hi_box[ax] - sax) / dax
        bs = np.minimum(a0_, a1_) - sc.margin
        be = np.maximum(a0_, a1_) + sc.margin
        kend = np.ceil(be - bs)
        P0 = Sb + bs[..., None] * d
        klo = np.full(bs.shape, -0.5)      # k >= 0  <=>  k > -0.5
        khi = kend - 0.5                   # k < kend
        for c in range(3):
            dc = d[..., c]
            with np.errstate(divide="ignore", invalid="ignore"):
                t0 = (-1.0 - P0[..., c]) / dc
                t1 = (N[c] - P0[..., c]) / dc
            inside = (P0[..., c] > -1.0) & (P0[..., c] < N[c])
            lo_c = np.where(dc != 0, np.minimum(t0, t1), np.where(inside, -np.inf, np.inf))
            hi_c = np.where(dc != 0, np.maximum(t0, t1), np.where(inside, np.inf, -np.inf))
            klo = np.maximum(klo, lo_c)
            khi = np.minimum(khi, hi_c)
        total += int(_open_count(klo, khi).sum())
    return total


def pair_bytes(geom: ConeGeometry, samples: int | None = None, batch: int = 1) -> dict:
    """Algorithmic bytes of A, matched A^T and the reference backprojector."""
    if samples is None:
        samples = sample_count(geom)
    X, Y, Z = geom.volume_size
    nvox = X * Y * Z
    rays = geom.n_rays
    fwd = batch * (32 * samples + 4 * rays)
    adj = batch * (32 * samples + 4 * nvox)
    bp = batch * (16 * nvox * geom.n_views + 4 * nvox)
    compulsory = batch * 4 * (2 * nvox + 2 * rays)
    return dict(samples=samples, forward=fwd, adjoint=adj, pair=fwd + adj, bp_ref=bp,
                compulsory_pair=compulsory, rays=rays * batch, voxels=nvox * batch)
