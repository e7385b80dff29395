"""Cone-beam geometries reduced to one canonical form for the HIP kernels.

Every geometry the projector supports is expressed as a table of FP64 view
vectors in the *voxel frame* -- the frame in which voxel (ix, iy, iz) of a
C-contiguous ``[Z, Y, X]`` volume has its sample centre at the integer point
(ix, iy, iz) -- plus a small *sampling scheme* that fixes where along each ray
the unit-spaced samples fall (their phase).  Per view the table row holds

    S  (3)  source position
    D0 (3)  position of detector pixel (iu=0, iv=0)
    U  (3)  step from pixel iu to iu+1
    V  (3)  step from pixel iv to iv+1

so that the ray of pixel (iu, iv) runs from S through D0 + iu*U + iv*V.  Rays
are normalised and sampled at beta_k = beta_start + k (one voxel apart); the
scheme gives beta_start from the box planes of the ray's dominant axis.

Three constructors restate the reference geometries:

* ``standard``  -- circular orbit from ``Standard/plug/kernel/forwardKernel.cu:
  22-86`` (world unit = one detector pixel, voxel = SID/SDD world units, the
  alpha range taken from box planes at +-N/2 *world* units, margin 3 world
  units, output scaled by SID/SDD).
* ``beijing``   -- 12-float projection vectors from ``Beijing/plug/kernel/
  forwardKernel.cu:17-76`` (source and detector z negated, pixel centres at
  +0.5, three-way dominant axis, margin 3, unscaled output; the Python layer
  multiplies by ``sampleInterval``, ``BeijingGeometry.py:17``).
* ``circular``  -- the physical-unit circular geometry the benchmark configs
  use (voxel size, detector pitch, SOD, SDD in mm); voxel centres are
  symmetric about the rotation axis, pixel centres at +0.5, output in
  attenuation x mm.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

# The reference's own value of pi (``#define PI 3.14159265359``,
# Standard/plug/kernel/forwardKernel.cu:14, backwardKernel.cu:14,
# Beijing/plug/kernel/backwardKernel.cu:12).
PI_REF = 3.14159265359

AXIS_XY = 2   # dominant axis chosen among x, y; |dx| >= |dy| picks x
AXIS_XYZ = 3  # dominant axis chosen among x, y, z; ties resolved x -> y -> z

KIND_TABLE = 0
KIND_STANDARD = 1
KIND_BEIJING = 2


@dataclass(frozen=True)
class Scheme:
    """Where samples fall along a ray and how the sum is scaled.

    alpha0/1 = (box_lo/hi[ax] - S[ax]) / d[ax] on the dominant axis ``ax``;
    beta_start = min(alpha0, alpha1) - margin, samples while
    beta_start + k < max(alpha0, alpha1) + margin, sum multiplied by ``scale``.
    """
    axis_mode: int
    box_lo: tuple
    box_hi: tuple
    margin: float
    scale: float


@dataclass
class ConeGeometry:
    """A cone-beam geometry in canonical voxel-frame form.

    ``views`` is float64 ``[n_views, 12]`` (S, D0, U, V).  ``kind`` records
    which reference geometry it restates; ``ref`` keeps the reference-form
    parameters (angles/SID/SDD or the float32 projection vectors) needed by
    the reference-mode backprojectors.
    """
    volume_size: tuple          # (X, Y, Z)
    detector_size: tuple        # (U, V)
    views: np.ndarray           # float64 [n_views, 12]
    scheme: Scheme
    kind: int = KIND_TABLE
    ref: dict = field(default_factory=dict)
    _dev_cache: dict = field(default_factory=dict, repr=False, compare=False)

    @property
    def n_views(self) -> int:
        return int(self.views.shape[0])

    @property
    def n_rays(self) -> int:
        return self.n_views * self.detector_size[0] * self.detector_size[1]

    def sino_shape(self, batch: int = 1):
        U, V = self.detector_size
        return (batch, 1, self.n_views, V, U)

    def volume_shape(self, batch: int = 1):
        X, Y, Z = self.volume_size
        return (batch, 1, Z, Y, X)

    def subset(self, idx: Sequence[int]) -> "ConeGeometry":
        """Geometry restricted to the given views (views are independent)."""
        idx = np.asarray(idx, dtype=np.int64)
        ref = dict(self.ref)
        if "angles" in ref:
            ref["angles"] = np.asarray(ref["angles"], dtype=np.float64)[idx]
        if "vectors" in ref:
            ref["vectors"] = np.asarray(ref["vectors"], dtype=np.float32)[idx]
        return ConeGeometry(self.volume_size, self.detector_size,
                            np.ascontiguousarray(self.views[idx]), self.scheme,
                            self.kind, ref)

    def device_views(self, device):
        """FP64 view table on ``device``, uploaded once and cached."""
        import torch
        key = str(device)
        t = self._dev_cache.get(key)
        if t is None:
            t = torch.from_numpy(np.ascontiguousarray(self.views)).to(device)
            self._dev_cache[key] = t
        return t


def _rotation(angles_deg):
    a = np.asarray(angles_deg, dtype=np.float64)
    theta = a / 180.0 * PI_REF - PI_REF / 2.0
    return np.cos(theta), np.sin(theta)


def standard(angles_deg, volume_size, detector_size, sid, sdd) -> ConeGeometry:
    """Circular geometry of ``JITStandardGeometry`` (forwardKernel.cu:22-86).

    World unit is one detector pixel; the source sits at -ex*SID, the
    detector line at +ex*(SDD-SID); pixel iu lies at ey*(iu - U/2) (no +0.5);
    voxels are SID/SDD world units wide and voxel coordinate
    p = world/(SID/SDD) + N/2.
    """
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    sid = float(sid)
    sdd = float(sdd)
    s = sid / sdd
    c, sn = _rotation(angles_deg)
    n = c.shape[0]
    views = np.zeros((n, 12), dtype=np.float64)
    Sx = -c * sid / s + X / 2.0
    Sy = -sn * sid / s + Y / 2.0
    Sz = np.full(n, Z / 2.0)
    views[:, 0], views[:, 1], views[:, 2] = Sx, Sy, Sz
    # pixel (0,0) relative to the source: ex*SDD + ey*(-U/2) + ez*(-V/2)
    views[:, 3] = Sx + (c * sdd + sn * (U / 2.0)) / s
    views[:, 4] = Sy + (sn * sdd - c * (U / 2.0)) / s
    views[:, 5] = Sz + (-V / 2.0) / s
    views[:, 6] = -sn / s
    views[:, 7] = c / s
    views[:, 8] = 0.0
    views[:, 9] = 0.0
    views[:, 10] = 0.0
    views[:, 11] = 1.0 / s
    half = (X / 2.0 / s, Y / 2.0 / s, Z / 2.0 / s)
    scheme = Scheme(AXIS_XY,
                    (X / 2.0 - half[0], Y / 2.0 - half[1], Z / 2.0 - half[2]),
                    (X / 2.0 + half[0], Y / 2.0 + half[1], Z / 2.0 + half[2]),
                    3.0 / s, s)
    ref = dict(angles=np.asarray(angles_deg, dtype=np.float64), sid=sid, sdd=sdd)
    return ConeGeometry((X, Y, Z), (U, V), views, scheme, KIND_STANDARD, ref)


def beijing(project_vector, volume_size, detector_size) -> ConeGeometry:
    """Vector geometry of ``JITBeijingGeometry`` (forwardKernel.cu:17-76).

    Row j of ``project_vector`` (float32 [n, 12]) holds source (0..2),
    detector centre (3..5), v (6..8), u (9..11); the kernel negates the z of
    source and detector and places pixel centres at +0.5.
    """
    P = np.asarray(project_vector, dtype=np.float32)
    if P.ndim != 2 or P.shape[1] != 12:
        raise RuntimeError("project vector's shape must be [angle's number, 12]")
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    Pd = P.astype(np.float64)
    n = Pd.shape[0]
    views = np.zeros((n, 12), dtype=np.float64)
    S = np.stack([Pd[:, 0] + X / 2.0, Pd[:, 1] + Y / 2.0, -Pd[:, 2] + Z / 2.0], 1)
    Dc = np.stack([Pd[:, 3] + X / 2.0, Pd[:, 4] + Y / 2.0, -Pd[:, 5] + Z / 2.0], 1)
    v = Pd[:, 6:9]
    u = Pd[:, 9:12]
    D0 = Dc + (0.5 - U / 2.0) * u + (0.5 - V / 2.0) * v
    views[:, 0:3] = S
    views[:, 3:6] = D0
    views[:, 6:9] = u
    views[:, 9:12] = v
    scheme = Scheme(AXIS_XYZ, (0.0, 0.0, 0.0), (float(X), float(Y), float(Z)), 3.0, 1.0)
    return ConeGeometry((X, Y, Z), (U, V), views, scheme, KIND_BEIJING,
                        dict(vectors=P.copy()))


def circular(angles_deg, volume_size, voxel_mm, detector_size, pitch_mm,
             sod, sdd) -> ConeGeometry:
    """Physical-unit circular cone-beam geometry (benchmark configurations).

    The orbit orientation follows ``standard`` (theta = angle - 90 deg).
    Voxels are ``voxel_mm`` cubes centred on the rotation axis; pixel
    (iu, iv) centre sits at ((iu + 0.5 - U/2)*du, (iv + 0.5 - V/2)*dv) on a
    flat detector at distance SDD from the source.  Samples are one voxel
    apart, phased from the true box like the vector scheme, and the sum is
    multiplied by ``voxel_mm`` so the output is a line integral in
    attenuation x mm.
    """
    X, Y, Z = (int(v) for v in volume_size)
    U, V = (int(v) for v in detector_size)
    vx = float(voxel_mm)
    if np.ndim(pitch_mm) == 0:
        du = dv = float(pitch_mm)
    else:
        du, dv = (float(p) for p in pitch_mm)
    sod = float(sod)
    sdd = float(sdd)
    c, sn = _rotation(angles_deg)
    n = c.shape[0]
    off = np.array([X / 2.0 - 0.5, Y / 2.0 - 0.5, Z / 2.0 - 0.5])
    ex = np.stack([c, sn, np.zeros(n)], 1)
    ey = np.stack([-sn, c, np.zeros(n)], 1)
    ez = np.tile(np.array([0.0, 0.0, 1.0]), (n, 1))
    S = -ex * (sod / vx) + off
    Dc = ex * ((sdd - sod) / vx) + off
    Uv = ey * (du / vx)
    Vv = ez * (dv / vx)
    D0 = Dc + (0.5 - U / 2.0) * Uv + (0.5 - V / 2.0) * Vv
    views = np.concatenate([S, D0, Uv, Vv], 1)
    scheme = Scheme(AXIS_XYZ, (-0.5, -0.5, -0.5), (X - 0.5, Y - 0.5, Z - 0.5), 3.0, vx)
    ref = dict(angles=np.asarray(angles_deg, dtype=np.float64), voxel_mm=vx,
               pitch_mm=(du, dv), sod=sod, sdd=sdd)
    return ConeGeometry((X, Y, Z), (U, V), np.ascontiguousarray(views), scheme,
                        KIND_TABLE, ref)


def uniform_angles(n_views: int, full_deg: float = 360.0):
    """n angles over ``full_deg`` with the endpoint excluded
    (``np.linspace(0, 360, n, False)``, StandardGeometry.py:6)."""
    return np.linspace(0.0, full_deg, int(n_views), endpoint=False)


def beijing_default_vectors(n_angles=144, planes=21, sub_detector=(80, 128),
                            sid=708.0, sdd=1143.0, pitch=(1.0, 1.0), plane_gap=0.0):
    """Synthetic stand-in for the reference's ``projection_matrix`` .mat file.

    The reference loads per-view vectors from an absolute path that does not
    ship with it (options.yaml:42).  This builds the same layout -- for each
    of ``n_angles`` source angles, ``planes`` flat sub-detectors tiled side by
    side along u -- in the kernel's voxel units (before its z negation), so
    the Beijing path can run without the file.  ``pitch`` = (u, v) pixel
    pitch in voxels.
    """
    U, V = sub_detector
    pu, pv = pitch
    rows = []
    ang = uniform_angles(n_angles)
    c, sn = _rotation(ang)
    for a in range(n_angles):
        ex = np.array([c[a], sn[a], 0.0])
        ey = np.array([-sn[a], c[a], 0.0])
        src = -ex * sid
        for p in range(planes):
            shift = (p - (planes - 1) / 2.0) * (U + plane_gap) * pu
            det = ex * (sdd - sid) + ey * shift
            u = ey * pu
            v = np.array([0.0, 0.0, pv])
            rows.append([src[0], src[1], -src[2], det[0], det[1], -det[2],
                         v[0], v[1], v[2], u[0], u[1], u[2]])
    return np.asarray(rows, dtype=np.float32)


def angle_groups(n_blocks: int, block: int, take: int, step: int = 1, start: int = 0):
    """Source-angle indices of the reference's sparse-view selections: from
    each of ``n_blocks`` consecutive blocks of ``block`` angles, ``take``
    angles ``step`` apart starting ``start`` into the block.  The reference
    scripts use 24 blocks of 45 of the scan's 1080 angles: the first 3 of
    each (iterations/eval_generateSparseView.py:46-48, 72 angles) or 6 taken
    8 apart (iterations/fista_eval.py:14-16, ``[j*45:(j+1)*45:45//6+1]``,
    144 angles)."""
    if take < 1 or step < 1 or start < 0 or start + (take - 1) * step >= block:
        raise ValueError("the selection must fit inside one block")
    b = np.arange(n_blocks, dtype=np.int64)[:, None] * block
    k = start + np.arange(take, dtype=np.int64)[None, :] * step
    return (b + k).reshape(-1)


def subsample_beijing(projection, vectors, angles, planes: int = 21):
    """Keep the given source angles of a Beijing scan.

    ``projection``: [..., n_angles * planes, V, U] in the projector's view
    order (view = angle * planes + plane, as the raw files store
    [angles, planes, V, U]); ``vectors``: [n_angles * planes, 12].  Returns
    both restricted to ``angles`` in the same layout, which is what the
    reference scripts build by reshaping to [angles, planes, ...] and
    slicing (iterations/fista_eval.py:12-20)."""
    angles = np.asarray(angles, dtype=np.int64)
    views = (angles[:, None] * planes + np.arange(planes, dtype=np.int64)[None, :]).reshape(-1)
    n_views = vectors.shape[0]
    if n_views % planes or (views.size and (views.min() < 0 or views.max() >= n_views)):
        raise ValueError("angle index out of range or view count not a multiple of planes")
    if projection.shape[-3] != n_views:
        raise ValueError(f"projection has {projection.shape[-3]} views, vectors {n_views}")
    import torch
    if isinstance(projection, torch.Tensor):
        idx = torch.as_tensor(views, device=projection.device)
        proj = projection.index_select(projection.dim() - 3, idx).contiguous()
    else:
        proj = np.ascontiguousarray(np.take(np.asarray(projection), views, axis=-3))
    if isinstance(vectors, torch.Tensor):
        vec = vectors.index_select(0, torch.as_tensor(views, device=vectors.device)).contiguous()
    else:
        vec = np.ascontiguousarray(np.asarray(vectors)[views])
    return proj, vec
