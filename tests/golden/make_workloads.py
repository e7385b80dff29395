"""Writes tests/golden/oracle_workloads.npz: FP64-oracle outputs at two
full-size workloads, sampled at seeded voxels so the file stays small.

* ``s256_*`` -- the matched adjoint A^T over all 60 views of the sparse-view
  workload (256^3 at 0.5 mm, 60 views, 384x384 detector at 0.8 mm, SOD 600,
  SDD 1100) applied to the i.i.d. N(0,1) sinogram of seed 21
  (``dtvproj.phantom.gaussian``, float32), evaluated by
  ``oracle.cone.circ_adjoint_at`` at 4096 seeded voxels (3584 uniform, 512
  on the faces of the volume).  The sinogram itself (35 MB) is not stored:
  its FP64 sum and sum of squares are, so a test can prove it regenerated
  the same input.
* ``s256f_*`` -- A over all 60 views of the same workload applied to the
  i.i.d. N(0,1) volume of seed 20, at 4096 seeded rays
  (``oracle.cone.circ_forward_at``), with the volume's FP64 sum.
* ``u16_*`` -- 20 FISTA-TV iterations (smoothed TV, Neumann differences,
  lam = eps = 1e-3, non-negativity, FISTA momentum of FISTA.py:23-25) on the
  smallest-work instance of the 16-instance uneven workload (103x103x75 at
  1 mm, 38 views, 153x91 detector, SOD 500 / SDD 1000), run entirely in FP64
  on the oracle's circ_forward / circ_adjoint and oracle.tv.tv_eps_grad,
  from the sinogram of its seeded 10-ellipsoid phantom (seed 102).  The step
  is 1 / (L + 12 lam / eps) with L from 12 FP64 power iterations of A^T A.
  Stored: the step, the iterate at 32768 seeded voxels, its peak, sum and
  norm.

* ``l512_*`` -- the 512x512x384 workload at 0.4 mm, 120 views, 768x576
  detector at 0.6 mm, SOD 700 / SDD 1200: A of the i.i.d. N(0,1) volume of
  seed 20 at 4096 seeded rays (``oracle.cone.circ_forward_at``) and the
  matched A^T of the N(0,1) sinogram of seed 21 at 4096 seeded voxels (3584
  uniform, 512 on the faces), with the FP64 sums and sums of squares of both
  inputs so a test can prove it regenerated them.

* ``u128_*`` -- the 128^3 / 90-view unroll geometry (1 mm voxels, 192x192
  detector at 1 mm, SOD 500 / SDD 1000): A of the N(0,1) volume of seed 30
  at 2048 seeded rays and A^T of the N(0,1) sinogram of seed 31 at 2048
  seeded voxels (1792 uniform, 256 on the faces).

* ``bj_*`` -- the 3024-view Beijing scan of the benchmark's loop (256x256x64,
  144 angles x 21 planes of 80x128 sub-detectors, SID 708 / SDD 1143, pitch
  from ``dtvproj.geometry.beijing_default_vectors``): the matched A^T of the
  N(0,1) sinogram of seed 41 at 1024 seeded voxels (896 uniform, 128 on the
  faces), evaluated by ``oracle.cone.table_adjoint_at`` over all 3024 views,
  with the sinogram's FP64 sum and sum of squares.

* ``std_*`` -- the reference's Standard options geometry (256x256x64, 48
  views over 360 degrees, 288x28 detector, SID 1000 / SDD 2000 detector
  pixels): the matched A^T of the N(0,1) sinogram of seed 42 over all views at
  1024 seeded voxels (896 uniform, 128 on the faces), by
  ``oracle.cone.table_adjoint_at`` on the scheme's view table, with the
  sinogram's FP64 sum and sum of squares.

* ``bjf_*`` / ``stdf_*`` -- A over all views of the same Beijing scan and
  Standard options geometry, applied to the N(0,1) volume of seed 40, at
  4096 seeded rays (``oracle.cone.table_forward_at``), with the volume's FP64
  sum and sum of squares.

Takes several minutes on 8 cores.  Run from the repository root:
    python tests/golden/make_workloads.py          # everything
    python tests/golden/make_workloads.py l512 u128 s256f bj std fwd   # (re)write only these arrays
"""
import math
import os
import sys

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "dtv-net_amd"))

from oracle import cone, tv  # noqa: E402
from dtvproj import phantom  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "oracle_workloads.npz")

S256 = dict(vol=(256, 256, 256), voxel=0.5, det=(384, 384), pitch=0.8, n_views=60,
            sod=600.0, sdd=1100.0, sino_seed=21, n_uniform=3584, n_face=512, vox_seed=2021)
U16 = dict(vol=(103, 103, 75), voxel=1.0, det=(153, 91), pitch=2.0, n_views=38, sod=500.0,
           sdd=1000.0, phantom_seed=102, n_ellipsoids=10, lam=1e-3, eps=1e-3, iters=20,
           power_iters=12, n_vox=32768, vox_seed=2022)


def angles(n):
    return np.linspace(0.0, 360.0, n, endpoint=False)


L512 = dict(vol=(512, 512, 384), voxel=0.4, det=(768, 576), pitch=0.6, n_views=120,
            sod=700.0, sdd=1200.0, vol_seed=20, sino_seed=21, n_rays=4096, n_uniform=3584,
            n_face=512, ray_seed=2023, vox_seed=2024)


BJ = dict(vol=(256, 256, 64), det=(80, 128), angles=144, planes=21, sid=708.0, sdd=1143.0,
          sino_seed=41, n_uniform=896, n_face=128, vox_seed=2041)


def bj_geometry():
    """The Beijing scan of the benchmark's loop, built as the bench builds it."""
    from dtvproj import geometry as G
    c = BJ
    X, Y, Z = c["vol"]
    U, V = c["det"]
    mag = c["sdd"] / c["sid"]
    pitch = (1.2 * np.hypot(X, Y) * mag / (U * c["planes"]), 1.2 * Z * mag / V)
    vec = G.beijing_default_vectors(c["angles"], c["planes"], (U, V), c["sid"], c["sdd"], pitch)
    return G.beijing(vec, (X, Y, Z), (U, V))


def bj_voxels():
    c = BJ
    return face_and_uniform_voxels(c["vol"], c["n_uniform"], c["n_face"], c["vox_seed"])


def build_bj(log=print):
    c = BJ
    g = bj_geometry()
    y = phantom.gaussian(g.sino_shape(1), c["sino_seed"]).double()
    vox = bj_voxels()
    sc = g.scheme
    adj = cone.table_adjoint_at(y, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                sc.box_lo, sc.box_hi, sc.margin, sc.scale, vox)
    log(f"bj: {g.n_views} views, {vox.shape[0]} voxels")
    return dict(bj_voxels=vox, bj_adjoint=adj[0].numpy(), bj_sino_sum=np.float64(y.sum()),
                bj_sino_sumsq=np.float64((y * y).sum()))


STD = dict(vol=(256, 256, 64), det=(288, 28), n_views=48, sid=1000.0, sdd=2000.0,
           sino_seed=42, n_uniform=896, n_face=128, vox_seed=2042)


def std_geometry():
    """The reference's Standard options geometry (options.yaml defaults)."""
    from dtvproj import geometry as G
    c = STD
    return G.standard(G.uniform_angles(c["n_views"]), c["vol"], c["det"], c["sid"], c["sdd"])


def std_voxels():
    c = STD
    return face_and_uniform_voxels(c["vol"], c["n_uniform"], c["n_face"], c["vox_seed"])


def build_std(log=print):
    c = STD
    g = std_geometry()
    y = phantom.gaussian(g.sino_shape(1), c["sino_seed"]).double()
    vox = std_voxels()
    sc = g.scheme
    adj = cone.table_adjoint_at(y, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                sc.box_lo, sc.box_hi, sc.margin, sc.scale, vox)
    log(f"std: {g.n_views} views, {vox.shape[0]} voxels")
    return dict(std_voxels=vox, std_adjoint=adj[0].numpy(), std_sino_sum=np.float64(y.sum()),
                std_sino_sumsq=np.float64((y * y).sum()))


def table_rays_pick(g, n, seed):
    U, V = g.detector_size
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(g.n_views * U * V, size=n, replace=False)).astype(np.int64)


def build_table_forwards(log=print):
    """A of the seed-40 N(0,1) volume at 4096 seeded rays over all views of
    the Beijing scan and of the Standard options geometry."""
    out = {}
    for key, g, seed in (("bjf", bj_geometry(), 2043), ("stdf", std_geometry(), 2044)):
        X, Y, Z = g.volume_size
        x = phantom.gaussian((1, 1, Z, Y, X), 40).double()
        rays = table_rays_pick(g, 4096, seed)
        sc = g.scheme
        fwd = cone.table_forward_at(x, g.views, g.volume_size, g.detector_size, sc.axis_mode,
                                    sc.box_lo, sc.box_hi, sc.margin, sc.scale, rays)
        log(f"{key}: {g.n_views} views, {rays.shape[0]} rays")
        out.update({f"{key}_rays": rays, f"{key}_forward": fwd[0].numpy(),
                    f"{key}_vol_sum": np.float64(x.sum()),
                    f"{key}_vol_sumsq": np.float64((x * x).sum())})
    return out


def face_and_uniform_voxels(vol, n_uniform, n_face, seed):
    X, Y, Z = vol
    rng = np.random.default_rng(seed)
    N = np.array([X, Y, Z])
    uni = rng.integers(0, N, size=(n_uniform, 3))
    face = rng.integers(0, N, size=(n_face, 3))
    ax = rng.integers(0, 3, size=n_face)
    hi = rng.integers(0, 2, size=n_face)
    face[np.arange(n_face), ax] = np.where(hi == 1, N[ax] - 1, 0)
    return np.concatenate([uni, face]).astype(np.int32)


def l512_rays():
    c = L512
    U, V = c["det"]
    rng = np.random.default_rng(c["ray_seed"])
    return np.sort(rng.choice(c["n_views"] * U * V, size=c["n_rays"], replace=False)).astype(np.int64)


def l512_voxels():
    c = L512
    return face_and_uniform_voxels(c["vol"], c["n_uniform"], c["n_face"], c["vox_seed"])


def build_l512(log=print):
    c = L512
    X, Y, Z = c["vol"]
    U, V = c["det"]
    args = (angles(c["n_views"]), c["vol"], c["voxel"], c["det"], c["pitch"], c["sod"], c["sdd"])
    x = phantom.gaussian((1, 1, Z, Y, X), c["vol_seed"]).double()
    rays = l512_rays()
    fwd = cone.circ_forward_at(x, *args, rays)
    log(f"l512 forward at {rays.shape[0]} rays")
    out = dict(l512_rays=rays, l512_forward=fwd[0].numpy(),
               l512_vol_sum=np.float64(x.sum()), l512_vol_sumsq=np.float64((x * x).sum()))
    del x
    y = phantom.gaussian((1, 1, c["n_views"], V, U), c["sino_seed"]).double()
    vox = l512_voxels()
    adj = cone.circ_adjoint_at(y, *args, vox)
    log(f"l512 adjoint at {vox.shape[0]} voxels")
    out.update(l512_voxels=vox, l512_adjoint=adj[0].numpy(),
               l512_sino_sum=np.float64(y.sum()), l512_sino_sumsq=np.float64((y * y).sum()))
    return out


U128 = dict(vol=(128, 128, 128), voxel=1.0, det=(192, 192), pitch=1.0, n_views=90, sod=500.0,
            sdd=1000.0, vol_seed=30, sino_seed=31, n_rays=2048, n_uniform=1792, n_face=256,
            ray_seed=2025, vox_seed=2026)


def u128_rays():
    c = U128
    U, V = c["det"]
    rng = np.random.default_rng(c["ray_seed"])
    return np.sort(rng.choice(c["n_views"] * U * V, size=c["n_rays"], replace=False)).astype(np.int64)


def u128_voxels():
    c = U128
    return face_and_uniform_voxels(c["vol"], c["n_uniform"], c["n_face"], c["vox_seed"])


def s256_rays():
    c = S256
    U, V = c["det"]
    rng = np.random.default_rng(2027)
    return np.sort(rng.choice(c["n_views"] * U * V, size=4096, replace=False)).astype(np.int64)


def build_s256f(log=print):
    c = S256
    X, Y, Z = c["vol"]
    args = (angles(c["n_views"]), c["vol"], c["voxel"], c["det"], c["pitch"], c["sod"], c["sdd"])
    x = phantom.gaussian((1, 1, Z, Y, X), 20).double()
    rays = s256_rays()
    fwd = cone.circ_forward_at(x, *args, rays)
    log(f"s256f: {rays.shape[0]} rays")
    return dict(s256f_rays=rays, s256f_forward=fwd[0].numpy(), s256f_vol_sum=np.float64(x.sum()))


def build_u128(log=print):
    c = U128
    X, Y, Z = c["vol"]
    U, V = c["det"]
    args = (angles(c["n_views"]), c["vol"], c["voxel"], c["det"], c["pitch"], c["sod"], c["sdd"])
    x = phantom.gaussian((1, 1, Z, Y, X), c["vol_seed"]).double()
    rays = u128_rays()
    fwd = cone.circ_forward_at(x, *args, rays)
    y = phantom.gaussian((1, 1, c["n_views"], V, U), c["sino_seed"]).double()
    vox = u128_voxels()
    adj = cone.circ_adjoint_at(y, *args, vox)
    log(f"u128: {rays.shape[0]} rays, {vox.shape[0]} voxels")
    return dict(u128_rays=rays, u128_forward=fwd[0].numpy(), u128_vol_sum=np.float64(x.sum()),
                u128_voxels=vox, u128_adjoint=adj[0].numpy(), u128_sino_sum=np.float64(y.sum()))


def s256_voxels():
    X, Y, Z = S256["vol"]
    rng = np.random.default_rng(S256["vox_seed"])
    N = np.array([X, Y, Z])
    uni = rng.integers(0, N, size=(S256["n_uniform"], 3))
    face = rng.integers(0, N, size=(S256["n_face"], 3))
    ax = rng.integers(0, 3, size=S256["n_face"])
    hi = rng.integers(0, 2, size=S256["n_face"])
    face[np.arange(S256["n_face"]), ax] = np.where(hi == 1, N[ax] - 1, 0)
    return np.concatenate([uni, face]).astype(np.int32)


def u16_voxels():
    X, Y, Z = U16["vol"]
    rng = np.random.default_rng(U16["vox_seed"])
    return np.sort(rng.choice(X * Y * Z, size=U16["n_vox"], replace=False)).astype(np.int64)


def build_s256():
    c = S256
    X, Y, Z = c["vol"]
    U, V = c["det"]
    y = phantom.gaussian((1, 1, c["n_views"], V, U), c["sino_seed"]).double()
    vox = s256_voxels()
    out = cone.circ_adjoint_at(y, angles(c["n_views"]), c["vol"], c["voxel"], c["det"],
                               c["pitch"], c["sod"], c["sdd"], vox)
    return dict(s256_voxels=vox, s256_adjoint=out[0].numpy(),
                s256_sino_sum=np.float64(y.sum()), s256_sino_sumsq=np.float64((y * y).sum()))


def build_u16(log=print):
    c = U16
    args = (angles(c["n_views"]), c["vol"], c["voxel"], c["det"], c["pitch"], c["sod"], c["sdd"])
    A = lambda x: cone.circ_forward(x, *args)          # noqa: E731
    AT = lambda s: cone.circ_adjoint(s, *args)         # noqa: E731
    ph = phantom.phantom(c["vol"], c["voxel"], c["n_ellipsoids"], seed=c["phantom_seed"])
    b = A(ph.double())
    # power iteration for L = ||A^T A|| from a seeded start
    v = phantom.gaussian(ph.shape, 7, dtype=torch.float64)
    v = v / v.norm()
    lip = 0.0
    for k in range(c["power_iters"]):
        w = AT(A(v))
        lip = float(w.norm())
        v = w / lip
        log(f"power iteration {k}: {lip:.6e}")
    lam, eps = c["lam"], c["eps"]
    step = 1.0 / (lip + lam * 12.0 / eps)
    x = torch.zeros(ph.shape, dtype=torch.float64)
    y = x.clone()
    t = 1.0
    for k in range(c["iters"]):
        tn = (1.0 + math.sqrt(1.0 + 4.0 * t * t)) / 2.0       # FISTA.py:23-25
        beta = (t - 1.0) / tn
        t = tn
        g = AT(A(y) - b).numpy()
        gtv = tv.tv_eps_grad(y.numpy()[0, 0], eps, c["vol"], "neumann")[None, None]
        xn = np.maximum(y.numpy() - step * (g + lam * gtv), 0.0)
        xn = torch.from_numpy(xn)
        y = xn + beta * (xn - x)
        x = xn
        log(f"fista iteration {k}: |x| = {float(x.norm()):.6e}")
    flat = x.reshape(-1)
    idx = u16_voxels()
    return dict(u16_step=np.float64(step), u16_lip=np.float64(lip), u16_voxels=idx,
                u16_values=flat[torch.from_numpy(idx)].numpy(),
                u16_peak=np.float64(flat.abs().max()), u16_sum=np.float64(flat.sum()),
                u16_norm=np.float64(flat.norm()))


if __name__ == "__main__":
    torch.set_num_threads(os.cpu_count() or 1)
    log = lambda s: print(s, flush=True)  # noqa: E731
    parts = dict(l512=build_l512, u128=build_u128, s256f=build_s256f, bj=build_bj, std=build_std,
                 fwd=build_table_forwards)
    if sys.argv[1:]:
        with np.load(OUT) as f:
            fx = {k: f[k] for k in f.files}
        for name in sys.argv[1:]:
            fx.update(parts[name](log))
    else:
        fx = build_s256()
        print(f"s256: {fx['s256_adjoint'].shape[0]} voxels", flush=True)
        fx.update(build_u16(log))
        fx.update(build_l512(log))
        fx.update(build_u128(log))
        fx.update(build_s256f(log))
        fx.update(build_bj(log))
        fx.update(build_std(log))
        fx.update(build_table_forwards(log))
    np.savez_compressed(OUT, **fx)
    print(f"wrote {OUT}: {len(fx)} arrays, {os.path.getsize(OUT)} bytes")
