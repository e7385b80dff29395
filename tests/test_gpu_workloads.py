"""GPU: the full-size workloads against FP64-oracle fixtures
(tests/golden/oracle_workloads.npz, written by tests/golden/make_workloads.py).

* the matched A^T over all 60 views of the 256^3 sparse-view workload (the
  headline call: two 32-view chunks and many fixed-point view groups per
  brick) at 4096 seeded voxels, rel L2 <= 1e-5;
* A over all 60 views of the 256^3 workload at 4096 seeded rays;
* the 512x512x384 / 120-view workload: A at 4096 seeded rays and the matched
  A^T at 4096 seeded voxels, rel L2 <= 1e-5 (default launches, the
  deterministic forward and a split A^T);
* the 4 x 128^3 / 90-view unroll geometry: A at 2048 rays and A^T at 2048
  voxels of the seeded item of a batch of four;
* the 3024-view Beijing scan (256x256x64, 144 angles x 21 planes of 80x128):
  the matched A^T over all views at 1024 seeded voxels, partitioned (view
  parts through the workspace) and unpartitioned;
* the reference's Standard options geometry (256x256x64, 48 views of
  288x28): the matched A^T over all views at 1024 seeded voxels;
* A over all views of both: 4096 seeded rays each, FP32 atomics and the
  deterministic fixed-point sums;
* the 16-instance uneven FISTA-TV workload: 20 iterations batched in one
  launch per operator, eager and hipGraph replay, against the plain
  per-volume loop (rel L2 <= 1e-6), and its smallest-work instance against
  the same 20 iterations run in FP64 on the oracle (RMSE / peak <= 1e-4)."""
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2

from dtvproj import batched, configs, geometry as G, ops, phantom, unroll

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "oracle_workloads.npz")


@pytest.fixture(scope="module")
def wfx():
    with np.load(FIX) as f:
        return {k: f[k] for k in f.files}


def _case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.mark.parametrize("launch", [None, dict(algorithm=1),
                                    dict(algorithm=2, brick=(16, 16, 8), block=256)])
def test_sparse256_full_adjoint_vs_oracle_fixture(cuda, wfx, launch):
    g = configs.sparse_256()
    _case("sparse_256 full A^T vs oracle voxels", launch=launch)
    y = phantom.gaussian(g.sino_shape(1), 21)
    y64 = y.double()
    assert abs(float(y64.sum()) - float(wfx["s256_sino_sum"])) <= 1e-9 * float(wfx["s256_sino_sumsq"])
    assert abs(float((y64 * y64).sum()) / float(wfx["s256_sino_sumsq"]) - 1.0) <= 1e-12
    vol = ops.adjoint(y.to(cuda), g, launch)
    vox = torch.from_numpy(wfx["s256_voxels"].astype(np.int64))
    X, Y, Z = g.volume_size
    flat = (vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]
    got = vol.reshape(-1)[flat.to(cuda)]
    ref = torch.from_numpy(wfx["s256_adjoint"])
    assert rel_l2(got, ref) <= 1e-5
    # the 512 face voxels on their own (brick edges and the volume border)
    assert rel_l2(got[-512:], ref[-512:]) <= 1e-5


@pytest.mark.parametrize("launch", [None, dict(algorithm=2, brick=(24, 24, 14), block=256,
                                               view_parts=1)])
def test_sparse256_full_forward_vs_oracle_fixture(cuda, wfx, launch):
    """The headline A over all 60 views (default: 26x26x14 bricks, views split
    over three workgroups per brick) at 4096 seeded rays, FP32 atomics and the
    deterministic fixed-point sums."""
    g = configs.sparse_256()
    _case("sparse_256 full A vs oracle rays", launch=launch)
    x = phantom.gaussian(g.volume_shape(1), 20)
    assert abs(float(x.double().sum()) - float(wfx["s256f_vol_sum"])) <= 1e-6
    rays = torch.from_numpy(wfx["s256f_rays"]).to(cuda)
    ref = torch.from_numpy(wfx["s256f_forward"])
    xd = x.to(cuda)
    for det in (False, True):
        sino = ops.forward(xd, g, launch, deterministic=det)
        assert rel_l2(sino.reshape(-1)[rays], ref) <= 1e-5, f"deterministic={det}"


def test_large512_forward_and_adjoint_vs_oracle_fixture(cuda, wfx):
    g = configs.large_512()
    X, Y, Z = g.volume_size
    _case("large_512 A at rays, A^T at voxels vs oracle")
    x = phantom.gaussian(g.volume_shape(1), 20)
    x64 = x.double()
    assert abs(float(x64.sum()) - float(wfx["l512_vol_sum"])) <= 1e-9 * float(wfx["l512_vol_sumsq"])
    assert abs(float((x64 * x64).sum()) / float(wfx["l512_vol_sumsq"]) - 1.0) <= 1e-12
    del x64
    rays = torch.from_numpy(wfx["l512_rays"]).to(cuda)
    ref = torch.from_numpy(wfx["l512_forward"])
    xd = x.to(cuda)
    del x
    for det in (False, True):
        sino = ops.forward(xd, g, deterministic=det)
        assert rel_l2(sino.reshape(-1)[rays], ref) <= 1e-5, f"deterministic={det}"
        del sino
    del xd
    y = phantom.gaussian(g.sino_shape(1), 21)
    y64 = y.double()
    assert abs(float(y64.sum()) - float(wfx["l512_sino_sum"])) <= 1e-9 * float(wfx["l512_sino_sumsq"])
    assert abs(float((y64 * y64).sum()) / float(wfx["l512_sino_sumsq"]) - 1.0) <= 1e-12
    del y64
    vox = torch.from_numpy(wfx["l512_voxels"].astype(np.int64))
    flat = ((vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]).to(cuda)
    aref = torch.from_numpy(wfx["l512_adjoint"])
    yd = y.to(cuda)
    for launch in (None, dict(algorithm=2, brick=(32, 16, 16), block=256, view_parts=3)):
        vol = ops.adjoint(yd, g, launch)
        got = vol.reshape(-1)[flat]
        assert rel_l2(got, aref) <= 1e-5, f"launch={launch}"
        assert rel_l2(got[-512:], aref[-512:]) <= 1e-5
        del vol


def test_unroll128_batch4_vs_oracle_fixture(cuda, wfx):
    """The 4 x 128^3 / 90-view unroll geometry at full size, the seeded
    volume (and sinogram) as item 2 of a batch of four: A at 2048 seeded rays
    and the matched A^T at 2048 seeded voxels."""
    g = configs.unroll_128()
    X, Y, Z = g.volume_size
    _case("unroll_128 batch 4 vs oracle rays / voxels")
    x = phantom.gaussian(g.volume_shape(1), 30)
    assert abs(float(x.double().sum()) - float(wfx["u128_vol_sum"])) <= 1e-6
    xb = torch.cat([phantom.gaussian(g.volume_shape(1), 40 + k) for k in range(2)] + [x] +
                   [phantom.gaussian(g.volume_shape(1), 43)]).to(cuda)
    rays = torch.from_numpy(wfx["u128_rays"]).to(cuda)
    ref = torch.from_numpy(wfx["u128_forward"])
    for det in (False, True):
        sino = ops.forward(xb, g, deterministic=det)
        assert rel_l2(sino[2].reshape(-1)[rays], ref) <= 1e-5, f"deterministic={det}"
    y = phantom.gaussian(g.sino_shape(1), 31)
    assert abs(float(y.double().sum()) - float(wfx["u128_sino_sum"])) <= 1e-6
    yb = torch.cat([phantom.gaussian(g.sino_shape(1), 50 + k) for k in range(2)] + [y] +
                   [phantom.gaussian(g.sino_shape(1), 53)]).to(cuda)
    vox = torch.from_numpy(wfx["u128_voxels"].astype(np.int64))
    flat = ((vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]).to(cuda)
    aref = torch.from_numpy(wfx["u128_adjoint"])
    two = dict(algorithm=2, brick=(16, 16, 16), block=256, view_parts=2)
    for launch, part in ((None, True), (two, True), (two, False)):
        vol = ops.adjoint(yb, g, launch, partition=part)
        assert rel_l2(vol[2].reshape(-1)[flat], aref) <= 1e-5, f"launch={launch} {part}"


def test_uneven16_batched_fista_tv(cuda, wfx):
    geoms = configs.uneven_16()
    g2 = geoms[2]
    assert g2.volume_size == (103, 103, 75) and g2.detector_size == (153, 91)
    assert g2.n_views == 38 and g2.ref["pitch_mm"] == (2.0, 2.0)
    _case("uneven_16 batched FISTA-TV", instances=len(geoms), iters=20)
    b = []
    for i, g in enumerate(geoms):
        ph = phantom.phantom(g.volume_size, 1.0, 10, seed=100 + i).to(cuda)
        b.append(ops.forward(ph, g))
    lam, eps, iters = 1e-3, 1e-3, 20
    steps = [1.0 / (unroll.power_norm(g, iters=8, device=cuda) + lam * 12.0 / eps)
             for g in geoms]
    steps[2] = float(wfx["u16_step"])
    plain = batched.plain_fista_tv(geoms, b, iters, lam, eps, steps)
    bf = batched.BatchedFistaTV(geoms, b, iters, lam, eps, steps)
    bf.reset()
    eager = [x.clone() for x in bf.run()]
    for i, (e, p) in enumerate(zip(eager, plain)):
        assert rel_l2(e, p) <= 1e-6, f"instance {i}"
    bf.capture()
    for rep in range(2):
        torch.cuda.synchronize()
        junk = torch.randn(1000, device=cuda).sum()   # eager work between replays
        out = bf.replay()
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(out, eager)):
            assert rel_l2(o, e) <= 1e-6, f"replay {rep} instance {i}"
        float(junk)
    # the smallest-work instance against the FP64 oracle iteration
    x = eager[2].reshape(-1).double().cpu()
    idx = torch.from_numpy(wfx["u16_voxels"])
    ref = torch.from_numpy(wfx["u16_values"])
    peak = float(wfx["u16_peak"])
    rmse = float(((x[idx] - ref) ** 2).mean().sqrt())
    print(f"uneven_16 instance 2: RMSE/peak {rmse / peak:.3e}", flush=True)
    assert rmse / peak <= 1e-4
    assert abs(float(x.sum()) - float(wfx["u16_sum"])) <= 1e-4 * float(wfx["u16_sum"])
    assert abs(float(x.abs().max()) - peak) <= 1e-4 * peak


def test_beijing_scan_adjoint_vs_oracle_fixture(cuda, wfx):
    """The matched A^T over all 3024 views of the Beijing scan (each brick
    culls the views it cannot reach; the default splits the rest over eight
    parts through the workspace) against the FP64 oracle at 1024 seeded
    voxels, 128 of them on the volume's faces."""
    X, Y, Z, U, V = 256, 256, 64, 80, 128
    mag = 1143.0 / 708.0
    pitch = (1.2 * np.hypot(X, Y) * mag / (U * 21), 1.2 * Z * mag / V)
    g = G.beijing(G.beijing_default_vectors(144, 21, (U, V), 708.0, 1143.0, pitch), (X, Y, Z),
                  (U, V))
    y = phantom.gaussian(g.sino_shape(1), 41)
    # the same input (sums agree up to the summation order of the CPU threads)
    assert abs(float(y.double().sum()) - float(wfx["bj_sino_sum"])) <= 1e-12 * float(
        wfx["bj_sino_sumsq"])
    vox = wfx["bj_voxels"].astype(np.int64)
    flat = torch.from_numpy((vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]).to(cuda)
    ref = torch.from_numpy(wfx["bj_adjoint"])
    yd = y.to(cuda)
    for part in (True, False):
        _case("beijing 3024-view A^T vs oracle voxels", partition=part)
        got = ops.adjoint(yd, g, partition=part).reshape(-1)[flat]
        assert rel_l2(got, ref) <= 1e-5, f"partition={part}"
        assert rel_l2(got[-128:], ref[-128:]) <= 1e-5, f"faces, partition={part}"


def test_standard_options_adjoint_vs_oracle_fixture(cuda, wfx):
    """The matched A^T over all 48 views of the reference's Standard options
    geometry (options.yaml: 256x256x64, 288x28 detector, SID 1000 / SDD 2000
    pixels), partitioned and unpartitioned, against the FP64 oracle at 1024
    seeded voxels; the voxels no ray reaches (the detector's 28 rows see a
    slab of the volume) are exact zeros."""
    g = G.standard(G.uniform_angles(48), (256, 256, 64), (288, 28), 1000.0, 2000.0)
    X, Y, Z = g.volume_size
    y = phantom.gaussian(g.sino_shape(1), 42)
    assert abs(float(y.double().sum()) - float(wfx["std_sino_sum"])) <= 1e-12 * float(
        wfx["std_sino_sumsq"])
    vox = wfx["std_voxels"].astype(np.int64)
    flat = torch.from_numpy((vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]).to(cuda)
    ref = torch.from_numpy(wfx["std_adjoint"])
    yd = y.to(cuda)
    for part in (True, False):
        _case("standard options A^T vs oracle voxels", partition=part)
        got = ops.adjoint(yd, g, partition=part).reshape(-1)[flat].cpu()
        assert rel_l2(got, ref) <= 1e-5, f"partition={part}"
        assert bool((got[ref == 0] == 0).all()), f"zeros, partition={part}"


@pytest.mark.parametrize("key", ["bjf", "stdf"])
def test_table_geometry_forward_vs_oracle_fixture(cuda, wfx, key):
    """A over all views of the 3024-view Beijing scan and of the Standard
    options geometry (default launch: 512-lane workgroups, view parts), with
    FP32 atomics and with the deterministic fixed-point sums, against the
    FP64 oracle at 4096 seeded rays."""
    if key == "bjf":
        X, Y, Z, U, V = 256, 256, 64, 80, 128
        mag = 1143.0 / 708.0
        pitch = (1.2 * np.hypot(X, Y) * mag / (U * 21), 1.2 * Z * mag / V)
        g = G.beijing(G.beijing_default_vectors(144, 21, (U, V), 708.0, 1143.0, pitch),
                      (X, Y, Z), (U, V))
    else:
        g = G.standard(G.uniform_angles(48), (256, 256, 64), (288, 28), 1000.0, 2000.0)
    X, Y, Z = g.volume_size
    x = phantom.gaussian((1, 1, Z, Y, X), 40)
    assert abs(float(x.double().sum()) - float(wfx[f"{key}_vol_sum"])) <= 1e-12 * float(
        wfx[f"{key}_vol_sumsq"])
    rays = torch.from_numpy(wfx[f"{key}_rays"]).to(cuda)
    ref = torch.from_numpy(wfx[f"{key}_forward"])
    xd = x.to(cuda)
    for det in (False, True):
        _case("table-geometry A vs oracle rays", key=key, deterministic=det)
        got = ops.forward(xd, g, deterministic=det).reshape(-1)[rays].cpu()
        assert rel_l2(got, ref) <= 1e-5, f"deterministic={det}"
        assert bool((got[ref == 0] == 0).all())
