"""CPU: pin the FP64 oracle (no GPU).

The reference ships no tests, fixtures or golden vectors, so the oracle is
pinned by (a) literal scalar restatements of the reference kernels
(oracle/brute.py), (b) autograd / dense-matrix transposition, (c) what the
mathematics fixes: trilinear exactness on affine fields, the closed-form
perspective position of a point object, chord lengths of a constant volume,
90-degree rotation equivariance, and the constant-sinogram response of the
reference backprojector.
"""
import math

import numpy as np
import pytest
import torch

from oracle import brute, cone
from dtvproj import geometry as G, phantom


def _rng(seed):
    return np.random.default_rng(seed)


# ---- (a) vectorised oracle == literal scalar loops ------------------------

@pytest.mark.parametrize("seed,sid,sdd,off", [(0, 10.0, 20.0, 3.0), (1, 14.0, 23.0, 45.0),
                                              (2, 9.0, 30.0, 0.0)])
def test_std_forward_matches_scalar_loop(seed, sid, sdd, off):
    X, Y, Z = 6, 5, 4
    U, V = 7, 5
    vol = _rng(seed).standard_normal((Z, Y, X))
    ang = np.linspace(0, 360, 6, False) + off
    o = cone.std_forward(torch.from_numpy(vol)[None, None], ang, (X, Y, Z), (U, V), sid, sdd)
    b = brute.std_forward(vol, ang, (X, Y, Z), (U, V), sid, sdd)
    np.testing.assert_allclose(o[0, 0].numpy(), b, rtol=0, atol=1e-12)


def test_bj_forward_matches_scalar_loop():
    X, Y, Z = 6, 7, 5
    U, V = 6, 5
    vol = _rng(3).standard_normal((Z, Y, X))
    P = G.beijing_default_vectors(3, 4, (U, V), 30.0, 50.0, (0.4, 0.5))
    P[:, 2] += 0.7   # off-plane source: exercises z-dominant logic paths
    o = cone.bj_forward(torch.from_numpy(vol)[None, None], P, (X, Y, Z), (U, V))
    b = brute.bj_forward(vol, P, (X, Y, Z), (U, V))
    np.testing.assert_allclose(o[0, 0].numpy(), b, rtol=0, atol=1e-12)


def test_std_bp_ref_matches_scalar_loop():
    """The oracle restates helper_geometry.h's general line intersection; the
    scalar loop uses the closed form u = SDD<c,ey>/L + U/2 -- they must agree."""
    X, Y, Z = 6, 5, 4
    U, V = 7, 5
    ang = np.linspace(0, 360, 5, False) + 3.0
    sino = _rng(4).standard_normal((5, V, U))
    o = cone.std_bp_ref(torch.from_numpy(sino)[None, None], ang, (X, Y, Z), (U, V), 10.0, 20.0)
    b = brute.std_bp_ref(sino, ang, (X, Y, Z), (U, V), 10.0, 20.0)
    np.testing.assert_allclose(o[0, 0].numpy(), b, rtol=1e-12, atol=1e-12)


def test_bj_bp_ref_matches_scalar_loop():
    X, Y, Z = 5, 6, 4
    U, V = 7, 5
    P = G.beijing_default_vectors(2, 21, (U, V), 30.0, 50.0, (0.2, 0.3))
    sino = _rng(5).standard_normal((42, V, U))
    o = cone.bj_bp_ref(torch.from_numpy(sino)[None, None], P, (X, Y, Z), (U, V))
    b = brute.bj_bp_ref(sino, P, (X, Y, Z), (U, V))
    np.testing.assert_allclose(o[0, 0].numpy(), b, rtol=1e-12, atol=1e-12)


# ---- (b) matched adjoint is the transpose --------------------------------

def test_adjoint_equals_autograd_vjp():
    X, Y, Z = 7, 6, 5
    ang = np.linspace(0, 360, 7, False) + 1.0
    args = ((X, Y, Z), 1.1, (9, 6), (1.3, 1.2), 30.0, 55.0)
    v = torch.from_numpy(_rng(6).standard_normal((2, 1, Z, Y, X))).requires_grad_(True)
    y = cone.circ_forward(v, ang, *args)
    g = torch.randn(y.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(7))
    y.backward(g)
    adj = cone.circ_adjoint(g, ang, *args)
    assert float((adj - v.grad).abs().max()) < 1e-12


def test_dense_matrix_transpose():
    """A built column by column from unit volumes; A^T built column by column
    from unit sinograms; A^T must equal A's transpose exactly."""
    X, Y, Z = 4, 4, 3
    U, V = 5, 4
    ang = np.array([0.0, 37.0, 121.0])
    args = (ang, (X, Y, Z), (U, V), 8.0, 16.0)
    nvox, nray = X * Y * Z, len(ang) * U * V
    A = np.zeros((nray, nvox))
    for j in range(nvox):
        e = torch.zeros(1, 1, Z, Y, X, dtype=torch.float64)
        e.view(-1)[j] = 1.0
        A[:, j] = cone.std_forward(e, *args).reshape(-1).numpy()
    AT = np.zeros((nvox, nray))
    for i in range(nray):
        e = torch.zeros(1, 1, len(ang), V, U, dtype=torch.float64)
        e.view(-1)[i] = 1.0
        AT[:, i] = cone.std_adjoint(e, *args).reshape(-1).numpy()
    assert np.abs(A).sum() > 0
    np.testing.assert_allclose(AT, A.T, rtol=0, atol=1e-13)


# ---- (c) what the mathematics fixes ---------------------------------------

def test_trilinear_exact_on_affine_field():
    X, Y, Z = 6, 5, 4
    a = np.array([0.3, -1.1, 0.7])
    b0 = 0.25
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    vol = torch.from_numpy(a[0] * xx + a[1] * yy + a[2] * zz + b0).reshape(-1)
    p = torch.from_numpy(_rng(8).uniform(0, 1, (200, 3)) * np.array([X - 1, Y - 1, Z - 1]))
    flat, w = cone._taps(p, (X, Y, Z))
    val = (w * vol[flat]).sum(1).numpy()
    np.testing.assert_allclose(val, p.numpy() @ a + b0, atol=1e-12)


def _project_point(ang_deg, c_mm, U, V, du, dv, sod, sdd):
    """Closed-form perspective projection of a world point (mm) onto the
    circular geometry's detector, as continuous pixel-centre indices."""
    th = ang_deg / 180 * G.PI_REF - G.PI_REF / 2
    ex = np.array([math.cos(th), math.sin(th), 0.0])
    ey = np.array([-math.sin(th), math.cos(th), 0.0])
    src = -ex * sod
    w = c_mm - src
    t = sdd / w.dot(ex)
    q = src + t * w
    return q.dot(ey) / du + U / 2 - 0.5, q[2] / dv + V / 2 - 0.5


def test_point_object_lands_at_perspective_projection():
    X, Y, Z = 16, 16, 12
    U, V = 40, 30
    vx, du, dv, sod, sdd = 1.0, 0.9, 1.1, 60.0, 130.0
    ix, iy, iz = 11, 4, 8
    vol = torch.zeros(1, 1, Z, Y, X, dtype=torch.float64)
    vol[0, 0, iz, iy, ix] = 1.0
    ang = np.array([0.0, 33.0, 90.0, 200.0, 311.0])
    sino = cone.circ_forward(vol, ang, (X, Y, Z), vx, (U, V), (du, dv), sod, sdd)[0, 0].numpy()
    c = np.array([(ix + 0.5 - X / 2) * vx, (iy + 0.5 - Y / 2) * vx, (iz + 0.5 - Z / 2) * vx])
    for a in range(len(ang)):
        img = sino[a]
        pv, pu = np.unravel_index(np.argmax(img), img.shape)
        # response-weighted centroid lands on the projected voxel centre
        m = img / img.sum()
        cu = (m.sum(0) * np.arange(U)).sum()
        cv = (m.sum(1) * np.arange(V)).sum()
        eu, ev = _project_point(ang[a], c, U, V, du, dv, sod, sdd)
        assert abs(pu - eu) <= 1.0 and abs(pv - ev) <= 1.0
        assert abs(cu - eu) < 0.25 and abs(cv - ev) < 0.25


def test_constant_volume_gives_chord_length():
    """Ray sums of a constant volume approximate the chord through the voxel
    box (trilinear ramps add at most half a voxel at each end)."""
    X, Y, Z = 20, 18, 16
    U, V = 25, 21
    vx, du, sod, sdd = 1.0, 1.0, 80.0, 160.0
    ang = np.array([10.0, 100.0, 233.0])
    vol = torch.ones(1, 1, Z, Y, X, dtype=torch.float64)
    sino = cone.circ_forward(vol, ang, (X, Y, Z), vx, (U, V), du, sod, sdd)[0, 0].numpy()
    lo = -np.array([X, Y, Z]) / 2 * vx
    hi = -lo
    for a in range(len(ang)):
        th = ang[a] / 180 * G.PI_REF - G.PI_REF / 2
        ex = np.array([math.cos(th), math.sin(th), 0.0])
        ey = np.array([-math.sin(th), math.cos(th), 0.0])
        src = -ex * sod
        for iv in range(V):
            for iu in range(U):
                pos = ex * (sdd - sod) + ey * (iu + 0.5 - U / 2) * du + \
                    np.array([0, 0, 1.0]) * (iv + 0.5 - V / 2) * du
                d = pos - src
                d /= np.linalg.norm(d)
                with np.errstate(divide="ignore"):
                    t0 = (lo - src) / d
                    t1 = (hi - src) / d
                tn = np.max(np.minimum(t0, t1))
                tf = np.min(np.maximum(t0, t1))
                chord = max(0.0, tf - tn)
                assert abs(sino[a, iv, iu] - chord) <= 1.0 * vx + 1e-9


def test_rotation_equivariance_90_degrees():
    """Rotating the object by +90 deg about z equals shifting the views by
    -90 deg (symmetric voxel grid of the physical geometry)."""
    X = Y = 12
    Z = 6
    U, V = 17, 9
    vol = torch.from_numpy(_rng(9).standard_normal((1, 1, Z, Y, X)))
    # p' = R p with R = +90 deg: (x, y) -> (-y, x); new[y', x'] = old[y, x]
    # with x' = X-1-y, y' = x  ->  new = rot90 of old over (y, x) axes
    rot = torch.flip(torch.transpose(vol, 3, 4), dims=(4,))
    ang = np.linspace(0, 360, 8, False) + 7.0
    args = ((X, Y, Z), 1.0, (U, V), 1.0, 40.0, 90.0)
    s_rot = cone.circ_forward(rot, ang, *args)
    s_shift = cone.circ_forward(vol, ang - 90.0, *args)
    np.testing.assert_allclose(s_rot.numpy(), s_shift.numpy(), rtol=0, atol=1e-9)


def test_std_bp_ref_constant_sinogram_closed_form():
    """A constant sinogram c backprojects to sum_a c (SDD/L_a)^2 pi / A for
    voxels whose (u, v) falls strictly inside the detector on every view."""
    X, Y, Z = 8, 8, 4
    U, V = 40, 30
    sid, sdd = 50.0, 100.0
    ang = np.linspace(0, 360, 6, False)
    c = 2.5
    sino = torch.full((1, 1, 6, V, U), c, dtype=torch.float64)
    out = cone.std_bp_ref(sino, ang, (X, Y, Z), (U, V), sid, sdd)[0, 0].numpy()
    s = sid / sdd
    exp = np.zeros((Y, X))
    for a in ang:
        th = a / 180 * G.PI_REF - G.PI_REF / 2
        ex = np.array([math.cos(th), math.sin(th)])
        iy, ix = np.meshgrid(np.arange(Y), np.arange(X), indexing="ij")
        L = sid + ((ix - X / 2) * s) * ex[0] + ((iy - Y / 2) * s) * ex[1]
        exp += c * (sdd / L) ** 2 * G.PI_REF / len(ang)
    for k in range(Z):
        np.testing.assert_allclose(out[k], exp, rtol=1e-12)


def test_float32_restatement_close_to_float64():
    """The CPU baseline runs the same oracle in FP32."""
    g = (np.linspace(0, 360, 6, False), (16, 16, 12), 1.0, (24, 20), 1.0, 60.0, 120.0)
    vol = phantom.gaussian((1, 1, 12, 16, 16), 1)
    a = cone.circ_forward(vol, *g)
    b = cone.circ_forward(vol, *g, dtype=torch.float32)
    assert float((a - b.double()).norm() / a.norm()) < 1e-5


def test_reference_backprojectors_at_selected_voxels():
    """The voxel-subset evaluation of both reference backprojectors is the
    full evaluation read at those voxels, bit for bit."""
    rng = np.random.default_rng(0)
    vs, ds = (10, 8, 6), (12, 7)
    ang = np.array([0.0, 90.0, 37.5, 200.0, 315.0])
    y = phantom.gaussian((2, 1, 5, 7, 12), 3, dtype=torch.float64)
    vox = np.stack([rng.integers(0, 10, 30), rng.integers(0, 8, 30), rng.integers(0, 6, 30)], 1)
    full = cone.std_bp_ref(y, ang, vs, ds, 60.0, 120.0)
    sub = cone.std_bp_ref(y, ang, vs, ds, 60.0, 120.0, voxels=vox)
    assert torch.equal(sub, full[:, 0][:, vox[:, 2], vox[:, 1], vox[:, 0]])
    vs, ds = (8, 8, 6), (6, 5)
    P = G.beijing_default_vectors(2, 21, ds, 40.0, 70.0, (0.6, 0.7))
    y = phantom.gaussian((1, 1, P.shape[0], 5, 6), 4, dtype=torch.float64)
    vox = np.stack([rng.integers(0, 8, 20), rng.integers(0, 8, 20), rng.integers(0, 6, 20)], 1)
    full = cone.bj_bp_ref(y, P, vs, ds)
    sub = cone.bj_bp_ref(y, P, vs, ds, voxels=vox)
    assert torch.equal(sub, full[:, 0][:, vox[:, 2], vox[:, 1], vox[:, 0]])


@pytest.mark.parametrize("case", [
    ((8, 8, 8), 1.0, (12, 10), 1.2, 6, 40.0, 80.0),
    # odd detector: the central column's ray has a zero direction component
    ((9, 7, 5), 1.1, (13, 9), 1.0, 5, 30.0, 120.0),
    ((16, 12, 10), 0.7, (21, 15), (0.9, 1.3), 7, 25.0, 60.0),
])
def test_matched_adjoint_at_selected_voxels(case):
    """``circ_adjoint_at`` (the per-voxel evaluation used for the full-size
    fixtures) equals the full matched adjoint at every voxel."""
    vs, vx, ds, pitch, n, sod, sdd = case
    X, Y, Z = vs
    U, V = ds
    ang = G.uniform_angles(n)
    y = phantom.gaussian((2, 1, n, V, U), 5, dtype=torch.float64)
    full = cone.circ_adjoint(y, ang, vs, vx, ds, pitch, sod, sdd).reshape(2, -1)
    vox = np.stack(np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij"),
                   -1).reshape(-1, 3)
    sub = cone.circ_adjoint_at(y, ang, vs, vx, ds, pitch, sod, sdd, vox)
    ref = full[:, (vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]]
    assert float((sub - ref).abs().max()) <= 1e-13 * float(ref.abs().max())


def _table_cases():
    """View tables of three schemes: a Beijing scan (tilted sub-detectors,
    source off the mid-plane), the Standard scheme and a circular orbit with
    the source close to the volume."""
    vs = (11, 9, 7)
    P = G.beijing_default_vectors(2, 21, (10, 6), 30.0, 50.0, (0.6, 0.7))
    bj = G.beijing(P, vs, (10, 6))
    st = G.standard(G.uniform_angles(5) + 3.0, vs, (13, 8), 40.0, 80.0)
    ci = G.circular(G.uniform_angles(6) + 1.0, vs, 1.0, (15, 9), 1.1, 14.0, 30.0)
    return [("beijing", bj), ("standard", st), ("circular", ci)]


@pytest.mark.parametrize("name,g", _table_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_table_adjoint_at_selected_voxels(name, g):
    """``table_adjoint_at`` (the per-voxel evaluation for any view table, used
    for the full-size Beijing fixture) equals the full ``table_adjoint`` at
    every voxel of the volume."""
    X, Y, Z = g.volume_size
    U, V = g.detector_size
    sc = g.scheme
    y = phantom.gaussian((2, 1, g.n_views, V, U), 8, dtype=torch.float64)
    args = (g.views, g.volume_size, g.detector_size, sc.axis_mode, sc.box_lo, sc.box_hi,
            sc.margin, sc.scale)
    full = cone.table_adjoint(y, *args).reshape(2, -1)
    vox = np.stack(np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij"),
                   -1).reshape(-1, 3)
    sub = cone.table_adjoint_at(y, *args, voxels=vox)
    ref = full[:, (vox[:, 2] * Y + vox[:, 1]) * X + vox[:, 0]]
    assert float(ref.abs().max()) > 0
    assert float((sub - ref).abs().max()) <= 1e-13 * float(ref.abs().max())


@pytest.mark.parametrize("name,g", _table_cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_table_forward_at_selected_rays(name, g):
    """``table_forward_at`` (rays built for the chosen rays' views only)
    equals the full ``table_forward`` at the chosen rays, bitwise."""
    X, Y, Z = g.volume_size
    U, V = g.detector_size
    sc = g.scheme
    args = (g.views, g.volume_size, g.detector_size, sc.axis_mode, sc.box_lo, sc.box_hi,
            sc.margin, sc.scale)
    x = phantom.gaussian((2, 1, Z, Y, X), 9, dtype=torch.float64)
    full = cone.table_forward(x, *args).reshape(2, -1)
    idx = np.random.default_rng(4).choice(g.n_views * U * V, size=60, replace=False)
    sub = cone.table_forward_at(x, *args, rays_idx=idx)
    assert torch.equal(sub, full[:, torch.from_numpy(idx)])


@pytest.mark.parametrize("case", [
    ((8, 8, 8), 1.0, (12, 10), 1.2, 6, 40.0, 80.0),
    ((9, 7, 5), 1.1, (13, 9), 1.0, 5, 30.0, 120.0),
])
def test_forward_at_selected_rays(case):
    """``circ_forward_at`` (the per-ray evaluation used for the full-size
    fixtures) equals the full forward at the chosen rays, bitwise (the same
    samples in the same order per ray)."""
    vs, vx, ds, pitch, n, sod, sdd = case
    X, Y, Z = vs
    U, V = ds
    ang = G.uniform_angles(n)
    x = phantom.gaussian((2, 1, Z, Y, X), 6, dtype=torch.float64)
    full = cone.circ_forward(x, ang, vs, vx, ds, pitch, sod, sdd).reshape(2, -1)
    idx = np.random.default_rng(3).choice(n * U * V, size=40, replace=False)
    sub = cone.circ_forward_at(x, ang, vs, vx, ds, pitch, sod, sdd, idx)
    assert torch.equal(sub, full[:, torch.from_numpy(idx)])
