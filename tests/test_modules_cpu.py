"""Module layer on the CPU: FBP weights and Ram-Lak taps pinned by their
geometric / spectral meaning, the Beijing plane-interleave permutes, the
``.mat`` geometry ingest and the raw float32 I/O and loaders
(model/ConeBeamLayers/*/*.py, data/*.py)."""
import os

import numpy as np
import pytest
import torch


def _std():
    import model.ConeBeamLayers.Standard.StandardGeometry as S
    return S


def _bj():
    import model.ConeBeamLayers.Beijing.BeijingGeometry as B
    return B


# ---- FBP weights -------------------------------------------------------------

def test_default_adjoint_mode_is_the_reference_wiring():
    """Reference callers keep their step sizes: both drop-in modules default
    to the reference's backprojector as A^T."""
    S, B = _std(), _bj()
    assert S.adjoint_mode == "reference" and B.adjoint_mode == "reference"


def test_standard_cos_weight_is_cosine_to_the_central_ray():
    S = _std()
    from options import standardSDD
    w = S.StandardGeometryWithFBP().cosWeight.numpy()[0, 0, 0]
    V, U = w.shape
    iv, iu = np.meshgrid(np.arange(V) - V / 2, np.arange(U) - U / 2, indexing="ij")
    rays = np.stack([np.full_like(iu, standardSDD), iu, iv], -1)
    cos = (rays @ np.array([1.0, 0.0, 0.0])) / np.linalg.norm(rays, axis=-1)
    np.testing.assert_allclose(w, cos, rtol=2e-7)
    assert w[V // 2, U // 2] == 1.0
    # same weight for every view
    all_w = S.StandardGeometryWithFBP().cosWeight.numpy()[0, 0]
    assert np.all(all_w == all_w[:1])


def test_beijing_cos_weight_keeps_the_reference_index_swap():
    B = _bj()
    from options import beijingSDD, beijingSubDetectorSize, beijingPlanes
    U, V = beijingSubDetectorSize
    W = U * beijingPlanes
    w = B.BeijingGeometryWithFBP().cosWeight.numpy()[0, 0, 0]
    assert w.shape == (V, W)
    # the row index is paired with the row *length* centre W/2 and the column
    # index with V/2 (BeijingGeometry.py:68-71): the largest weight sits in
    # the last row (closest to W/2 > V) at column V/2, not at the centre
    i, j = np.unravel_index(np.argmax(w), w.shape)
    assert (i, j) == (V - 1, V // 2)
    d = np.hypot(V - 1 - W / 2, 0.0)
    assert np.isclose(w[V - 1, V // 2], np.cos(np.arctan2(d, beijingSDD)), rtol=1e-6)


@pytest.mark.parametrize("width", [288, 1680])
def test_ram_lak_taps_have_a_ramp_spectrum(width):
    S = _std()
    h = S.StandardGeometryWithFBP().__conv__(width).numpy().reshape(-1).astype(np.float64)
    assert h.size == width + 1
    mid = width // 2
    assert h[mid] == np.float32(1.0 / 8)
    assert np.all(h[mid + 2::2] == 0) and np.all(h[mid - 2::-2] == 0)
    # symmetric, and its frequency response is |w| / (4 pi) on [-pi, pi]
    # up to the truncation of the 1/n^2 tail
    assert np.array_equal(h, h[::-1])
    n = np.arange(width + 1) - mid
    for om in np.linspace(0.05, np.pi, 9):
        H = np.sum(h * np.cos(om * n))
        assert abs(H - om / (4 * np.pi)) < 2.0 / (np.pi ** 2 * mid)


def test_beijing_ramp_matches_standard_taps():
    S, B = _std(), _bj()
    a = S.StandardGeometryWithFBP().__conv__(160).numpy()
    b = B.BeijingGeometryWithFBP().__conv__(160).numpy()
    assert np.array_equal(a, b)


# ---- Beijing layout and geometry ingest ------------------------------------------

def test_beijing_plane_interleave_permutes():
    B = _bj()
    from options import beijingAngleNum as A, beijingPlanes as P, beijingSubDetectorSize
    U, V = beijingSubDetectorSize
    t = torch.arange(2 * A * P * V * U, dtype=torch.float32).reshape(2, 1, A * P, V, U)
    pub = B._to_public(t)
    assert pub.shape == t.shape
    src = t.reshape(2, A, P, V, U)
    dst = pub.reshape(2, A, V, P, U)
    rng = np.random.default_rng(0)
    for _ in range(200):
        b, a, p, v, u = (int(rng.integers(0, n)) for n in (2, A, P, V, U))
        assert dst[b, a, v, p, u] == src[b, a, p, v, u]
    assert torch.equal(B._to_projector(pub), t)


def test_projection_matrix_ingest(tmp_path, monkeypatch):
    import scipy.io as sco
    B = _bj()
    from data import realLoader
    m = np.random.default_rng(1).standard_normal((42, 12))
    p = str(tmp_path / "geom.mat")
    sco.savemat(p, {"projection_matrix": m})
    got = realLoader.load_projection_matrix(p)
    assert got.dtype == torch.float32 and got.shape == (42, 12) and got.is_contiguous()
    np.testing.assert_array_equal(got.numpy(), m.astype(np.float32))
    # the module's own loader takes the file when it exists
    monkeypatch.setattr(B, "beijingParameterRoot", p)
    np.testing.assert_array_equal(B._load_parameters().numpy(), m.astype(np.float32))
    sco.savemat(p, {"projection_matrix": m[:, :11]})
    with pytest.raises(ValueError, match="expected \\[views, 12\\]"):
        realLoader.load_projection_matrix(p)


def test_synthetic_parameters_have_the_reference_layout():
    B = _bj()
    from options import beijingAngleNum, beijingPlanes
    P = B.parameters
    assert P.dtype == torch.float32 and P.shape == (beijingAngleNum * beijingPlanes, 12)
    # the planes of one angle share the source
    src = P.reshape(beijingAngleNum, beijingPlanes, 12)[:, :, 0:3]
    assert torch.all(src == src[:, :1])


# ---- raw float32 I/O and loaders --------------------------------------------------

def test_raw_roundtrip_scale_and_size_check(tmp_path):
    from dtvproj import rawio
    a = np.random.default_rng(2).standard_normal((1, 4, 5, 6)).astype(np.float32)
    p = str(tmp_path / "v.raw")
    rawio.write_raw(p, torch.from_numpy(a))
    assert os.path.getsize(p) == a.nbytes
    np.testing.assert_array_equal(rawio.read_raw(p, a.shape), a)
    np.testing.assert_array_equal(np.fromfile(p, "float32").reshape(a.shape), a)
    np.testing.assert_array_equal(rawio.read_raw(p, a.shape, mmap=True), a)
    np.testing.assert_allclose(rawio.read_raw(p, a.shape, scale=6e4), a * np.float32(6e4))
    with pytest.raises(ValueError, match="expected 484 for float32"):
        rawio.read_raw(p, (1, 11, 11))


class _Projector:
    """CPU stand-in for the Beijing ForwardProjection (tests the loaders'
    bookkeeping only; the GPU tests run the real projector)."""

    @staticmethod
    def apply(x):
        return x.sum(dim=(-1,), keepdim=True) * 2.0


def test_simulate_loaders(tmp_path, monkeypatch):
    from data import simulateLoader as L
    from dtvproj import rawio
    monkeypatch.setattr(L, "ForwardProjection", _Projector)
    scale = (1, 3, 4, 5)
    (tmp_path / "lab").mkdir()
    (tmp_path / "inp").mkdir()
    rng = np.random.default_rng(3)
    vols = {}
    for name in ("a.raw", "b.raw"):
        v = rng.standard_normal(scale).astype(np.float32)
        rawio.write_raw(str(tmp_path / "lab" / name), v)
        rawio.write_raw(str(tmp_path / "inp" / name), v * 0.5)
        vols[name] = v
    ds = L.Stimulated256(str(tmp_path / "lab"), "cpu", scale)
    assert len(ds) == 2
    inp, proj, img = ds[0]
    name = os.path.basename(ds.dir[0])
    assert inp.shape == img.shape == scale and torch.all(inp == 0)
    np.testing.assert_array_equal(img.numpy(), vols[name])
    np.testing.assert_array_equal(proj.numpy(), _Projector.apply(torch.from_numpy(vols[name])[None])[0])
    ds2 = L.Stimulated256Input(str(tmp_path / "lab"), str(tmp_path / "inp"), "cpu", scale)
    inp, proj, img = ds2[1]
    name = os.path.basename(ds2.dir[1])
    np.testing.assert_array_equal(inp.numpy(), vols[name] * 0.5)


def test_real_loader_scales_clips_and_switches_geometry(tmp_path):
    import scipy.io as sco
    B = _bj()
    from data import realLoader as R
    from dtvproj import rawio
    scale, pscale = (1, 2, 3, 4), (1, 6, 2, 3)
    for d in ("label", "input", "matlab", "sino"):
        (tmp_path / "s1" / d).mkdir(parents=True)
    rng = np.random.default_rng(4)
    lab = rng.standard_normal(scale).astype(np.float32)
    sino = rng.standard_normal(pscale).astype(np.float32)
    rawio.write_raw(str(tmp_path / "s1" / "label" / "x.raw"), lab)
    rawio.write_raw(str(tmp_path / "s1" / "input" / "x.raw"), lab * 2)
    rawio.write_raw(str(tmp_path / "s1" / "sino" / "x.raw"), sino)
    m = rng.standard_normal((6, 12))
    sco.savemat(str(tmp_path / "s1" / "matlab" / "x.mat"), {"projection_matrix": m})
    old = B.parameters
    try:
        ds = R.Real256Input([str(tmp_path / "s1")], "cpu", scale, pscale)
        inp, proj, img = ds[0]
        np.testing.assert_allclose(img.numpy(), np.maximum(lab * np.float32(6e4), 0))
        np.testing.assert_allclose(inp.numpy(), np.maximum(lab * 2 * np.float32(6e4), 0))
        np.testing.assert_allclose(proj.numpy(), sino * np.float32(6e4))
        np.testing.assert_array_equal(B.parameters.numpy(), m.astype(np.float32))
    finally:
        B.parameters = old


# ---- drop-in next to a checkout of the callers -------------------------------

def test_packages_merge_with_a_callers_checkout(tmp_path):
    """The callers' tree (no __init__ files) keeps model.PDHG, model.FISTA and
    data.* beside model.ConeBeamLayers; with this package first on sys.path
    those resolve from the checkout while the projector modules are these."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.join(root, "dtv-net_amd")
    (tmp_path / "model" / "PDHG").mkdir(parents=True)
    (tmp_path / "model" / "ConeBeamLayers" / "Beijing").mkdir(parents=True)
    (tmp_path / "data").mkdir()
    (tmp_path / "model" / "PDHG" / "Caller.py").write_text(
        "from model.ConeBeamLayers.Beijing.BeijingGeometry import ForwardProjection, "
        "BackProjection\n")
    (tmp_path / "model" / "ConeBeamLayers" / "Beijing" / "BeijingGeometry.py").write_text(
        "raise ImportError('the checkout copy must not be imported')\n")
    (tmp_path / "data" / "evalLoader.py").write_text("VALUE = 7\n")
    code = (
        "import sys; sys.path[:0] = [%r, %r]\n"
        "import model.PDHG.Caller as c, data.evalLoader as e\n"
        "import model.ConeBeamLayers.Beijing.BeijingGeometry as B\n"
        "assert c.ForwardProjection is B.ForwardProjection\n"
        "assert B.__file__.startswith(%r), B.__file__\n"
        "assert e.VALUE == 7\n"
        "from data.simulateLoader import Stimulated256\n"
        "print('ok')\n" % (pkg, str(tmp_path), pkg))
    out = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path),
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout


# ---- sparse-view selection of Beijing scans ----------------------------------

def test_angle_groups_match_the_reference_slicings():
    from dtvproj import geometry as G
    a = np.arange(1080)
    # fista_eval.py: projection[j*45:(j+1)*45:45//6+1] for 24 blocks
    want = np.concatenate([a[j * 45:(j + 1) * 45:45 // 6 + 1] for j in range(24)])
    assert np.array_equal(G.angle_groups(24, 45, 6, 8), want)
    # eval_generateSparseView.py: projection[j*45:j*45+3]
    want = np.concatenate([a[j * 45:j * 45 + 3] for j in range(24)])
    assert np.array_equal(G.angle_groups(24, 45, 3), want)
    with pytest.raises(ValueError):
        G.angle_groups(24, 45, 7, 8)


def test_subsample_beijing_layout_and_geometry():
    """The selected views equal the reference's reshape-to-[angles, planes]
    and slice, for arrays and tensors, and the geometry built from the
    selected vectors is the full geometry's view subset."""
    from dtvproj import geometry as G
    n_ang, planes, U, V = 12, 21, 6, 5
    rng = np.random.default_rng(4)
    proj = rng.normal(size=(1, 1, n_ang * planes, V, U)).astype(np.float32)
    vec = G.beijing_default_vectors(n_ang, planes, (U, V), 40.0, 70.0, (0.6, 0.7))
    ang = G.angle_groups(3, 4, 2, 2)                      # angles 0, 2, 4, 6, 8, 10
    sp, sv = G.subsample_beijing(proj, vec, ang, planes)
    want_p = proj.reshape(n_ang, planes, V, U)[ang].reshape(1, 1, -1, V, U)
    want_v = vec.reshape(n_ang, planes, 12)[ang].reshape(-1, 12)
    assert np.array_equal(sp, want_p) and np.array_equal(sv, want_v)
    tp, tv_ = G.subsample_beijing(torch.from_numpy(proj), torch.from_numpy(vec), ang, planes)
    assert torch.equal(tp, torch.from_numpy(want_p)) and torch.equal(tv_, torch.from_numpy(want_v))
    full = G.beijing(vec, (8, 8, 6), (U, V))
    views = (ang[:, None] * planes + np.arange(planes)[None, :]).reshape(-1)
    assert np.array_equal(G.beijing(sv, (8, 8, 6), (U, V)).views, full.subset(views).views)
    with pytest.raises(ValueError):
        G.subsample_beijing(proj, vec, [n_ang], planes)
