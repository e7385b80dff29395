"""GPU: the module layer of the drop-in (reference modules and loaders) on
the gfx950 kernels, against compositions built from the FP64 oracle.

Module globals are patched down to small geometries so the oracle finishes
in seconds; the code paths are the ones the reference configuration runs."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle import cone

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _correlate_rows(x, taps, weight=None):
    """FP64 conv2d(x * weight, taps, padding=(0, (len-1)/2)) over rows."""
    x = np.asarray(x, dtype=np.float64)
    W = x.shape[-1]
    rows = x.reshape(-1, W).copy()
    if weight is not None:
        w = np.asarray(weight, dtype=np.float64).reshape(-1, W)
        rows *= np.tile(w, (rows.shape[0] // w.shape[0], 1))
    t = np.asarray(taps, dtype=np.float64).reshape(-1)
    pad = (t.size - 1) // 2
    padded = np.pad(rows, ((0, 0), (pad, pad)))
    out = np.stack([np.correlate(r, t, mode="valid") for r in padded])
    return out.reshape(x.shape)


@pytest.mark.parametrize("W,ntaps,wrows", [(288, 289, 28), (1680, 1681, 5), (31, 31, 1),
                                           (100, 7, 0), (8, 17, 3)])
def test_row_filter_matches_fp64_correlation(cuda, W, ntaps, wrows):
    from dtvproj import fbp
    print(f"[case] row_filter W={W} ntaps={ntaps} wrows={wrows}", flush=True)
    g = torch.Generator().manual_seed(W + ntaps)
    rows = 3 * max(wrows, 1) * 2
    x = torch.randn((rows, W), generator=g)
    t = torch.randn(ntaps, generator=g)
    w = torch.rand((wrows, W), generator=g) + 0.5 if wrows else None
    out = fbp.row_filter(x.to(cuda), t, None if w is None else w.to(cuda))
    ref = _correlate_rows(x.numpy(), t.numpy(), None if w is None else w.numpy())
    assert rel_l2(out, torch.from_numpy(ref)) <= 1e-6


def test_row_filter_rejects_bad_arguments(cuda):
    from dtvproj import fbp
    x = torch.zeros((4, 16), device=cuda)
    with pytest.raises(RuntimeError, match="odd"):
        fbp.row_filter(x, torch.ones(16))
    with pytest.raises(RuntimeError, match="too large"):
        fbp.row_filter(torch.zeros((1, 9000), device=cuda), torch.ones(9001))


# ---- Standard modules -------------------------------------------------------

SMALL_STD = dict(vol=(24, 20, 12), det=(40, 14), n_views=12)


@pytest.fixture
def std_small(monkeypatch):
    import model.ConeBeamLayers.Standard.StandardGeometry as S
    X, Y, Z = SMALL_STD["vol"]
    U, V = SMALL_STD["det"]
    n = SMALL_STD["n_views"]
    monkeypatch.setattr(S, "angles", torch.from_numpy(np.linspace(0, 360, n, False)))
    monkeypatch.setattr(S, "volumeSize", torch.tensor([X, Y, Z]))
    monkeypatch.setattr(S, "detectorSize", torch.tensor([U, V]))
    monkeypatch.setattr(S, "standardAngleNum", n)
    return S


@pytest.mark.parametrize("mode", ["matched", "reference"])
def test_standard_fbp_module(cuda, std_small, monkeypatch, mode):
    S = std_small
    monkeypatch.setattr(S, "adjoint_mode", mode)
    print(f"[case] standard fbp module mode={mode}", flush=True)
    X, Y, Z = SMALL_STD["vol"]
    U, V = SMALL_STD["det"]
    n = SMALL_STD["n_views"]
    g = torch.Generator().manual_seed(5)
    x = torch.randn((1, 1, Z, Y, X), generator=g)
    p = torch.randn((1, 1, n, V, U), generator=g)
    m = S.StandardGeometryWithFBP().to(cuda)
    out = m(x.to(cuda), p.to(cuda))
    ang = S.angles.numpy()
    args = (ang, SMALL_STD["vol"], SMALL_STD["det"], S.standardSID, S.standardSDD)
    r = cone.std_forward(x, *args) - p.double()
    r = _correlate_rows(r.numpy(), m.ramp.cpu().numpy(), m.cosWeight[0, 0, 0].cpu().numpy())
    r = torch.from_numpy(r)
    back = cone.std_adjoint(r, *args) if mode == "matched" else cone.std_bp_ref(r, *args)
    assert rel_l2(out, back) <= TOL


@pytest.mark.parametrize("mode", ["matched", "reference"])
def test_standard_step_module(cuda, std_small, monkeypatch, mode):
    S = std_small
    monkeypatch.setattr(S, "adjoint_mode", mode)
    print(f"[case] standard step module mode={mode}", flush=True)
    X, Y, Z = SMALL_STD["vol"]
    U, V = SMALL_STD["det"]
    n = SMALL_STD["n_views"]
    g = torch.Generator().manual_seed(6)
    x = torch.randn((2, 1, Z, Y, X), generator=g)
    p = torch.randn((2, 1, n, V, U), generator=g)
    out = S.StandardGeometry().to(cuda)(x.to(cuda), p.to(cuda))
    args = (S.angles.numpy(), SMALL_STD["vol"], SMALL_STD["det"], S.standardSID, S.standardSDD)
    back = cone.std_adjoint if mode == "matched" else cone.std_bp_ref
    ref = x.double() - back(cone.std_forward(x, *args) - p.double(), *args)
    assert rel_l2(out, ref) <= TOL


@pytest.mark.parametrize("mode", ["matched", "reference"])
def test_standard_functions_backward(cuda, std_small, monkeypatch, mode):
    """autograd through the drop-in Functions, batch 2: the backward of
    ForwardProjection is A^T of the mode (the reference's FDK backprojector,
    StandardGeometry.py:17-21, or the matched transpose), the backward of
    BackProjection is A (:30-34)."""
    S = std_small
    monkeypatch.setattr(S, "adjoint_mode", mode)
    print(f"[case] standard functions backward mode={mode}", flush=True)
    X, Y, Z = SMALL_STD["vol"]
    U, V = SMALL_STD["det"]
    n = SMALL_STD["n_views"]
    g = torch.Generator().manual_seed(16)
    x = torch.randn((2, 1, Z, Y, X), generator=g)
    w = torch.randn((2, 1, n, V, U), generator=g)
    c = torch.randn((2, 1, Z, Y, X), generator=g)
    args = (S.angles.numpy(), SMALL_STD["vol"], SMALL_STD["det"], S.standardSID, S.standardSDD)
    back = cone.std_adjoint if mode == "matched" else cone.std_bp_ref
    xg = x.to(cuda).requires_grad_(True)
    out = S.ForwardProjection.apply(xg)
    assert rel_l2(out, cone.std_forward(x, *args)) <= TOL
    (gx,) = torch.autograd.grad((out * w.to(cuda)).sum(), xg)
    assert rel_l2(gx, back(w.double(), *args)) <= TOL
    wg = w.to(cuda).requires_grad_(True)
    vol = S.BackProjection.apply(wg)
    assert rel_l2(vol, back(w.double(), *args)) <= TOL
    vol.backward(c.to(cuda))
    assert rel_l2(wg.grad, cone.std_forward(c.double(), *args)) <= TOL


def test_standard_reference_caller_step_sizes(cuda, std_small, monkeypatch):
    """A reference caller iterating StandardGeometry (x <- x - lamb B(Ax - p),
    lamb = 1.0 as tuned for the FDK backprojector, StandardGeometry.py:36-43):
    in the default (reference) mode five iterates meet the same iteration on
    the oracle's A and B_ref; in matched mode the same step would diverge
    (||A^T A|| >> 2), and the documented rescaling lamb = 1 / ||A^T A||
    makes the iteration reduce the residual."""
    from dtvproj import geometry as G, unroll
    S = std_small
    print("[case] standard reference caller", flush=True)
    X, Y, Z = SMALL_STD["vol"]
    U, V = SMALL_STD["det"]
    n = SMALL_STD["n_views"]
    args = (S.angles.numpy(), SMALL_STD["vol"], SMALL_STD["det"], S.standardSID, S.standardSDD)
    x_true = torch.randn((1, 1, Z, Y, X), generator=torch.Generator().manual_seed(17)).abs()
    p64 = cone.std_forward(x_true.double(), *args)
    p = p64.float().to(cuda)
    m = S.StandardGeometry().to(cuda)
    x = torch.zeros_like(x_true).to(cuda)
    x64 = torch.zeros_like(x_true, dtype=torch.float64)
    for _ in range(5):
        x = m(x, p)
        x64 = x64 - 1.0 * cone.std_bp_ref(cone.std_forward(x64, *args) - p64, *args)
        assert rel_l2(x, x64) <= TOL
    monkeypatch.setattr(S, "adjoint_mode", "matched")
    geom = G.standard(S.angles.numpy(), SMALL_STD["vol"], SMALL_STD["det"], S.standardSID,
                      S.standardSDD)
    L = unroll.power_norm(geom, device=cuda)
    assert L > 2.0      # the reference step would diverge on the matched pair
    m.lamb = 1.0 / L
    x = torch.zeros_like(x_true).to(cuda)
    r0 = float(p.norm())
    for _ in range(20):
        x = m(x, p)
    assert float((S.ForwardProjection.apply(x) - p).norm()) < 0.5 * r0


# ---- Beijing modules --------------------------------------------------------

SMALL_BJ = dict(vol=(20, 20, 10), sub=(6, 8), angles=3, planes=21)


@pytest.fixture
def bj_small(monkeypatch):
    import model.ConeBeamLayers.Beijing.BeijingGeometry as B
    from dtvproj import geometry as G
    X, Y, Z = SMALL_BJ["vol"]
    U, V = SMALL_BJ["sub"]
    A, P = SMALL_BJ["angles"], SMALL_BJ["planes"]
    vec = G.beijing_default_vectors(A, P, (U, V), 60.0, 100.0, (0.12, 0.9))
    for name, val in (("parameters", torch.from_numpy(vec)),
                      ("volumeSize", torch.IntTensor([X, Y, Z])),
                      ("detectorSize", torch.IntTensor([U, V])),
                      ("beijingAngleNum", A), ("beijingPlanes", P),
                      ("beijingSubDetectorSize", [U, V]), ("beijingSDD", 100.0)):
        monkeypatch.setattr(B, name, val)
    return B, vec


def _bj_public(sino, A, P, V, U):
    """projector layout [A*P, V, U] -> the module's public interleave."""
    return sino.reshape(sino.shape[0], A, P, V, U).permute(0, 1, 3, 2, 4) \
        .reshape(sino.shape[0], 1, A * P, V, U)


def _bj_projector(sino, A, P, V, U):
    return sino.reshape(sino.shape[0], A, V, P, U).permute(0, 1, 3, 2, 4) \
        .reshape(sino.shape[0], 1, A * P, V, U)


@pytest.mark.parametrize("mode", ["matched", "reference"])
def test_beijing_step_and_fbp_modules(cuda, bj_small, monkeypatch, mode):
    B, vec = bj_small
    monkeypatch.setattr(B, "adjoint_mode", mode)
    print(f"[case] beijing modules mode={mode}", flush=True)
    X, Y, Z = SMALL_BJ["vol"]
    U, V = SMALL_BJ["sub"]
    A, P = SMALL_BJ["angles"], SMALL_BJ["planes"]
    si = B.sampleInterval
    g = torch.Generator().manual_seed(7)
    x = torch.randn((1, 1, Z, Y, X), generator=g)
    p = torch.randn((1, 1, A * P, V, U), generator=g)
    vs, ds = SMALL_BJ["vol"], SMALL_BJ["sub"]
    Ax = _bj_public(cone.bj_forward(x, vec, vs, ds) * si, A, P, V, U)

    def back(r):
        rp = _bj_projector(r, A, P, V, U)
        if mode == "matched":
            return cone.bj_adjoint(rp, vec, vs, ds) * si
        return cone.bj_bp_ref(rp, vec, vs, ds)

    step = B.BeijingGeometry().to(cuda)(x.to(cuda), p.to(cuda))
    assert rel_l2(step, -B.BeijingGeometry().lamb * back(Ax - p.double())) <= TOL
    m = B.BeijingGeometryWithFBP().to(cuda)
    out = m(x.to(cuda), p.to(cuda))
    r = (Ax - p.double()).reshape(1, 1, A, V, U * P).numpy()
    r = _correlate_rows(r, m.ramp.cpu().numpy(), m.cosWeight[0, 0, 0].cpu().numpy())
    r = torch.from_numpy(r).reshape(1, 1, A * P, V, U)
    assert rel_l2(out, x.double() - back(r)) <= TOL


@pytest.mark.parametrize("mode", ["matched", "reference"])
def test_beijing_functions_backward(cuda, bj_small, monkeypatch, mode):
    """autograd through the Beijing drop-in Functions, batch 2, with the
    public (angles, V, planes, U) interleave.  Reference mode keeps the
    reference's factor placement: sampleInterval in ForwardProjection.forward
    and BackProjection.backward, none in ForwardProjection.backward or
    BackProjection.forward (BeijingGeometry.py:17, 20-25, 29-32, 38)."""
    B, vec = bj_small
    monkeypatch.setattr(B, "adjoint_mode", mode)
    print(f"[case] beijing functions backward mode={mode}", flush=True)
    X, Y, Z = SMALL_BJ["vol"]
    U, V = SMALL_BJ["sub"]
    A, P = SMALL_BJ["angles"], SMALL_BJ["planes"]
    vs, ds = SMALL_BJ["vol"], SMALL_BJ["sub"]
    si = B.sampleInterval
    g = torch.Generator().manual_seed(18)
    x = torch.randn((2, 1, Z, Y, X), generator=g)
    w = torch.randn((2, 1, A * P, V, U), generator=g)      # public layout
    c = torch.randn((2, 1, Z, Y, X), generator=g)

    def fwd(v):
        return _bj_public(cone.bj_forward(v.double(), vec, vs, ds) * si, A, P, V, U)

    def back(s, with_si):
        rp = _bj_projector(s.double(), A, P, V, U)
        if mode == "matched":
            return cone.bj_adjoint(rp, vec, vs, ds) * si
        return cone.bj_bp_ref(rp, vec, vs, ds) * (si if with_si else 1.0)

    xg = x.to(cuda).requires_grad_(True)
    out = B.ForwardProjection.apply(xg)
    assert rel_l2(out, fwd(x)) <= TOL
    (gx,) = torch.autograd.grad((out * w.to(cuda)).sum(), xg)
    assert rel_l2(gx, back(w, with_si=False)) <= TOL
    wg = w.to(cuda).requires_grad_(True)
    vol = B.BackProjection.apply(wg)
    assert rel_l2(vol, back(w, with_si=False)) <= TOL
    vol.backward(c.to(cuda))
    assert rel_l2(wg.grad, fwd(c)) <= TOL


def test_simulate_loader_projects_on_gpu(cuda, bj_small, tmp_path):
    B, vec = bj_small
    from data import simulateLoader as L
    from dtvproj import rawio
    print("[case] simulate loader", flush=True)
    X, Y, Z = SMALL_BJ["vol"]
    scale = (1, Z, Y, X)
    v = torch.randn(scale, generator=torch.Generator().manual_seed(8)).numpy()
    (tmp_path / "lab").mkdir()
    rawio.write_raw(str(tmp_path / "lab" / "a.raw"), v)
    inp, proj, img = L.Stimulated256(str(tmp_path / "lab"), cuda, scale)[0]
    assert proj.device.type == "cpu" and torch.all(inp == 0)
    ref = B.ForwardProjection.apply(torch.from_numpy(v)[None].to(cuda))[0].cpu()
    assert torch.equal(proj, ref)


def test_sparse_view_selection_projects_the_view_subset(cuda):
    """Views are independent: projecting with the vectors of the selected
    angles gives the selected views of the full projection, and the
    selected-view projection meets the FP64 oracle."""
    from dtvproj import geometry as G, ops, phantom
    n_ang, planes, U, V = 8, 21, 10, 9
    vs = (16, 14, 10)
    vec = G.beijing_default_vectors(n_ang, planes, (U, V), 40.0, 70.0, (0.6, 0.7))
    ang = G.angle_groups(4, 2, 1)                         # every other angle
    print(f"[case] sparse view selection angles={list(ang)}", flush=True)
    x = phantom.gaussian((1, 1, vs[2], vs[1], vs[0]), 41)
    full = ops.forward(x.to(cuda), G.beijing(vec, vs, (U, V)))
    sub_full, sv = G.subsample_beijing(full, vec, ang, planes)
    sub = ops.forward(x.to(cuda), G.beijing(sv, vs, (U, V)))
    assert rel_l2(sub, sub_full) <= 1e-6
    assert rel_l2(sub, cone.bj_forward(x.double(), sv, vs, (U, V))) <= TOL
