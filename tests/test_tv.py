"""TV operators: torch restatement vs the reference's sparse matrices (CPU)
and the fused gfx950 step vs the FP64 formulation (GPU)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

from oracle import tv as otv
from dtvproj import tv, phantom

SIZES = [(5, 4, 3), (6, 6, 4), (4, 7, 5)]


@pytest.mark.parametrize("vs", SIZES)
@pytest.mark.parametrize("mode", ["reference", "neumann"])
@pytest.mark.parametrize("axis", ["x", "y", "z"])
def test_difference_operators_match_sparse_matrices(vs, mode, axis):
    X, Y, Z = vs
    f = phantom.gaussian((2, 1, Z, Y, X), 3, dtype=torch.float64)
    M = (otv.reference_matrix if mode == "reference" else otv.neumann_matrix)(vs, axis)
    np.testing.assert_allclose(tv.D(f, axis, mode).numpy(), otv.apply(M, f.numpy()), atol=1e-13)
    np.testing.assert_allclose(tv.DT(f, axis, mode).numpy(), otv.apply(M.T.tocsr(), f.numpy()),
                               atol=1e-13)


@pytest.mark.parametrize("mode", ["reference", "neumann"])
def test_pdhg_dtv_step_torch_matches_sparse_matrices(mode):
    """The torch PDHG step (reference for the fused kernels) against the
    same update written with the reference's sparse matrices."""
    X, Y, Z = 6, 5, 4
    shp = (2, 1, Z, Y, X)
    f, fbar, atw = (phantom.gaussian(shp, s, dtype=torch.float64) for s in (41, 42, 43))
    p = [phantom.gaussian(shp, 44 + a, dtype=torch.float64) for a in range(3)]
    sigma, tau, lam = 0.6, 0.3, (0.4, 0.7, 0.25)
    fn, fb, pn = tv.pdhg_dtv_step_torch(f, fbar, atw, p, sigma, tau, lam, mode)
    build = otv.reference_matrix if mode == "reference" else otv.neumann_matrix
    Ms = [build((X, Y, Z), a) for a in "xyz"]
    want_p = [np.clip(p[a].numpy() + sigma * otv.apply(Ms[a], fbar.numpy()), -lam[a], lam[a])
              for a in range(3)]
    grad = atw.numpy() + sum(otv.apply(Ms[a].T.tocsr(), want_p[a]) for a in range(3))
    want_f = np.maximum(f.numpy() - tau * grad, 0.0)
    for a in range(3):
        np.testing.assert_allclose(pn[a].numpy(), want_p[a], atol=1e-13)
    np.testing.assert_allclose(fn.numpy(), want_f, atol=1e-13)
    np.testing.assert_allclose(fb.numpy(), 2 * want_f - f.numpy(), atol=1e-13)
    assert (want_f == 0).any() and (np.abs(want_p[0]) == lam[0]).any()


def test_reference_dy_uses_y_stride_quirk():
    """Gradients.py:21-23 strides D_y by volumeSize[1]; for X != Y that is
    not the y-neighbour."""
    X, Y, Z = 5, 3, 2
    f = torch.arange(X * Y * Z, dtype=torch.float64).reshape(1, 1, Z, Y, X)
    d = tv.D(f, "y", "reference").reshape(-1)
    assert float(d[0]) == 3.0          # f[0 + Y] - f[0], Y = 3 (a true y-step would give 5)
    assert float(d[-1]) == -float(f.reshape(-1)[-1])


@pytest.mark.parametrize("mode", ["reference", "neumann"])
def test_tv_eps_grad_matches_oracle_and_autograd(mode):
    vs = (6, 5, 4)
    f = phantom.gaussian((1, 1, 4, 5, 6), 4, dtype=torch.float64).requires_grad_(True)
    g = tv.tv_eps_grad(f.detach(), 0.1, mode)
    np.testing.assert_allclose(g.numpy(), otv.tv_eps_grad(f.detach().numpy(), 0.1, vs, mode),
                               atol=1e-12)
    tv.tv_eps(f, 0.1, mode).backward()
    np.testing.assert_allclose(g.numpy(), f.grad.numpy(), atol=1e-12)


def test_shrink():
    v = torch.tensor([-3.0, -0.5, 0.0, 0.2, 2.0])
    assert torch.equal(tv.shrink(v, 1.0), torch.tensor([-2.0, -0.0, 0.0, 0.0, 1.0]))


@pytest.mark.parametrize("mode", ["reference", "neumann"])
def test_oracle_asd_gradient_is_the_gradient_of_its_potential(mode):
    """The ASD-POCS TV direction D^T (D f / (|D f| + eps)) is the exact
    gradient of sum_c [r_c - eps log(1 + r_c / eps)]: FP64 central
    differences of that potential pin the oracle's formula."""
    vs = (5, 4, 3)
    f = phantom.gaussian((1, 1, 3, 4, 5), 12, dtype=torch.float64).numpy()
    g = otv.asd_tv_grad(f, 1.0, vs, mode)
    rng = np.random.default_rng(5)
    for _ in range(4):
        v = rng.normal(size=f.shape)
        h = 1e-5
        fd = (otv.asd_potential(f + h * v, 1.0, vs, mode)
              - otv.asd_potential(f - h * v, 1.0, vs, mode)) / (2 * h)
        assert abs(fd - float((g * v).sum())) <= 1e-7 * max(1.0, abs(fd))
    # and a step along -g lowers the potential
    assert otv.asd_potential(f - 1e-3 * g, 1.0, vs, mode) < otv.asd_potential(f, 1.0, vs, mode)


@pytest.mark.parametrize("mode", ["reference", "neumann"])
def test_asd_tv_grad_matches_oracle_and_autograd(mode):
    vs = (6, 5, 4)
    f = phantom.gaussian((2, 1, 4, 5, 6), 13, dtype=torch.float64).requires_grad_(True)
    g = tv.asd_tv_grad(f.detach(), 0.7, mode)
    np.testing.assert_allclose(g.numpy(), otv.asd_tv_grad(f.detach().numpy(), 0.7, vs, mode),
                               atol=1e-12)
    tv.asd_potential(f, 0.7, mode).backward()
    np.testing.assert_allclose(g.numpy(), f.grad.numpy(), atol=1e-12)
    step = tv.tv_descent_torch(f.detach(), 0.3, 0.7, mode)
    np.testing.assert_allclose(step.numpy(), f.detach().numpy() - 0.3 * g.numpy(), atol=1e-14)


def _l1_projection_exact(v, b):
    """Euclidean projection onto {w : sum|w| <= b} by sorting (Duchi et al.,
    ICML 2008): the threshold is the lam with sum relu(|v| - lam) = b."""
    a = np.sort(np.abs(v).ravel())[::-1]
    cs = np.cumsum(a)
    k = np.arange(1, a.size + 1)
    rho = np.nonzero(a - (cs - b) / k > 0)[0][-1]
    lam = (cs[rho] - b) / (rho + 1)
    return np.sign(v) * np.maximum(np.abs(v) - lam, 0)


def test_oracle_l1ball_is_the_l1_projection():
    rng = np.random.default_rng(3)
    for n, b in ((50, 3.0), (2000, 100.0), (2000, 1500.0)):
        v = rng.normal(size=n)
        w = otv.l1ball(v, b)
        np.testing.assert_allclose(w, _l1_projection_exact(v, b), atol=1e-6)
        assert abs(np.abs(w).sum() - b) <= 1e-4
    # inside the ball: unchanged (L1ball.py:5-6)
    v = rng.normal(size=30)
    assert np.array_equal(otv.l1ball(v, np.abs(v).sum() + 1.0), v)


def test_l1ball_matches_oracle():
    v = phantom.gaussian((1, 1, 6, 7, 8), 14, dtype=torch.float64)
    for b in (5.0, 60.0, 1e4):
        want = otv.l1ball(v.numpy(), b)
        np.testing.assert_allclose(tv.l1ball_torch(v, b).numpy(), want, atol=1e-12)
        got32 = tv.l1ball_torch(v.float(), b).double().numpy()
        assert np.linalg.norm(got32 - want) <= 1e-5 * max(np.linalg.norm(want), 1e-30)


def test_power_method_bounds_and_approaches_the_top_singular_value():
    vs = (5, 4, 3)
    M = otv.reference_matrix(vs, "y").toarray()
    smax = np.linalg.svd(M, compute_uv=False)[0]
    from dtvproj import unroll
    Mt = torch.from_numpy(M)
    op = lambda x: (Mt @ x.reshape(-1)).reshape(x.shape)
    op_t = lambda x: (Mt.T @ x.reshape(-1)).reshape(x.shape)
    s = unroll.power_method(op, op_t, (1, 1, 3, 4, 5), iters=200, dtype=torch.float64)
    assert s <= smax * (1 + 1e-12)
    assert s >= 0.99 * smax


def test_asd_pocs_without_tv_steps_is_landweber_to_the_solution():
    """ng = 0 leaves the data steps: f <- f - beta A^T (A f - b), which for a
    consistent, well-conditioned system with a positive solution converges
    to that solution (the final relu then changes nothing)."""
    from dtvproj import unroll
    rng = np.random.default_rng(8)
    A = torch.from_numpy(rng.normal(size=(40, 24)) + 3 * np.eye(40, 24))
    x = torch.from_numpy(rng.uniform(0.5, 1.5, size=(1, 1, 2, 3, 4)))
    b = (A @ x.reshape(-1)).reshape(1, 1, 40)
    proj = lambda f: (A @ f.reshape(-1)).reshape(1, 1, 40)
    back = lambda y: (A.T @ y.reshape(-1)).reshape(1, 1, 2, 3, 4)
    L = float(torch.linalg.matrix_norm(A, 2)) ** 2
    out = unroll.asd_pocs(torch.zeros_like(x), b, proj, back, cascades=1500, beta=1.0 / L,
                          beta_red=1.0, ng=0, fused=False)
    assert float((out - x).abs().max()) <= 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["neumann", "reference"])
@pytest.mark.parametrize("shp", [(2, 1, 13, 11, 17), (1, 1, 20, 16, 16), (1, 1, 35, 70, 70),
                                 (2, 1, 3, 9, 130), (1, 1, 19, 64, 64), (1, 1, 35, 204, 204),
                                 (2, 1, 3, 20, 512)])
def test_fused_fista_tv_step(cuda, mode, shp):
    """The fused step vs the torch formulation (whose D and D^T meet the
    sparse matrices); the shapes cover several x-y tiles and z chunks of the
    z-marching kernels, one voxel per lane and four (X = 204, 512), in
    Neumann and reference mode (X == Y), and the per-voxel kernel (reference
    mode with X != Y)."""
    print(f"[case] fused_fista_tv_step {mode} {shp}", flush=True)
    y = phantom.gaussian(shp, 1)
    g = phantom.gaussian(shp, 2)
    xp = phantom.gaussian(shp, 3)
    args = dict(step=0.3, lam=0.7, eps=0.05, beta=0.4, nonneg=True, mode=mode)
    ref_x, ref_y = tv.fista_tv_step_torch(y.double(), g.double(), xp.double(), **args)
    x, yn = tv.fista_tv_step(y.to(cuda), g.to(cuda), xp.to(cuda), **args)
    assert rel_l2(x, ref_x) <= 1e-6
    assert rel_l2(yn, ref_y) <= 1e-6
    # x_prev may share storage with the output x
    xp_dev = xp.to(cuda)
    x2, _ = tv.fista_tv_step(y.to(cuda), g.to(cuda), xp_dev, xnew=xp_dev, **args)
    assert rel_l2(x2, ref_x) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "neumann"])
@pytest.mark.parametrize("shp", [(2, 1, 13, 11, 17), (1, 1, 1, 1, 5), (1, 1, 9, 1, 1),
                                 (3, 1, 4, 6, 5), (1, 1, 20, 16, 16), (2, 1, 35, 70, 70),
                                 (2, 1, 35, 68, 68), (1, 1, 4, 204, 204)])
def test_fused_tv_descent_matches_oracle(cuda, mode, shp):
    print(f"[case] tv_descent {mode} {shp}", flush=True)
    f = phantom.gaussian(shp, 21)
    Z, Y, X = shp[-3:]
    want = f.double().numpy() - 0.25 * otv.asd_tv_grad(f.numpy(), 1.0, (X, Y, Z), mode)
    got = tv.tv_descent(f.to(cuda), 0.25, 1.0, mode)
    assert rel_l2(got, torch.from_numpy(want)) <= 1e-6
    with pytest.raises(RuntimeError, match="share storage"):
        fd = f.to(cuda)
        tv.tv_descent(fd, 0.25, 1.0, mode, out=fd)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["neumann", "reference"])
@pytest.mark.parametrize("shp", [(2, 1, 13, 11, 17), (1, 1, 1, 1, 5), (1, 1, 9, 1, 1),
                                 (3, 1, 4, 6, 5), (2, 1, 5, 12, 16), (1, 1, 3, 8, 260),
                                 (1, 1, 2, 3, 4)])
def test_fused_pdhg_dtv_step(cuda, mode, shp):
    """The two fused launches vs the torch step in FP64 (whose D and D^T are
    checked against the sparse matrices above); the clamp and the relu both
    act on these inputs."""
    print(f"[case] pdhg_dtv_step {mode} {shp}", flush=True)
    f = phantom.gaussian(shp, 31) - 0.3
    fbar = phantom.gaussian(shp, 32)
    atw = phantom.gaussian(shp, 33)
    p = torch.stack([phantom.gaussian(shp, 34 + a) - 0.4 for a in range(3)])
    sigma, tau, lam = 0.7, 0.4, (0.3, 0.5, 0.2)
    rf, rfb, rp = tv.pdhg_dtv_step_torch(f.double(), fbar.double(), atw.double(), p.double(),
                                         sigma, tau, lam, mode)
    if f.numel() >= 100:
        assert float((rf == 0).double().mean()) > 0.05
        assert float((rp[0].abs() == lam[0]).double().mean()) > 0.05
    fd, fbd, pd = f.to(cuda), fbar.to(cuda), p.to(cuda)
    tv.pdhg_dtv_step(fd, fbd, atw.to(cuda), pd, sigma, tau, lam, mode)
    assert rel_l2(fd, rf) <= 1e-6
    assert rel_l2(fbd, rfb) <= 1e-6
    for a in range(3):
        assert rel_l2(pd[a], rp[a]) <= 1e-6
    with pytest.raises(RuntimeError, match="share storage"):
        tv.pdhg_dtv_step(fd, fd, atw.to(cuda), pd, sigma, tau, lam, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 5, 4096, 1 << 21])
def test_l1ball_on_the_gpu_matches_oracle(cuda, n):
    """dtv_l1ball (the Newton iteration on the device) vs the FP64 oracle
    restatement of L1ball.py:4-17, from one value to 2M values (every
    workgroup of the reduction busy), inside and outside the ball."""
    print(f"[case] l1ball gpu n={n}", flush=True)
    v = phantom.gaussian((n,), 22)
    total = float(v.double().abs().sum())
    for b in (0.0, 0.01 * total, 0.5 * total, 1.1 * total):
        want = torch.from_numpy(otv.l1ball(v.double().numpy(), b))
        got = tv.l1ball(v.to(cuda), b)
        if b == 0.0:
            # the radius-0 ball: what remains is below the stopping tolerance
            assert float(got.double().abs().sum()) <= 1e-3
            assert float(want.abs().sum()) <= 1e-3
        else:
            assert rel_l2(got, want) <= 1e-5, b
        if b > total:
            assert torch.equal(got.cpu(), v)   # inside: v itself (L1ball.py:5-6)


@pytest.mark.gpu
def test_l1ball_batched_and_dual_update(cuda):
    """Three vectors of different size and radius in one call, as DTV-CP's
    dual update (mode 1): p <- p - sigma sign(p) l1ball(|p| / sigma, b),
    against the same composition on the FP64 oracle; and the projection mode
    on the same vectors."""
    print("[case] l1ball batched dual", flush=True)
    sigma = 1e-3
    ps = [phantom.gaussian((1, 1, 5, 6, 7), 41) * 1e-3, phantom.gaussian((70000,), 42) * 2e-3,
          phantom.gaussian((3, 4), 43) * 1e-3]
    radii = [0.2 * float((p.double().abs() / sigma).sum()) for p in ps[:2]] + [1e9]
    want = []
    for p, b in zip(ps, radii):
        pd = p.double().numpy()
        want.append(pd - sigma * np.sign(pd) * otv.l1ball(np.abs(pd) / sigma, b))
    got = tv.l1ball_dual([p.to(cuda) for p in ps], sigma, radii)
    for g, w in zip(got, want):
        if np.linalg.norm(w) == 0.0:
            assert float(g.abs().max()) <= 1e-9
        else:
            assert rel_l2(g, torch.from_numpy(w)) <= 1e-5
    outs = [torch.empty_like(p).to(cuda) for p in ps]
    tv._l1ball_device([p.to(cuda) for p in ps], radii, outs, 0)
    for o, p, b in zip(outs, ps, radii):
        w = torch.from_numpy(otv.l1ball(p.double().numpy(), b))
        assert rel_l2(o, w) <= 1e-5


@pytest.mark.gpu
def test_l1ball_iteration_count_and_tolerance_follow_the_reference(cuda):
    """max_iter and tol are the reference's loop controls: with max_iter = 1
    only the first Newton step (from the sums at lam = 0) is taken, as in the
    oracle run with max_iter = 1."""
    print("[case] l1ball max_iter", flush=True)
    v = phantom.gaussian((10000,), 44)
    b = 0.3 * float(v.double().abs().sum())
    for it in (1, 2, 3):
        want = torch.from_numpy(otv.l1ball(v.double().numpy(), b, max_iter=it))
        got = tv.l1ball(v.to(cuda), b, max_iter=it)
        assert rel_l2(got, want) <= 1e-5
    full = torch.from_numpy(otv.l1ball(v.double().numpy(), b))
    assert rel_l2(tv.l1ball(v.to(cuda), b, max_iter=1), full) > 1e-3


@pytest.mark.gpu
def test_fused_shrink_step_closed_form(cuda):
    print("[case] shrink step", flush=True)
    y = phantom.gaussian((2, 1, 9, 10, 11), 51)
    g = phantom.gaussian((2, 1, 9, 10, 11), 52)
    xp = phantom.gaussian((2, 1, 9, 10, 11), 53)
    step, thr, beta = 0.3, 0.25, 0.6
    d = y.double() - step * g.double()
    x = torch.sign(d) * torch.clamp(d.abs() - thr, min=0)
    assert float((x == 0).double().mean()) > 0.1
    yn = x + beta * (x - xp.double())
    gx, gy = tv.shrink_step(y.to(cuda), g.to(cuda), xp.to(cuda), step, thr, beta)
    assert rel_l2(gx, x) <= 1e-6 and rel_l2(gy, yn) <= 1e-6
    xpd = xp.to(cuda)
    gx2, gy2 = tv.shrink_step(y.to(cuda), g.to(cuda), xpd, step, thr, beta, xnew=xpd)
    assert rel_l2(gx2, x) <= 1e-6 and rel_l2(gy2, yn) <= 1e-6
    # outputs are checked before the launch: size, dtype, aliasing
    yd, gd = y.to(cuda), g.to(cuda)
    small = torch.empty(5, device=cuda)
    with pytest.raises(RuntimeError, match="shape"):
        tv.shrink_step(yd, gd, xpd, step, thr, beta, xnew=small)
    with pytest.raises(RuntimeError, match="shape"):
        tv.shrink_step(yd, gd, xpd, step, thr, beta, ynew=torch.empty_like(yd).double())
    both = torch.empty_like(yd)
    with pytest.raises(RuntimeError, match="distinct"):
        tv.shrink_step(yd, gd, xpd, step, thr, beta, xnew=both, ynew=both)


def _fdpg_fp64(f, g, sigma, v, t, beta, mode):
    """One DTVFISTA.py:29-41 step in FP64 on the explicit sparse matrices."""
    Z, Y, X = f.shape[-3:]
    mats = [(otv.reference_matrix if mode == "reference" else otv.neumann_matrix)((X, Y, Z), a)
            for a in "xyz"]
    fb = f.reshape(-1, X * Y * Z)
    z = np.maximum(fb - sigma * g.reshape(-1, X * Y * Z), 0.0)
    d = [np.stack([M @ zi for zi in z]) for M in mats]
    m = [float((np.abs(da) - va).max()) for da, va in zip(d, v)]
    fnew = z.copy()
    for M, da, ma, ta in zip(mats, d, m, t):
        fnew = fnew + ta * np.stack([M.T @ (np.sign(di) * ma) for di in da])
    return (fnew + beta * (fnew - fb)).reshape(f.shape)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["reference", "neumann"])
@pytest.mark.parametrize("shp", [(2, 1, 7, 9, 11), (1, 1, 1, 1, 6), (1, 1, 5, 70, 3)])
def test_fused_fdpg_step_matches_sparse_matrices(cuda, mode, shp):
    print(f"[case] fdpg step {mode} {shp}", flush=True)
    f = phantom.gaussian(shp, 61) + 0.5
    g = phantom.gaussian(shp, 62)
    sigma, v, t, beta = 0.3, (0.4, 0.2, 0.7), (0.05, 0.08, 0.03), 0.25
    want = _fdpg_fp64(f.double().numpy(), g.double().numpy(), sigma, v, t, beta, mode)
    got = tv.fdpg_step(f.to(cuda), g.to(cuda), sigma, v, t, beta, mode)
    assert rel_l2(got, torch.from_numpy(want)) <= 1e-6
    got_t = tv.fdpg_step_torch(f.double(), g.double(), sigma, v, t, beta, mode)
    assert rel_l2(got_t, torch.from_numpy(want)) <= 1e-12
    with pytest.raises(RuntimeError, match="shape"):
        tv.fdpg_step(f.to(cuda), g.to(cuda), sigma, v, t, beta, mode,
                     out=torch.empty(3, device=cuda))


def test_asd_pocs_step_schedule(monkeypatch):
    """The TV step is alpha * mean|first data step| and, as in the reference
    (f0 aliases f there), never shrinks; with adaptive=True it shrinks by
    alpha_red after a cascade whose TV change exceeds rmax times its data
    change."""
    from dtvproj import unroll
    rng = np.random.default_rng(9)
    A = torch.from_numpy(rng.normal(size=(30, 24)))
    x = torch.from_numpy(rng.uniform(0.5, 1.5, size=(1, 1, 2, 3, 4)))
    b = (A @ x.reshape(-1)).reshape(1, 1, 30)
    proj = lambda f: (A @ f.reshape(-1)).reshape(1, 1, 30)
    back = lambda y: (A.T @ y.reshape(-1)).reshape(1, 1, 2, 3, 4)
    seen = []
    real = tv.tv_descent_torch
    monkeypatch.setattr(tv, "tv_descent_torch",
                        lambda f, step, eps=1.0, mode="reference": (seen.append(step),
                                                                    real(f, step, eps, mode))[1])
    kw = dict(cascades=3, beta=1e-3, ng=2, alpha=50.0, fused=False)
    f0 = torch.zeros_like(x)
    unroll.asd_pocs(f0, b, proj, back, **kw)
    assert len(seen) == 6 and len(set(seen)) == 1 and seen[0] > 0
    dp0 = float((1e-3 * back(proj(f0) - b)).abs().mean())
    assert abs(seen[0] - 50.0 * dp0) <= 1e-12 * seen[0]
    seen.clear()
    unroll.asd_pocs(f0, b, proj, back, adaptive=True, alpha_red=0.5, **kw)
    # a large TV step moves f far more than the data step: every cascade halves it
    assert seen[0:2] == [seen[0]] * 2 and abs(seen[2] - 0.5 * seen[0]) <= 1e-15 * seen[0]
    assert abs(seen[4] - 0.25 * seen[0]) <= 1e-15 * seen[0]


def test_dtv_cp_without_regularisation_solves_the_data_term():
    """With v = 0 and mu = 0 the DTV Chambolle-Pock cascade is a primal-dual
    iteration on 1/2 ||A f - b||^2 alone (dual w, over-relaxed primal); for a
    consistent, well-conditioned system it converges to the solution."""
    from dtvproj import unroll
    rng = np.random.default_rng(8)
    A = torch.from_numpy(rng.normal(size=(40, 24)) + 3 * np.eye(40, 24))
    x = torch.from_numpy(rng.uniform(0.5, 1.5, size=(1, 1, 2, 3, 4)))
    b = (A @ x.reshape(-1)).reshape(1, 1, 40)
    proj = lambda f: (A @ f.reshape(-1)).reshape(1, 1, 40)
    back = lambda y: (A.T @ y.reshape(-1)).reshape(1, 1, 2, 3, 4)
    L = float(torch.linalg.matrix_norm(A, 2)) ** 2
    out = unroll.dtv_cp(torch.zeros_like(x), b, proj, back, cascades=2000, v=(0.0, 0.0, 0.0),
                        mu=0.0, sigma=0.5 / L, tau=0.5 / L, fused=False)
    assert float((out - x).abs().max()) <= 1e-8
