"""CPU: the reference unroll cascades pinned by what they solve, independently
of their own formulas (tiny dense systems, FP64).

* DTV Chambolle-Pock (model/PDHG/DTVCP.py:29-60): with its radii held, its
  limit solves  min 1/2 ||A f - b||^2  s.t.  ||D_a f||_1 <= t_a (a = x, y,
  z), f >= 0; the constrained problem is solved by scipy's SLSQP in an
  epigraph form (|D_a f| <= s_a, sum s_a <= t_a).
* ASD-POCS (model/PDHG/ASDPOCS.py:28-47): each cascade is the data step
  f - beta A^T (A f - b) followed by ng gradient-descent steps on the
  potential sum_c [r_c - eps log(1 + r_c / eps)], r = |D f|, with the step
  fixed from the first cascade -- here the gradient comes from central
  differences of that potential, not from the closed form the product uses.
* FISTA with soft threshold (model/FISTA/FISTA.py:16-27): converges to the
  LASSO minimiser, here solved by scikit-learn's coordinate descent.
"""
import numpy as np
import pytest
import torch

from oracle import tv as otv
from dtvproj import unroll

VS = (4, 2, 2)        # X, Y, Z: 16 unknowns
SHAPE = (1, 1, 2, 2, 4)


def _system(seed, m):
    rng = np.random.default_rng(seed)
    n = int(np.prod(VS))
    A = rng.normal(size=(m, n)) + 2.0 * np.eye(m, n)
    x = rng.uniform(0.2, 1.0, size=n)
    b = A @ x + 0.3 * rng.normal(size=m)
    At = torch.from_numpy(A)
    proj = lambda f: (At @ f.reshape(-1)).reshape(1, 1, m)          # noqa: E731
    back = lambda y: (At.T @ y.reshape(-1)).reshape(SHAPE)          # noqa: E731
    return A, b, proj, back


def test_dtv_cp_solves_the_tv_constrained_least_squares():
    """With the radii held (refresh beyond the cascade count; the reference
    refreshes them from the iterate every 10 cascades, which lets them follow
    it), the cascade converges to the solution of the constrained problem at
    the radii of its start, found here by SLSQP.  v and mu scale the dual
    operators (5 sqrt(L), inside the step condition sigma tau ||K||^2 < 1)."""
    from scipy.optimize import minimize
    A, b, proj, back = _system(31, 24)
    n = A.shape[1]
    L = np.linalg.norm(A, 2) ** 2
    x_ls = np.linalg.lstsq(A, b, rcond=None)[0]
    rng = np.random.default_rng(36)
    f0 = torch.from_numpy((0.5 * x_ls + 0.2 + 0.05 * rng.normal(size=n)).reshape(SHAPE))
    mats = [otv.reference_matrix(VS, a).toarray() for a in "xyz"]
    t = [float(np.abs(M @ f0.reshape(-1).numpy()).sum()) for M in mats]
    # the least-squares solution breaks the radii: the constraints bind
    assert all(np.abs(M @ x_ls).sum() > 1.2 * ta for M, ta in zip(mats, t))
    w = 5.0 * np.sqrt(L)
    out = unroll.dtv_cp(f0, torch.from_numpy(b).reshape(1, 1, -1), proj, back, cascades=12000,
                        v=(w, w, w), mu=w, sigma=0.5 / L, tau=0.5 / L, mode="reference",
                        fused=False, refresh=10 ** 9)
    f = out.reshape(-1).numpy()

    # epigraph form: z = (f, s_x, s_y, s_z), |D_a f| <= s_a, sum s_a <= t_a, f >= 0
    def obj(z):
        r = A @ z[:n] - b
        return 0.5 * r @ r

    def grad(z):
        g = np.zeros_like(z)
        g[:n] = A.T @ (A @ z[:n] - b)
        return g

    # one stacked linear inequality J z + c >= 0: s_a - D_a f >= 0,
    # s_a + D_a f >= 0, t_a - sum s_a >= 0
    rows, offs = [], []
    for k, (M, ta) in enumerate(zip(mats, t)):
        E = np.zeros((n, 4 * n))
        E[:, n * (k + 1):n * (k + 2)] = np.eye(n)
        Mz = np.hstack([M, np.zeros((n, 3 * n))])
        rows += [E - Mz, E + Mz, -E.sum(0)[None]]
        offs += [np.zeros(n), np.zeros(n), np.array([ta])]
    J, c = np.vstack(rows), np.concatenate(offs)
    cons = [dict(type="ineq", fun=lambda z: J @ z + c, jac=lambda z: J)]
    z0 = np.concatenate([f0.reshape(-1).numpy()] + [np.abs(M @ f0.reshape(-1).numpy())
                                                    for M in mats])
    res = minimize(obj, z0, jac=grad, constraints=cons, bounds=[(0.0, None)] * (4 * n),
                   method="SLSQP", options=dict(ftol=1e-12, maxiter=5000))
    assert res.success, res.message
    f_ref = res.x[:n]
    assert np.linalg.norm(f - f_ref) <= 1e-4 * np.linalg.norm(f_ref)
    for M, ta in zip(mats, t):
        assert abs(np.abs(M @ f).sum() - ta) <= 1e-4 * ta     # on the boundary


def _asd_potential_grad_fd(f, eps, h=1e-5):
    """Central differences of sum_c [r_c - eps log(1 + r_c/eps)], r = |D f|."""
    g = np.zeros(f.size)
    flat = f.reshape(-1)
    for i in range(f.size):
        e = np.zeros(f.size)
        e[i] = h
        up = otv.asd_potential((flat + e).reshape(f.shape), eps, VS, "reference")
        dn = otv.asd_potential((flat - e).reshape(f.shape), eps, VS, "reference")
        g[i] = (up - dn) / (2 * h)
    return g.reshape(f.shape)


def test_asd_pocs_cascades_are_descent_on_the_stated_potential():
    A, b, proj, back = _system(32, 30)
    beta, alpha, ng, cascades = 1.0 / np.linalg.norm(A, 2) ** 2, 0.3, 3, 2
    f0 = torch.from_numpy(np.random.default_rng(33).uniform(0, 1, size=SHAPE))
    bt = torch.from_numpy(b).reshape(1, 1, -1)
    got = unroll.asd_pocs(f0, bt, proj, back, cascades=cascades, beta=beta, ng=ng, alpha=alpha,
                          mode="reference", fused=False).numpy()
    # the same cascades written from the published description, gradient by FD
    f = f0.numpy().copy()
    step = None
    for i in range(cascades):
        f_new = f - beta * (A.T @ (A @ f.reshape(-1) - b)).reshape(SHAPE)
        if i == 0:
            step = alpha * np.abs(f - f_new).mean()     # alpha * mean|f0 - f| of cascade 0
        f = f_new
        for _ in range(ng):
            phi0 = otv.asd_potential(f, 1.0, VS, "reference")
            f = f - step * _asd_potential_grad_fd(f, 1.0)
            assert otv.asd_potential(f, 1.0, VS, "reference") < phi0   # descent
        beta *= 0.9999
    want = np.maximum(f, 0.0)
    assert np.abs(got - want).max() <= 1e-7 * np.abs(want).max()


def test_fista_l1_converges_to_the_lasso_minimiser():
    from sklearn.linear_model import Lasso
    A, b, proj, back = _system(34, 40)
    m, n = A.shape
    L = np.linalg.norm(A, 2) ** 2
    lam = 20.0
    x = unroll.fista_l1(torch.zeros(SHAPE, dtype=torch.float64),
                        torch.from_numpy(b).reshape(1, 1, -1), proj, back, cascades=3000,
                        lamb=lam, L=1.0 / L, fused=False).reshape(-1).numpy()
    # sklearn: (1 / (2 m)) ||b - A w||^2 + alpha ||w||_1  ==  the FISTA objective / m
    ref = Lasso(alpha=lam / m, fit_intercept=False, tol=1e-14, max_iter=100000).fit(A, b).coef_
    assert (ref == 0).any() and (x == 0).any()          # the threshold is active
    assert np.linalg.norm(x - ref) <= 1e-6 * np.linalg.norm(ref)
