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


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["neumann", "reference"])
def test_fused_fista_tv_step(cuda, mode):
    print(f"[case] fused_fista_tv_step {mode}", flush=True)
    shp = (2, 1, 13, 11, 17)
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
