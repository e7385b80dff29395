"""GPU: the two largest workloads, checked through size-independent
properties (the FP64 oracle would take hours at these sizes).

* 512x512x384 at 0.4 mm, 120 views, 768x576 at 0.6 mm (40-ellipsoid
  phantom, Poisson I0 = 5e4): adjoint dot-product identity, linearity, and
  the brick kernels against the ray-driven ones at full size.
* batch 4 of 128^3, 90 views, 192^2 (Poisson I0 = 1e5): one training step of
  the checkpointed PDHG/DTV network -- batch independence of the forward,
  finite gradients, and a loss that an SGD step lowers.
"""
import pytest
import torch

from conftest import rel_l2
from dtvproj import configs, ops, phantom

pytestmark = pytest.mark.gpu


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def test_large_512_dot_linearity_and_kernel_families(cuda):
    g = configs.large_512()
    print(f"[case] large_512 views={g.n_views} vol={g.volume_size} det={g.detector_size}", flush=True)
    x = phantom.phantom(g.volume_size, 0.4, 40, seed=5).to(cuda)
    y = ops.forward(x, g)
    # noisy measurement of the phantom at I0 = 5e4 as the A^T input
    yn = phantom.poisson_noise(y, 5e4, seed=6)
    z = ops.adjoint(yn, g)
    lhs, rhs = _dot(y, yn), _dot(x, z)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # linearity of A on x + 2 u (u: N(0,1))
    u = phantom.gaussian(g.volume_shape(1), 7).to(cuda)
    assert rel_l2(ops.forward(x + 2 * u, g), y + 2 * ops.forward(u, g)) <= 1e-5
    del u
    # brick kernels vs the one-lane-per-ray kernels at full size
    ray = dict(algorithm=1)
    assert rel_l2(ops.forward(x, g, ray), y) <= 1e-6
    assert rel_l2(ops.adjoint(yn, g, ray), z) <= 5e-6


def test_unroll_128_batch4_training_step(cuda):
    from dtvproj import unroll
    g = configs.unroll_128()
    print(f"[case] unroll_128 batch 4 views={g.n_views}", flush=True)
    truth = torch.cat([phantom.phantom(g.volume_size, 1.0, 10, seed=30 + i) for i in range(4)])
    truth = (truth * 0.02).to(cuda)  # attenuation per mm
    with torch.no_grad():
        y = ops.forward(truth, g)
    y = phantom.poisson_noise(y, 1e5, seed=31)
    torch.manual_seed(0)
    net = unroll.PDHGDTVNet(g, iters=3, norm=unroll.power_norm(g, iters=6, device=cuda),
                            checkpointing=True).to(cuda)
    out = net(y)
    # batch independence of the whole network
    with torch.no_grad():
        one = net(y[2:3])
    assert rel_l2(out.detach()[2:3], one) <= 1e-5
    loss = torch.nn.functional.mse_loss(out, truth)
    loss.backward()
    grads = [p.grad for p in net.parameters()]
    assert all(gr is not None and torch.isfinite(gr).all() for gr in grads)
    assert any(float(gr.abs().max()) > 0 for gr in grads)
    with torch.no_grad():
        for p in net.parameters():
            p -= 1e-2 * p.grad / (p.grad.norm() + 1e-30)
        loss2 = torch.nn.functional.mse_loss(net(y), truth)
    assert float(loss2) < float(loss)
